"""HIP Jacobi-PCG (generic over HIPSpace, fused FP64, mixed refined) against the oracle."""
import os

import numpy as np
import pytest

import oracle
from cie_amd import systems
from cie_amd._lib import Statistics

pytestmark = pytest.mark.gpu


def to_dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda:0")


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def true_relative_residual(csr, x, b):
    r = oracle.csr_product(csr, 1.0, x, -1.0, np.array(b, dtype=np.float64))  # r = b - A x (in place on b copy)
    return np.linalg.norm(r) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def five_point(hip):
    w = systems.five_point_2d(256)
    x, stats, hist, _ = oracle.pcg(w.matrix, w.rhs, relative_tolerance=w.relative_tolerance, history=True)
    return w, hip.CSRMatrix(w.matrix), x, stats, hist


def test_generic_cg_over_hipspace_matches_oracle(hip, torch_gpu, five_point):
    """ConjugateGradients<HIPSpace<double>> with makeDiagonalOperator, as poisson2D/solver.cpp:82-108 builds it."""
    w, matrix, x_ref, s_ref, h_ref = five_point
    space = hip.HIPSpace()
    solver = hip.ConjugateGradients(hip.HIPCSROperator(matrix, space), space, hip.makeDiagonalOperator(matrix, space),
                                    Statistics(5000, 0.0, w.relative_tolerance))
    x = space.makeVector(w.matrix.rows)
    space.fill(x, 0)
    solver.product(0, to_dev(torch_gpu, w.rhs), 1, x)
    stats = solver.getStats()
    assert abs(int(stats.iterationCount) - int(s_ref.iterationCount)) <= 1
    assert rel(x.cpu().numpy(), x_ref) < 1e-10
    n = min(len(h_ref), len(solver.history))
    assert np.allclose(solver.history[:n], h_ref[:n], rtol=1e-6)


def test_generic_cg_over_float_hipspace(hip, torch_gpu, five_point):
    """The reference's FP32 instantiation ConjugateGradients<...<float>> (ConjugateGradients.cpp:10,16) over
    HIPSpace(float32) with the FP32 Jacobi operator, to relative 1e-6 as the driver sets it (solver.cpp:82-108).
    Bounded against the FP64 oracle by its FP64 true residual and the 5-point closed-form condition number."""
    from test_gpu_facade import assert_fp32_solution_bound

    w, matrix, x_ref, s_ref, _ = five_point
    space = hip.HIPSpace(torch_gpu.float32)
    solver = hip.ConjugateGradients(hip.HIPCSROperator(matrix, space), space, hip.makeDiagonalOperator(matrix, space),
                                    Statistics(5000, 0.0, 1e-6))
    x = space.makeVector(w.matrix.rows)
    space.fill(x, 0)
    solver.product(0, torch_gpu.from_numpy(w.rhs.astype(np.float32)).to("cuda:0"), 1, x)
    stats = solver.getStats()
    assert stats.relativeResidual < 1e-6 and 0 < stats.iterationCount < s_ref.iterationCount
    # the recursion's norm history follows the FP64 oracle's to FP32 accuracy early on
    _, _, h64, _ = oracle.pcg(w.matrix, w.rhs, relative_tolerance=1e-6, history=True)
    assert np.allclose(solver.history[:20], h64[:20], rtol=1e-3)
    error, true32 = assert_fp32_solution_bound(w, x.cpu().numpy().astype(np.float64), x_ref)
    print(f"FP32 generic CG: {stats.iterationCount} iterations, true residual {true32:.2e}, error {error:.2e}")


def test_fused_pcg_matches_oracle(hip, torch_gpu, five_point):
    w, matrix, x_ref, s_ref, _ = five_point
    solver = hip.FusedPCG(matrix, Statistics(5000, 0.0, w.relative_tolerance))
    x = torch_gpu.zeros(w.matrix.rows, dtype=torch_gpu.float64, device="cuda:0")
    solver.product(0, to_dev(torch_gpu, w.rhs), 1, x)
    stats = solver.getStats()
    assert solver.info.converged == 1
    assert abs(int(stats.iterationCount) - int(s_ref.iterationCount)) <= 1
    assert stats.relativeResidual < w.relative_tolerance
    assert rel(x.cpu().numpy(), x_ref) < 1e-10
    assert abs(solver.info.trueRelativeResidual - true_relative_residual(w.matrix, x.cpu().numpy(), w.rhs)) < 1e-12


def test_fused_pcg_bitwise_reproducible(hip, torch_gpu, five_point):
    w, matrix, _, _, _ = five_point
    outs = []
    for _ in range(2):
        solver = hip.FusedPCG(matrix, Statistics(5000, 0.0, w.relative_tolerance))
        x = torch_gpu.zeros(w.matrix.rows, dtype=torch_gpu.float64, device="cuda:0")
        solver.product(0, to_dev(torch_gpu, w.rhs), 1, x)
        outs.append((x.cpu().numpy(), solver.getStats().iterationCount, solver.getStats().absoluteResidual))
    assert np.array_equal(outs[0][0], outs[1][0]) and outs[0][1:] == outs[1][1:]


@pytest.mark.parametrize("in_scale,out_scale", [(1.0, 1.0), (0.5, -2.0), (1.0, -1.0)])
def test_fused_pcg_output_scaling_and_warm_start(hip, torch_gpu, in_scale, out_scale):
    w = systems.five_point_2d(40)
    matrix = hip.CSRMatrix(w.matrix)
    x0 = np.random.default_rng(2).standard_normal(w.matrix.rows)
    x_ref, s_ref, _, _ = oracle.pcg(w.matrix, w.rhs, x0=x0, relative_tolerance=1e-9, in_scale=in_scale,
                                    out_scale=out_scale)
    solver = hip.FusedPCG(matrix, Statistics(5000, 0.0, 1e-9))
    x = to_dev(torch_gpu, x0)
    solver.product(in_scale, to_dev(torch_gpu, w.rhs), out_scale, x)
    assert abs(int(solver.getStats().iterationCount) - int(s_ref.iterationCount)) <= 1
    assert rel(x.cpu().numpy(), x_ref) < 1e-9


@pytest.mark.parametrize("in_scale,out_scale", [(1.0, 1.0), (0.5, -2.0)])
@pytest.mark.parametrize("form", ["generic", "fused", "mixed"])
def test_early_exit_leaves_output_untouched(hip, torch_gpu, form, in_scale, out_scale):
    """ConjugateGradients_impl.hpp:100-111 return before the output step of :190-198: an initial guess that
    meets the tolerance leaves `out` bit for bit as it was, whatever the scales. Integer x* on the integer
    5-point stencil makes b = A x* exact in FP64, so the FP64 start residual is exactly zero."""
    w = systems.five_point_2d(24)
    x_star = np.random.default_rng(6).integers(-9, 10, w.matrix.rows).astype(np.float64)
    b_host = oracle.csr_product(w.matrix, 0.0, x_star, 1.0, np.zeros(w.matrix.rows))
    x_oracle, s_oracle, _, _ = oracle.pcg(w.matrix, b_host, x0=x_star, absolute_tolerance=1e-12,
                                          relative_tolerance=1e-8, in_scale=in_scale, out_scale=out_scale)
    assert np.array_equal(x_oracle, x_star) and s_oracle.iterationCount == 0
    matrix = hip.CSRMatrix(w.matrix)
    settings = Statistics(100, 1e-12, 1e-8)
    if form == "generic":
        space = hip.HIPSpace()
        solver = hip.ConjugateGradients(hip.HIPCSROperator(matrix, space), space,
                                        hip.makeDiagonalOperator(matrix, space), settings)
    elif form == "fused":
        solver = hip.FusedPCG(matrix, settings)
    else:
        solver = hip.MixedPCG(matrix, settings)
    x = to_dev(torch_gpu, x_star)
    solver.product(in_scale, to_dev(torch_gpu, b_host), out_scale, x)
    assert np.array_equal(x.cpu().numpy(), x_star)
    stats = solver.getStats()
    assert stats.iterationCount == 0 and stats.absoluteResidual == 0.0 and stats.relativeResidual == 1.0
    if form != "generic":
        assert solver.info.converged == 1 and solver.info.refinements == 0


@pytest.mark.parametrize("kernel", ["autotuned", "tiled", "blocks"])
@pytest.mark.parametrize("dim,order,n", [(3, 2, 10), (2, 4, 12), (3, 4, 4)])
def test_fused_pcg_lagrange_lognormal(hip, torch_gpu, dim, order, n, kernel):
    count = n ** dim
    k = np.exp(systems.normal(2, 1.0, count))
    csr = systems.lagrange(dim, order, n, conductivity=k)
    b = systems.uniform(1, csr.rows)
    x_ref, s_ref, _, _ = oracle.pcg(csr, b, relative_tolerance=1e-10)
    matrix = hip.CSRMatrix(csr)
    if kernel == "tiled":
        matrix.set_launch(64, "tiled", 256, 8)
    elif kernel == "blocks":
        matrix.set_launch(64, "blocks", 3, 8)
    else:
        matrix.autotune(64, 3, kernels=("staged", "vector", "tiled", "blocks"))
    solver = hip.FusedPCG(matrix, Statistics(20000, 0.0, 1e-10))
    x = torch_gpu.zeros(csr.rows, dtype=torch_gpu.float64, device="cuda:0")
    solver.product(0, to_dev(torch_gpu, b), 1, x)
    assert abs(int(solver.getStats().iterationCount) - int(s_ref.iterationCount)) <= 2
    assert rel(x.cpu().numpy(), x_ref) < 1e-8
    assert solver.info.trueRelativeResidual < 2e-10


@pytest.mark.parametrize("restart,tiled", [(False, False), (True, False), (False, True), (False, "blocks")])
def test_mixed_pcg_reaches_fp64_tolerance(hip, torch_gpu, restart, tiled):
    count = 8 ** 3
    k = np.exp(systems.normal(2, 1.0, count))
    csr = systems.lagrange(3, 4, 8, conductivity=k)
    b = systems.uniform(1, csr.rows)
    x_ref, _, _, _ = oracle.pcg(csr, b, relative_tolerance=1e-14, max_iterations=20000)
    matrix = hip.CSRMatrix(csr)
    if tiled == "blocks":
        matrix.set_launch(64, "blocks", 3, 16)
        matrix.set_launch(32, "blocks", 3, 8)
    elif tiled:
        matrix.set_launch(64, "tiled", 256, 16)
        matrix.set_launch(32, "tiled", 256, 8)
    solver = hip.MixedPCG(matrix, Statistics(20000, 0.0, 1e-10), restart=restart)
    x = torch_gpu.zeros(csr.rows, dtype=torch_gpu.float64, device="cuda:0")
    solver.product(0, to_dev(torch_gpu, b), 1, x)
    info = solver.info
    assert info.converged == 1 and info.refinements >= 1
    got = x.cpu().numpy()
    truth = true_relative_residual(csr, got, b)
    assert truth < 1e-10
    assert abs(info.trueRelativeResidual - truth) < 1e-12
    assert solver.getStats().relativeResidual < 1e-10
    assert rel(got, x_ref) < 1e-6


@pytest.mark.parametrize("restart", [False, True])
@pytest.mark.parametrize("rows", [1, 2, 3, 300])
def test_mixed_pcg_when_the_fp32_recurrence_solves_exactly(hip, torch_gpu, rows, restart):
    """Diagonal systems with power-of-two entries: the first Jacobi-preconditioned FP32 step is exact (alpha = 1,
    x = D^-1 b), so r and then p become zero and the next p.q is 0. The mixed solver folds without a step, and
    the FP64 true residual (zero) ends the solve as converged, not as a breakdown; x equals the oracle's."""
    diagonal = 2.0 ** (np.arange(rows) % 5 + 1)
    csr = systems.CSR(rows, rows, np.arange(rows + 1, dtype=np.int32), np.arange(rows, dtype=np.int32), diagonal)
    b = (np.arange(rows) % 7 + 1) * 0.5
    x_ref, s_ref, _, _ = oracle.pcg(csr, b, relative_tolerance=1e-10)
    matrix = hip.CSRMatrix(csr)
    solver = hip.MixedPCG(matrix, Statistics(100, 0.0, 1e-10), restart=restart)
    x = torch_gpu.zeros(rows, dtype=torch_gpu.float64, device="cuda:0")
    solver.product(0, to_dev(torch_gpu, b), 1, x)
    assert solver.info.converged == 1
    assert np.array_equal(x.cpu().numpy(), x_ref) and np.array_equal(x_ref, b / diagonal)
    assert int(solver.getStats().iterationCount) <= int(s_ref.iterationCount) + 1


@pytest.mark.parametrize("seed", range(6))
def test_mixed_pcg_restarts_below_the_fp32_floor(hip, torch_gpu, seed):
    """Single rows and tiny systems with lognormal entries: one FP32 step takes the recursive residual far
    below the FP32 floor of the true one (the correction is accumulated in FP32), so the first refinement's
    replaced residual exceeds the recursive one many times over and the direction restarts, instead of
    combining with the stale one; every solve converges to the oracle's solution."""
    rng = np.random.default_rng(seed)
    for elements in (1, 2):
        csr = systems.lagrange(2, 2, elements, conductivity=np.exp(rng.normal(0.0, 1.0, elements * elements)))
        b = rng.uniform(-1.0, 1.0, csr.rows)
        x_ref, _, _, _ = oracle.pcg(csr, b, relative_tolerance=1e-14, max_iterations=1000)
        solver = hip.MixedPCG(hip.CSRMatrix(csr), Statistics(1000, 0.0, 1e-10))
        x = torch_gpu.zeros(csr.rows, dtype=torch_gpu.float64, device="cuda:0")
        solver.product(0, to_dev(torch_gpu, b), 1, x)
        assert solver.info.converged == 1, (seed, elements)
        got = x.cpu().numpy()
        assert true_relative_residual(csr, got, b) < 1e-10
        assert rel(got, x_ref) < 1e-8


@pytest.mark.parametrize("restart", [False, True])
def test_mixed_pcg_iteration_cap_and_cadence(hip, torch_gpu, restart):
    """The mixed form stops at the cap with x holding every correction (the last trip folds) and reports the
    FP64 true residual of that x; refinements recompute r = b - A x on the trip after the fold."""
    w = systems.five_point_2d(48)
    matrix = hip.CSRMatrix(w.matrix)
    b = to_dev(torch_gpu, w.rhs)
    solver = hip.MixedPCG(matrix, Statistics(25, 0.0, 1e-30), restart=restart)
    x = torch_gpu.zeros_like(b)
    solver.product(0, b, 1, x)
    info = solver.info
    assert solver.getStats().iterationCount == 25 and info.converged == 0
    truth = true_relative_residual(w.matrix, x.cpu().numpy(), w.rhs)
    assert solver.getStats().relativeResidual == pytest.approx(truth, rel=1e-9)
    assert info.trueRelativeResidual == pytest.approx(truth, rel=1e-9)
    # the same 25 trips uncapped reach the same x up to the rounding of the fold
    longer = hip.MixedPCG(matrix, Statistics(20000, 0.0, 1e-10), restart=restart)
    x2 = torch_gpu.zeros_like(b)
    longer.product(0, b, 1, x2)
    assert longer.info.converged == 1 and longer.info.trueRelativeResidual < 1e-10
    assert longer.info.refinements >= 2


def test_fused_pcg_errors(hip, torch_gpu):
    # r^T z <= 0 before the loop (ConjugateGradients_impl.hpp:119-123)
    neg = systems.CSR(4, 4, np.arange(5, dtype=np.int32), np.arange(4, dtype=np.int32), -np.ones(4))
    solver = hip.FusedPCG(hip.CSRMatrix(neg), Statistics(100, 0.0, 1e-10))
    x = torch_gpu.zeros(4, dtype=torch_gpu.float64, device="cuda:0")
    with pytest.raises(hip._lib.Breakdown):
        solver.product(0, to_dev(torch_gpu, np.ones(4)), 1, x)
    # p^T A p <= 0 (:145-149): symmetric indefinite with a positive diagonal
    a = np.array([[1.0, 2.0], [2.0, 1.0]])
    indefinite = systems.CSR(2, 2, np.array([0, 2, 4], np.int32), np.array([0, 1, 0, 1], np.int32), a.ravel())
    solver = hip.FusedPCG(hip.CSRMatrix(indefinite), Statistics(100, 0.0, 1e-10))
    with pytest.raises(hip._lib.NotPositiveDefinite):
        solver.product(0, to_dev(torch_gpu, np.array([1.0, -1.0])), 1, torch_gpu.zeros(2, dtype=torch_gpu.float64,
                                                                                        device="cuda:0"))
    # iteration cap: stats, no exception (:134, :200)
    w = systems.five_point_2d(32)
    solver = hip.FusedPCG(hip.CSRMatrix(w.matrix), Statistics(7, 0.0, 1e-30))
    x = torch_gpu.zeros(w.matrix.rows, dtype=torch_gpu.float64, device="cuda:0")
    solver.product(0, to_dev(torch_gpu, w.rhs), 1, x)
    assert solver.getStats().iterationCount == 7 and solver.info.converged == 0
    _, capped, _, _ = oracle.pcg(w.matrix, w.rhs, relative_tolerance=1e-30, max_iterations=7)
    assert solver.getStats().absoluteResidual == pytest.approx(capped.absoluteResidual, rel=1e-10)
    # zero iterations: x is only rescaled
    solver = hip.FusedPCG(hip.CSRMatrix(w.matrix), Statistics(0, 0.0, 1e-8))
    solver.product(0, to_dev(torch_gpu, w.rhs), 1, x)
    assert solver.getStats().iterationCount == 0
    # exact initial guess: early exit before the loop
    solver = hip.FusedPCG(hip.CSRMatrix(w.matrix), Statistics(100, 1e-12, 1e-8))
    solver.product(0, torch_gpu.zeros_like(x), 1, torch_gpu.zeros_like(x))
    assert solver.getStats().iterationCount == 0 and solver.info.converged == 1
    with pytest.raises(hip._lib.ShapeError):
        hip.FusedPCG(hip.CSRMatrix(w.matrix), Statistics(10, 0.0, 1e-8)).product(0, x[:-1], 1, x[:-1])


def host_threads():
    return max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))))


@pytest.fixture(scope="module")
def q2_64():
    """3D Q2 64^3 (2,048,383 DOFs) and the oracle's threaded FP64 Jacobi-PCG solution to 1e-10."""
    w = systems.q2_laplace_3d(64)
    assert (w.matrix.rows, w.matrix.nnz) == (2048383, 127263527)
    x_ref, s_ref, _, _ = oracle.pcg(w.matrix, w.rhs, relative_tolerance=1e-10, max_iterations=20000,
                                    threads=host_threads())
    return w, x_ref, s_ref


# Mixed refinement stops on the FP64 true residual (1e-10), not on the FP64 recurrence, so its solution
# differs from the oracle's by the errors both leave at a 1e-10 residual. Against a 1e-14 solve the FP64
# solution is 1.2e-9 away and the mixed one 3.6e-9 on this system, so the stated refinement tolerance of the
# mixed solution against the oracle's 1e-10 solution is 1e-8 relative.
MIXED_SOLUTION_TOLERANCE = 1e-8


@pytest.mark.parametrize("launch", ["sorted", "sorted-4/8", "paired-2/4", "paired-4/4", "autotuned"])
def test_full_size_q2_64_fp64_and_mixed(hip, torch_gpu, q2_64, launch):
    """3D Q2 64^3: the FP64 solution equals the oracle's (threaded CPU Jacobi-PCG) to 1e-10 relative with the
    iteration count within one; the mixed solution reaches FP64 true relative residual 1e-10 and agrees with
    the oracle within the refinement tolerance. "sorted" pins sorted 1 x 3 tiles of 2048 blocks for both
    precisions, "sorted-4/8" FP64 tiles of 1024 blocks and FP32 of 2048, "paired-U/V" paired 1 x 5 / 1 x 3
    tiles of 256 U (FP64) and 256 V (FP32) blocks per width, the launches the benchmark chooses among."""
    w, x_ref, s_ref = q2_64
    matrix = hip.CSRMatrix(w.matrix)
    if launch.startswith("sorted"):
        matrix.set_launch(64, "sorted", 3, 4 if launch == "sorted-4/8" else 8)
        matrix.set_launch(32, "sorted", 3, 8)
    elif launch.startswith("paired"):
        unroll64, unroll32 = map(int, launch.split("-")[1].split("/"))
        matrix.set_launch(64, "paired", 53, unroll64)
        matrix.set_launch(32, "paired", 53, unroll32)
    else:
        matrix.autotune(64, 3)
        matrix.autotune(32, 3)
    b = to_dev(torch_gpu, w.rhs)
    fp64 = hip.FusedPCG(matrix, Statistics(20000, 0.0, 1e-10))
    x64 = torch_gpu.zeros_like(b)
    fp64.product(0, b, 1, x64)
    assert fp64.info.converged == 1 and fp64.info.trueRelativeResidual < 2e-10
    assert abs(int(fp64.getStats().iterationCount) - int(s_ref.iterationCount)) <= 1
    error64 = rel(x64.cpu().numpy(), x_ref)
    assert error64 < 1e-10
    mixed = hip.MixedPCG(matrix, Statistics(20000, 0.0, 1e-10))
    xm = torch_gpu.zeros_like(b)
    mixed.product(0, b, 1, xm)
    assert mixed.info.converged == 1 and mixed.info.trueRelativeResidual < 1e-10
    host = xm.cpu().numpy()
    assert true_relative_residual(w.matrix, host, w.rhs) < 1e-10
    error_mixed = rel(host, x_ref)
    assert error_mixed < MIXED_SOLUTION_TOLERANCE
    # with the FP32 plane's row-sum correction the inner solve needs the FP64 iteration count
    assert int(mixed.getStats().iterationCount) <= int(fp64.getStats().iterationCount) + 6
    print(f"{launch} {matrix.launch(64)} {matrix.launch(32)}: fp64 {fp64.getStats().iterationCount} "
          f"iterations (oracle {s_ref.iterationCount}) error vs oracle {error64:.2e}; mixed "
          f"{mixed.getStats().iterationCount} iterations, {mixed.info.refinements} refinements, error vs oracle "
          f"{error_mixed:.2e}")


def test_full_size_q4_32_lognormal_mixed(hip, torch_gpu):
    """3D Q4 32^3 with k_e = exp(N(0,1)) (2,048,383 DOFs, 423,564,751 nonzeros, up to 729 per row): the
    mixed solve reaches FP64 true relative residual 1e-10 and is as close to a 1e-14 FP64 solve as the
    FP64 solve to 1e-10 is."""
    w = systems.q4_lognormal_3d(32)
    assert (w.matrix.rows, w.matrix.nnz) == (2048383, 423564751)
    matrix = hip.CSRMatrix(w.matrix)
    b = to_dev(torch_gpu, w.rhs)
    fp64 = hip.FusedPCG(matrix, Statistics(20000, 0.0, 1e-10))
    x64 = torch_gpu.zeros_like(b)
    fp64.product(0, b, 1, x64)
    assert fp64.info.converged == 1 and fp64.info.trueRelativeResidual < 2e-10
    mixed = hip.MixedPCG(matrix, Statistics(20000, 0.0, 1e-10))
    xm = torch_gpu.zeros_like(b)
    mixed.product(0, b, 1, xm)
    assert mixed.info.converged == 1 and mixed.info.trueRelativeResidual < 1e-10
    tight = hip.FusedPCG(matrix, Statistics(40000, 0.0, 1e-14))
    xt = torch_gpu.zeros_like(b)
    tight.product(0, b, 1, xt)
    assert tight.info.converged == 1
    host, reference = xm.cpu().numpy(), xt.cpu().numpy()
    assert true_relative_residual(w.matrix, host, w.rhs) < 1e-10
    error_mixed, error_fp64 = rel(host, reference), rel(x64.cpu().numpy(), reference)
    assert error_mixed < 10 * error_fp64 + 1e-12
    print(f"fp64 {fp64.getStats().iterationCount} iterations {fp64.info.seconds:.3f}s error {error_fp64:.2e}; mixed "
          f"{mixed.getStats().iterationCount} iterations, {mixed.info.refinements} refinements "
          f"{mixed.info.seconds:.3f}s error {error_mixed:.2e}; speedup {fp64.info.seconds / mixed.info.seconds:.2f}")

"""Masked operators, Jacobi smoothers and the p-multigrid through the C ABI
against the oracle (MaskedCSROperator.cpp, MaskedIdentityOperator.cpp,
JacobiOperator.cpp, MaskedJacobiOperator.cpp, ApproximateSubstitution.cpp,
poisson2D/solver.cpp:111-232)."""
import time

import numpy as np
import pytest
import scipy.sparse

import oracle
from cie_amd import systems
from cie_amd._lib import MissingDiagonal, NotImplementedForArguments, ShapeError, Statistics

pytestmark = pytest.mark.gpu

SCALES = [(0.0, 1.0), (0.0, -1.0), (0.0, 2.5), (1.0, 1.0), (1.0, -1.0), (1.0, 0.5), (-1.0, 1.0), (-1.0, -1.0),
          (-1.0, 3.0), (2.0, 1.0), (2.0, -1.0), (0.5, -2.0)]


def _dev(torch, array):
    return torch.from_numpy(np.ascontiguousarray(array)).cuda()


def integer_csr(seed, rows=3000, density=0.004):
    rng = np.random.default_rng(seed)
    a = scipy.sparse.random(rows, rows, density=density, random_state=seed, format="csr")
    a = (a + scipy.sparse.eye(rows)).tocsr()
    a.sort_indices()
    a.data = rng.integers(-9, 10, a.nnz).astype(np.float64)
    return systems.CSR(rows, rows, a.indptr.astype(np.int32), a.indices.astype(np.int32), a.data)


def test_masked_products_and_identity_are_exact_on_integer_data(hip, torch_gpu):
    csr = integer_csr(1)
    rng = np.random.default_rng(2)
    mask_values = rng.integers(0, 4, csr.rows).astype(np.uint16)
    mask = hip.Mask(mask_values)
    matrix = hip.CSRMatrix(csr)
    space = hip.HIPSpace()
    x = rng.integers(-5, 6, csr.rows).astype(np.float64)
    y0 = rng.integers(-5, 6, csr.rows).astype(np.float64)
    for threshold in (0, 1, 3, 4):
        op = hip.MaskedCSROperator(matrix, mask, threshold, space)
        identity = hip.MaskedIdentityOperator(space, mask, threshold)
        for in_scale, out_scale in SCALES:
            y = _dev(torch_gpu, y0)
            op.product(in_scale, _dev(torch_gpu, x), out_scale, y)
            expected = oracle.masked_csr_product(csr, mask_values, threshold, in_scale, x, out_scale, y0.copy())
            assert np.array_equal(y.cpu().numpy(), expected), (threshold, in_scale, out_scale)
            y = _dev(torch_gpu, y0)
            identity.product(in_scale, _dev(torch_gpu, x), out_scale, y)
            expected = oracle.masked_identity(mask_values, threshold, in_scale, x, out_scale, y0.copy())
            assert np.array_equal(y.cpu().numpy(), expected)
    with pytest.raises(ShapeError):
        hip.maskedMatrix(matrix, hip.Mask(mask_values[:-1]), 2)


def _legendre(order, n):
    csr = systems.lagrange(2, order, n, basis="legendre")
    return csr, systems.ansatz_mask(2, order, n)


@pytest.mark.parametrize("threshold", [None, 2, 3])
def test_jacobi_operator_matches_oracle(hip, torch_gpu, threshold):
    csr, mask_values = _legendre(3, 12)
    space = hip.HIPSpace()
    matrix = hip.CSRMatrix(csr)
    lhs = matrix if threshold is None else hip.maskedMatrix(matrix, hip.Mask(mask_values), threshold)
    jacobi = hip.JacobiOperator(space, csr.rows, lhs, hip.makeDiagonalOperator(matrix, space), 6, 2.0 / 3.0)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(csr.rows)
    if threshold is not None:
        b = np.where(mask_values < threshold, b, 0.0)
    y0 = rng.standard_normal(csr.rows)
    for in_scale, out_scale in ((0.0, 1.0), (1.0, -1.0), (2.0, 0.5)):
        y = _dev(torch_gpu, y0)
        jacobi.product(in_scale, _dev(torch_gpu, b), out_scale, y)
        expected = oracle.jacobi(csr, 6, 2.0 / 3.0, in_scale, b, out_scale, y0.copy(),
                                 mask=None if threshold is None else mask_values, threshold=threshold or 0)
        assert np.allclose(y.cpu().numpy(), expected, rtol=1e-12, atol=1e-12 * np.abs(expected).max())


def test_masked_jacobi_and_approximate_substitution_match_oracle(hip, torch_gpu):
    csr, mask_values = _legendre(2, 10)
    space = hip.HIPSpace()
    matrix = hip.CSRMatrix(csr)
    mask = hip.Mask(mask_values)
    rng = np.random.default_rng(4)
    b, y0 = rng.standard_normal(csr.rows), rng.standard_normal(csr.rows)
    masked = hip.MaskedJacobiOperator(matrix, 3, 0.7, mask, 2, space)
    y = _dev(torch_gpu, y0)
    masked.product(0.0, _dev(torch_gpu, b), 1.0, y)
    expected = oracle.masked_jacobi(csr, mask_values, 2, 3, 0.7, 0.0, b, 1.0, y0.copy())
    assert np.allclose(y.cpu().numpy(), expected, rtol=1e-12, atol=1e-12 * np.abs(expected).max())
    single = hip.MaskedJacobiOperator(matrix, 1, 0.7, mask, 3, space)
    y = _dev(torch_gpu, y0)
    single.product(2.0, _dev(torch_gpu, b), -1.0, y)
    expected = oracle.masked_jacobi(csr, mask_values, 3, 1, 0.7, 2.0, b, -1.0, y0.copy())
    assert np.allclose(y.cpu().numpy(), expected, rtol=1e-12, atol=1e-12 * np.abs(expected).max())
    with pytest.raises(NotImplementedForArguments):
        masked.product(1.0, _dev(torch_gpu, b), 1.0, _dev(torch_gpu, y0))
    substitution = hip.ApproximateSubstitution(matrix, 2, space)
    y = _dev(torch_gpu, y0)
    substitution.product(0.0, _dev(torch_gpu, b), 1.0, y)
    expected = oracle.approximate_substitution(csr, 2, 0.0, b, 1.0, y0.copy())
    assert np.allclose(y.cpu().numpy(), expected, rtol=1e-12, atol=1e-12 * np.abs(expected).max())
    missing = systems.CSR(2, 2, np.array([0, 1, 2], dtype=np.int32), np.array([1, 0], dtype=np.int32),
                          np.array([1.0, 1.0]))
    with pytest.raises(MissingDiagonal):
        hip.ApproximateSubstitution(hip.CSRMatrix(missing), 1, space).product(
            0.0, _dev(torch_gpu, np.ones(2)), 1.0, _dev(torch_gpu, np.zeros(2)))


def test_nested_product_operator(hip, torch_gpu):
    csr, mask_values = _legendre(2, 6)
    space = hip.HIPSpace()
    matrix = hip.CSRMatrix(csr)
    mask = hip.Mask(mask_values)
    nested = hip.NestedProductOperator(space, hip.HIPCSROperator(matrix, space),
                                       hip.MaskedIdentityOperator(space, mask, 2), csr.rows)
    x = systems.uniform(5, csr.rows)
    y = space.makeVector(csr.rows)
    nested.product(0, _dev(torch_gpu, x), 1, y)
    expected = csr.to_scipy() @ np.where(mask_values < 2, x, 0.0)
    assert np.allclose(y.cpu().numpy(), expected, rtol=1e-13, atol=1e-13 * np.abs(expected).max())


@pytest.mark.parametrize("order,n,blocks", [(1, 20, False), (2, 16, False), (3, 12, False), (4, 10, False),
                                            (4, 32, False), (3, 12, True), (4, 32, True), (3, 12, "sorted"),
                                            (4, 32, "sorted")])
def test_pmg_matches_oracle(hip, torch_gpu, order, n, blocks):
    """blocks: the finest grid's sweeps and updates walk 1 x 3 column blocks, plain or column-sorted."""
    csr, mask_values = _legendre(order, n)
    b = systems.uniform(1, csr.rows)
    x_ref, history, coarse = oracle.pmg(csr, mask_values, order, b)
    matrix = hip.CSRMatrix(csr)
    if blocks:
        matrix.set_launch(64, "sorted" if blocks == "sorted" else "blocks", 3, 8)
    pmg = hip.PMultigrid(matrix, hip.Mask(mask_values), order)
    x = torch_gpu.zeros(csr.rows, dtype=torch_gpu.float64, device="cuda:0")
    info = pmg.solve(_dev(torch_gpu, b), x)
    assert info.converged == 1
    assert info.cycles == len(history)
    assert info.coarseIterations == int(coarse.sum())
    assert abs(info.residual - history[-1]) <= 1e-6 * history[-1]
    got = x.cpu().numpy()
    assert np.linalg.norm(got - x_ref) <= 1e-9 * np.linalg.norm(x_ref)


@pytest.mark.parametrize("launch", [("blocks", 2, 4), ("sorted", 3, 8), ("tiled", 256, 16)])
def test_pmg_grid_launches(hip, torch_gpu, launch):
    """Every grid's matrix is reachable (the finest grid, where the mask keeps every row, is the plan's own
    matrix; the grids below it are compacted to their kept rows and columns, grid 0 and -1 being the
    renumbered coarse matrix), and a launch set on the masked grids and the coarse matrix changes only
    summation orders: the oracle's cycle and coarse CG counts, x to 1e-9."""
    order, n = 3, 6
    csr, mask_values = _legendre(order, n)
    b = systems.uniform(1, csr.rows)
    x_ref, history, coarse = oracle.pmg(csr, mask_values, order, b)
    matrix = hip.CSRMatrix(csr)
    pmg = hip.PMultigrid(matrix, hip.Mask(mask_values), order)
    grids = [pmg.grid(g) for g in range(order)]
    assert grids[order - 1].handle.value == matrix.handle.value  # mask < order + 1 keeps every row
    for g in range(0, order - 1):
        kept = int(np.count_nonzero(mask_values < g + 2))
        assert grids[g].rows == kept and grids[g].columns == kept
    assert [g.nnz for g in grids] == sorted(g.nnz for g in grids) and grids[0].nnz < csr.nnz
    coarse_matrix = pmg.grid(-1)
    assert coarse_matrix.rows == int(np.count_nonzero(mask_values < 2))
    assert coarse_matrix.handle.value == grids[0].handle.value
    for g in grids[:order - 1]:
        g.set_launch(64, *launch)
        assert tuple(g.launch(64)) == launch
    x = torch_gpu.zeros(csr.rows, dtype=torch_gpu.float64, device="cuda:0")
    info = pmg.solve(_dev(torch_gpu, b), x)
    assert info.converged == 1 and info.cycles == len(history) and info.coarseIterations == int(coarse.sum())
    assert np.linalg.norm(x.cpu().numpy() - x_ref) <= 1e-9 * np.linalg.norm(x_ref)
    with pytest.raises(hip._lib.InvalidArgument):
        pmg.grid(order)
    del grids, coarse_matrix  # views: dropping them leaves the plan's matrices alone
    x.zero_()
    again = pmg.solve(_dev(torch_gpu, b), x)
    assert again.cycles == info.cycles


def test_pmg_rejects_bad_input(hip, torch_gpu):
    csr, mask_values = _legendre(2, 4)
    with pytest.raises(ShapeError):
        hip.PMultigrid(hip.CSRMatrix(csr), hip.Mask(mask_values[:-1]), 2)


def test_full_size_pmg_q4_legendre_2d(hip, torch_gpu):
    """2D Q4 integrated Legendre, 512^2 elements (4,190,209 rows): the p-multigrid to 1e-6 of ||b||, timed,
    beside the fused Jacobi-PCG on the same system (the driver's --solver multigrid vs cg)."""
    order, n = 4, 512
    csr = systems.lagrange(2, order, n, basis="legendre", threads=16)
    mask_values = systems.ansatz_mask(2, order, n)
    b = systems.uniform(1, csr.rows)
    matrix = hip.CSRMatrix(csr)
    mask = hip.Mask(mask_values)
    t0 = time.perf_counter()
    pmg = hip.PMultigrid(matrix, mask, order)
    setup = time.perf_counter() - t0
    b_dev = _dev(torch_gpu, b)
    x = torch_gpu.zeros(csr.rows, dtype=torch_gpu.float64, device="cuda:0")
    pmg.solve(b_dev, x)  # warm-up (tile cuts, narrow planes)
    x.zero_()
    info = pmg.solve(b_dev, x)
    assert info.converged == 1 and info.residual <= 1e-6 * info.initialResidual
    xp = torch_gpu.zeros_like(x)
    solver = hip.FusedPCG(matrix, Statistics(100000, 0.0, 1e-6))
    solver.product(0, b_dev, 1, xp)
    print(f"\n2D Q4 Legendre {n}^2: {csr.rows} rows, {csr.nnz} nonzeros; p-multigrid: {info.cycles} cycles, "
          f"{info.coarseIterations} coarse CG steps, {info.seconds * 1e3:.1f} ms (set-up {setup * 1e3:.0f} ms); "
          f"Jacobi-PCG to 1e-6: {solver.info.stats.iterationCount} iterations, {solver.info.seconds * 1e3:.1f} ms")

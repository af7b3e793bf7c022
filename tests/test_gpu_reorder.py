"""Reordering through the C ABI (reorder.cpp): the host Cuthill-McKee map
against the oracle, the device CSR and vector permutations against the
oracle bitwise, and the solve's invariance under the permutation."""
import time

import numpy as np
import pytest
import scipy.sparse

import oracle
from cie_amd import systems
from cie_amd._lib import ShapeError, Statistics

pytestmark = pytest.mark.gpu


def _dev(torch, array):
    return torch.from_numpy(np.ascontiguousarray(array)).cuda()


def _shuffled(csr, seed):
    shuffle = np.random.default_rng(seed).permutation(csr.rows).astype(np.int32)
    return systems.CSR(csr.rows, csr.rows, *oracle.reorder_csr(csr, shuffle))


def _patterns():
    yield "q2-3d", systems.lagrange(3, 2, 6)
    yield "q2-3d-shuffled", _shuffled(systems.lagrange(3, 2, 6), 1)
    yield "q4-2d-shuffled", _shuffled(systems.lagrange(2, 4, 6), 2)
    # several components of uneven size, and isolated rows
    blocks = [systems.lagrange(2, 1, k) for k in (3, 2, 5, 2, 4)]  # 2 elements: one isolated row
    yield "components", systems.block_diagonal(blocks)[0]
    rng = np.random.default_rng(3)
    a = scipy.sparse.random(700, 700, density=0.01, random_state=5, format="csr")
    a = (a + a.T + scipy.sparse.eye(700)).tocsr()
    a.sort_indices()
    yield "random", systems.CSR(700, 700, a.indptr.astype(np.int32), a.indices.astype(np.int32),
                                rng.standard_normal(a.nnz))


@pytest.mark.parametrize("strategy", ["cuthill-mckee", "reverse-cuthill-mckee", "none"])
def test_make_reordering_matches_oracle(hip, torch_gpu, strategy):
    for name, csr in _patterns():
        assert np.array_equal(hip.makeReordering(csr, strategy), oracle.make_reordering(csr, strategy)), name


def test_make_reordering_rejects_bad_patterns(hip, torch_gpu):
    csr = systems.lagrange(2, 1, 3)
    bad = systems.CSR(csr.rows, csr.rows, csr.row_extents, csr.column_indices.copy(), csr.entries)
    bad.column_indices[3] = csr.rows
    with pytest.raises(ShapeError):
        hip.makeReordering(bad, "cuthill-mckee")
    with pytest.raises(ValueError):
        hip.makeReordering(csr, "sloan")


def _long_row_matrix(rng):
    rows = 1500
    lengths = rng.integers(0, 6, rows)
    lengths[[3, 800]] = [1400, 1025]  # ranked out of L2 (beyond the LDS row buffer)
    lengths[[10, 11]] = [1024, 0]
    ext = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int32)
    cols = np.concatenate([np.sort(rng.choice(rows, n, replace=False)) for n in lengths]).astype(np.int32)
    return systems.CSR(rows, rows, ext, cols, rng.standard_normal(int(ext[-1])))


@pytest.mark.parametrize("reverse", [False, True])
def test_csr_reorder_matches_oracle_bitwise(hip, torch_gpu, reverse):
    rng = np.random.default_rng(7)
    cases = [csr for _, csr in _patterns()] + [_long_row_matrix(rng)]
    for csr in cases:
        mapping = rng.permutation(csr.rows).astype(np.int32)
        matrix = hip.CSRMatrix(csr)
        reordered = matrix.reorder(mapping, reverse=reverse)
        got = reordered.copy_out()
        extents, columns, entries = oracle.reorder_csr(csr, mapping, reverse=reverse)
        assert np.array_equal(got.row_extents, extents)
        assert np.array_equal(got.column_indices, columns)
        assert np.array_equal(got.entries, entries)
        # the FP32 plane follows: A32' x' equals the permuted A32 x up to FP32 rounding of the sums
        x = rng.standard_normal(csr.rows)
        space32 = hip.HIPSpace(torch_gpu.float32)
        y = space32.makeVector(csr.rows)
        hip.HIPCSROperator(reordered, space32).product(0, _dev(torch_gpu, x.astype(np.float32)), 1, y)
        ref = oracle.csr_product(systems.CSR(csr.rows, csr.rows, extents, columns, entries), 0.0, x, 1.0,
                                 np.zeros(csr.rows))
        assert np.linalg.norm(y.cpu().numpy() - ref) <= 1e-5 * max(1.0, np.linalg.norm(ref))
    with pytest.raises(ShapeError):
        hip.CSRMatrix(cases[0]).reorder(np.zeros(cases[0].rows, dtype=np.int32))


@pytest.mark.parametrize("reverse", [False, True])
def test_csr_reorder_prepares_maps_of_several_scan_blocks(hip, torch_gpu, reverse):
    """The new row extents are scanned on the device in runs of 4096 rows: a matrix of 3 * 4096 + 17 rows
    with rows of 0 to 40 entries, against the oracle bitwise; maps that are not permutations (an index out of
    range at either end, one index twice) are refused in both directions."""
    rng = np.random.default_rng(21)
    rows = 3 * 4096 + 17
    lengths = rng.integers(0, 41, rows)
    ext = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int32)
    cols = np.concatenate([np.sort(rng.choice(rows, n, replace=False)) for n in lengths]).astype(np.int32)
    csr = systems.CSR(rows, rows, ext, cols, rng.standard_normal(int(ext[-1])))
    mapping = rng.permutation(rows).astype(np.int32)
    matrix = hip.CSRMatrix(csr)
    got = matrix.reorder(mapping, reverse=reverse).copy_out()
    extents, columns, entries = oracle.reorder_csr(csr, mapping, reverse=reverse)
    assert np.array_equal(got.row_extents, extents)
    assert np.array_equal(got.column_indices, columns)
    assert np.array_equal(got.entries, entries)
    for position, value in ((5, -1), (rows - 1, rows), (4100, int(mapping[9000]))):
        bad = mapping.copy()
        bad[position] = value
        with pytest.raises(ShapeError):
            matrix.reorder(bad, reverse=reverse)
    again = matrix.reorder(mapping, reverse=reverse).copy_out()  # a refused map leaves nothing behind
    assert np.array_equal(again.column_indices, columns)


def test_vector_reorder_gather_scatter_and_errors(hip, torch_gpu):
    rng = np.random.default_rng(8)
    n = 100003
    mapping = rng.permutation(n).astype(np.int32)
    values = rng.standard_normal(n)
    m = _dev(torch_gpu, mapping)
    gathered = hip.reorder(m, _dev(torch_gpu, values))
    assert np.array_equal(gathered.cpu().numpy(), oracle.reorder_vector(mapping, values))
    back = hip.reorder(m, gathered, reverse=True)
    assert np.array_equal(back.cpu().numpy(), values)
    g32 = hip.reorder(m, _dev(torch_gpu, values.astype(np.float32)))
    assert np.array_equal(g32.cpu().numpy(), values.astype(np.float32)[mapping])
    bad = mapping.copy()
    bad[17] = n
    with pytest.raises(ShapeError):
        hip.reorder(_dev(torch_gpu, bad), _dev(torch_gpu, values))


def test_reordered_solve_is_the_permuted_solution(hip, torch_gpu):
    """The driver's flow (poisson2D/solver.cpp:495-516): RCM map, reordered matrix and right-hand side,
    solve, and the solution read back through the inverse map."""
    dim, order, n = 3, 2, 10
    count = systems.lagrange_size(dim, order, n)[2]
    csr = _shuffled(systems.lagrange(dim, order, n, conductivity=np.exp(systems.normal(4, 1.0, count))), 9)
    b = systems.uniform(1, csr.rows)
    matrix = hip.CSRMatrix(csr)
    settings = Statistics(10000, 0.0, 1e-11)
    x_natural = torch_gpu.zeros(csr.rows, dtype=torch_gpu.float64, device="cuda:0")
    solver = hip.FusedPCG(matrix, settings)
    solver.product(0, _dev(torch_gpu, b), 1, x_natural)
    mapping = hip.makeReordering(csr, "reverse-cuthill-mckee")
    reordered = matrix.reorder(mapping)
    m = _dev(torch_gpu, mapping)
    b_new = hip.reorder(m, _dev(torch_gpu, b))
    x_new = torch_gpu.zeros_like(b_new)
    solver_new = hip.FusedPCG(reordered, settings)
    solver_new.product(0, b_new, 1, x_new)
    x_back = hip.reorder(m, x_new, reverse=True)
    ref = x_natural.cpu().numpy()
    assert np.linalg.norm(x_back.cpu().numpy() - ref) <= 1e-9 * np.linalg.norm(ref)
    assert abs(int(solver_new.info.stats.iterationCount) - int(solver.info.stats.iterationCount)) <= 2


def test_full_size_rcm_recovers_gather_locality(hip, torch_gpu):
    """3D Q2 64^3 (127,263,527 nonzeros) numbered at random, as an unstructured mesh generator might leave
    it: the RCM map against the oracle, the reordered product against the permuted one, and SpMV times."""
    natural = systems.lagrange(3, 2, 64, threads=16)
    rows = natural.rows
    shuffle = np.random.default_rng(11).permutation(rows).astype(np.int32)
    base = hip.CSRMatrix(natural)
    shuffled = base.reorder(shuffle)
    host_shuffled = shuffled.copy_out()
    t0 = time.perf_counter()
    mapping = hip.makeReordering(host_shuffled, "reverse-cuthill-mckee")
    map_seconds = time.perf_counter() - t0
    assert np.array_equal(mapping, oracle.make_reordering(host_shuffled, "reverse-cuthill-mckee"))
    torch_gpu.cuda.synchronize()
    t0 = time.perf_counter()
    reordered = shuffled.reorder(mapping)
    permute_seconds = time.perf_counter() - t0
    x = systems.uniform(3, rows)
    space = hip.HIPSpace()

    def product(matrix, vector):
        y = space.makeVector(rows)
        hip.HIPCSROperator(matrix, space).product(0, vector, 1, y)
        return y

    m_shuffle, m_rcm = _dev(torch_gpu, shuffle), _dev(torch_gpu, mapping)
    x_dev = _dev(torch_gpu, x)
    x_shuffled = hip.reorder(m_shuffle, x_dev)
    x_rcm = hip.reorder(m_rcm, x_shuffled)
    y_natural = product(base, x_dev)
    y_back = hip.reorder(m_shuffle, hip.reorder(m_rcm, product(reordered, x_rcm), reverse=True), reverse=True)
    ref = y_natural.cpu().numpy()
    assert np.max(np.abs(y_back.cpu().numpy() - ref)) <= 1e-11 * np.max(np.abs(ref))
    times = {}
    for name, matrix, vector in (("natural", base, x_dev), ("shuffled", shuffled, x_shuffled),
                                 ("rcm", reordered, x_rcm)):
        best = matrix.autotune(64, repetitions=5)
        times[name] = best
    print(f"\nRCM map of {rows} rows: {map_seconds * 1e3:.0f} ms (host); device permutation of "
          f"{natural.nnz} nonzeros: {permute_seconds * 1e3:.1f} ms; FP64 SpMV (autotuned): "
          + ", ".join(f"{k} {v * 1e3:.3f} ms" for k, v in times.items()))
    assert times["rcm"] < 0.6 * times["shuffled"]

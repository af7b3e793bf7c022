"""Column-sorted block tiles (CIE_HIP_SPMV_SORTED): the column-block SpMV with each tile's blocks stored in
column order. Products are staged at each block's tile position, so every result is bitwise that of the
column-block kernel with the same width and unroll; both are checked against the oracle's rounding bound.
The sorted values follow every change of the entries, and a matrix whose tiles span more columns than the
keys hold is refused (and skipped by the autotuner)."""
import numpy as np
import pytest

import oracle
from cie_amd import systems
from cie_amd._lib import InvalidArgument, Statistics

pytestmark = pytest.mark.gpu
EPS64 = np.finfo(np.float64).eps
EPS32 = np.finfo(np.float32).eps


def _dev(torch, array, dtype=None):
    tensor = torch.from_numpy(np.ascontiguousarray(array)).cuda()
    return tensor if dtype is None else tensor.to(dtype)


def _product(hip, torch, matrix, x, y0, in_scale, out_scale, dtype):
    space = hip.HIPSpace(dtype)
    y = _dev(torch, y0, dtype)
    hip.HIPCSROperator(matrix, space).product(in_scale, _dev(torch, x, dtype), out_scale, y)
    return y.cpu().numpy()


def _bound(csr, x, y0, in_scale, out_scale, eps):
    absolute = np.abs(csr.entries) * np.abs(x[csr.column_indices])
    lengths = np.diff(csr.row_extents).astype(np.int64)
    sums = np.bincount(np.repeat(np.arange(csr.rows), lengths), weights=absolute, minlength=csr.rows)
    return 2 * (lengths + 3) * eps * (abs(out_scale) * sums + abs(in_scale) * np.abs(y0)) + 1e-300


def _mixed_rows(rng, unroll, cols):
    """Rows with more blocks than a tile, runs of adjacent columns cut at every offset, scattered rows,
    entries in the first and last columns, empty rows."""
    rows = 2000
    lengths = rng.integers(1, 9, rows)
    lengths[rng.choice(rows, 40, replace=False)] = 0
    lengths[[0, 700, 701, 1999]] = [5000, 256 * unroll + 1, 3 * 256 * unroll, 4000]
    rows_columns = []
    for n in lengths:
        if n >= 256:
            rows_columns.append(np.sort(rng.choice(cols, n, replace=False)))
        else:
            starts = np.sort(rng.choice(cols // 16, max(1, n // 3), replace=False)) * 16
            run = np.unique(np.concatenate([s + np.arange(rng.integers(1, 6)) for s in starts]))
            rows_columns.append(run[:n])
    rows_columns[5] = np.array([cols - 1])
    rows_columns[6] = np.array([0, cols - 2, cols - 1])
    lengths = np.array([len(c) for c in rows_columns])
    ext = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int32)
    return systems.CSR(rows, cols, ext, np.concatenate(rows_columns).astype(np.int32),
                       rng.standard_normal(int(ext[-1])))


def _systems(rng, width, unroll):
    yield "mixed rows, 16000 columns", _mixed_rows(rng, unroll, 16000)
    yield "mixed rows, 200000 columns", _mixed_rows(rng, unroll, 200000)
    yield "lognormal 3D Q2", systems.lagrange(3, 2, 7, conductivity=np.exp(systems.normal(4, 1.0, 7 ** 3)))
    yield "2D Q4", systems.lagrange(2, 4, 9)


@pytest.mark.parametrize("width,unroll", [(2, 4), (3, 8), (3, 16), (4, 8)])
def test_sorted_blocks_equal_column_blocks_bitwise(hip, torch_gpu, width, unroll):
    rng = np.random.default_rng(width * 100 + unroll)
    for name, csr in _systems(rng, width, unroll):
        plain, sorted_ = hip.CSRMatrix(csr), hip.CSRMatrix(csr)
        for precision in (64, 32):
            plain.set_launch(precision, "blocks", width, unroll)
            sorted_.set_launch(precision, "sorted", width, unroll)
        x = rng.standard_normal(csr.columns)
        y0 = rng.standard_normal(csr.rows)
        for in_scale, out_scale in [(0, 1), (1, -1), (0.5, 2.0)]:
            expected = oracle.csr_product(csr, in_scale, x, out_scale, y0.copy())
            for dtype, eps, slack in ((torch_gpu.float64, EPS64, 1), (torch_gpu.float32, EPS32, 4)):
                got = _product(hip, torch_gpu, sorted_, x, y0, in_scale, out_scale, dtype)
                again = _product(hip, torch_gpu, sorted_, x, y0, in_scale, out_scale, dtype)
                blocks = _product(hip, torch_gpu, plain, x, y0, in_scale, out_scale, dtype)
                assert np.array_equal(got, again), name  # bitwise run to run
                assert np.array_equal(got, blocks), name  # bitwise the column-block kernel's
                assert np.all(np.abs(got - expected) <= slack * _bound(csr, x, y0, in_scale, out_scale, eps)), name


def test_sorted_values_follow_entry_updates(hip, torch_gpu):
    csr = systems.lagrange(3, 2, 6)
    matrix = hip.CSRMatrix(csr)
    matrix.set_launch(64, "sorted", 3, 8)
    matrix.set_launch(32, "sorted", 3, 8)
    x = systems.uniform(3, csr.rows)
    zero = np.zeros(csr.rows)
    first = _product(hip, torch_gpu, matrix, x, zero, 0, 1, torch_gpu.float64)
    first32 = _product(hip, torch_gpu, matrix, x, zero, 0, 1, torch_gpu.float32)
    matrix.update_entries(_dev(torch_gpu, 2.0 * csr.entries))
    assert np.array_equal(_product(hip, torch_gpu, matrix, x, zero, 0, 1, torch_gpu.float64), 2.0 * first)
    assert np.array_equal(_product(hip, torch_gpu, matrix, x, zero, 0, 1, torch_gpu.float32), 2.0 * first32)
    dofs = systems.element_dofs(3, 2, 6)
    element = _dev(torch_gpu, systems.element_matrix(3, 2, 6))
    assembly = hip.Assembly(matrix, dofs)
    assembly.assemble(element, _dev(torch_gpu, np.full(dofs.shape[0], 4.0)))
    assert np.array_equal(_product(hip, torch_gpu, matrix, x, zero, 0, 1, torch_gpu.float64), 4.0 * first)


def test_sorted_refused_where_a_tile_spans_too_many_columns(hip, torch_gpu):
    """Two rows in one tile whose blocks lie 1.2M columns apart: the 20-bit column offsets cannot hold
    them, so the sorted launch fails with an error and the autotuner keeps another kernel."""
    cols = 1_200_000
    csr = systems.CSR(2, cols, np.array([0, 3, 6], np.int32),
                      np.array([0, 1, 2, cols - 3, cols - 2, cols - 1], np.int32), np.arange(1.0, 7.0))
    matrix = hip.CSRMatrix(csr)
    matrix.set_launch(64, "sorted", 3, 8)
    x = np.ones(cols)
    with pytest.raises(InvalidArgument):
        _product(hip, torch_gpu, matrix, x, np.zeros(2), 0, 1, torch_gpu.float64)
    matrix.autotune(64, 1, kernels=("blocks", "sorted"))
    assert matrix.launch(64)[0] == "blocks"
    got = _product(hip, torch_gpu, matrix, x, np.zeros(2), 0, 1, torch_gpu.float64)
    assert np.array_equal(got, np.array([6.0, 15.0]))


def test_pcg_on_sorted_blocks(hip, torch_gpu):
    """The fused FP64 solver on sorted tiles gives the column-block kernel's result bitwise (its products use
    one block per tile with either kernel); the mixed solver's refinement runs one resident round of blocks,
    whose partial sums may group differently, so it agrees to rounding."""
    k = np.exp(systems.normal(2, 1.0, 8 ** 3))
    csr = systems.lagrange(3, 2, 8, conductivity=k)
    b = systems.uniform(1, csr.rows)
    x_ref, s_ref, _, _ = oracle.pcg(csr, b, relative_tolerance=1e-10)
    results = {}
    for kernel in ("blocks", "sorted"):
        matrix = hip.CSRMatrix(csr)
        matrix.set_launch(64, kernel, 3, 8)
        matrix.set_launch(32, kernel, 3, 8)
        out = []
        for solver in (hip.FusedPCG(matrix, Statistics(20000, 0.0, 1e-10)),
                       hip.MixedPCG(matrix, Statistics(20000, 0.0, 1e-10))):
            x = torch_gpu.zeros(csr.rows, dtype=torch_gpu.float64, device="cuda:0")
            solver.product(0, _dev(torch_gpu, b), 1, x)
            assert solver.info.converged == 1 and solver.info.trueRelativeResidual < 2e-10
            out.append((x.cpu().numpy(), int(solver.getStats().iterationCount)))
        results[kernel] = out
    (fs, fits), (ms, mits) = results["sorted"]
    (fb, fitb), (mb, mitb) = results["blocks"]
    assert fits == fitb and np.array_equal(fs, fb)
    assert abs(mits - mitb) <= 2 and np.linalg.norm(ms - mb) / np.linalg.norm(mb) < 1e-8
    assert abs(results["sorted"][0][1] - int(s_ref.iterationCount)) <= 2
    assert np.linalg.norm(results["sorted"][0][0] - x_ref) / np.linalg.norm(x_ref) < 1e-8

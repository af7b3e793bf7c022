"""Paired block tiles (CIE_HIP_SPMV_PAIRED): rows covered by 1 x wide and 1 x narrow column blocks, each
width's blocks column-sorted within its tile.

- The device cover is checked block count for block count against the rule restated row by row here: a
  block starts at the first column not yet covered (moved left to columns - width at the last columns) and
  is wide when an entry of [start + narrow, start + wide) follows.
- Products are checked against the oracle's row loop within the rounding bound of a reordered sum, for
  every pair and unroll, several scalings, both precisions; they are bitwise reproducible run to run and
  follow every change of the entries.
- A matrix with a row of more blocks of one width than a tile holds, or with a tile whose blocks of one
  width span more columns than the key holds, is refused at its first product and skipped by the autotuner.
- The fused FP64 and the mixed solvers on paired launches agree with the oracle's Jacobi-PCG.
"""
import numpy as np
import pytest

import oracle
from cie_amd import systems
from cie_amd._lib import InvalidArgument, Statistics

pytestmark = pytest.mark.gpu
EPS64 = np.finfo(np.float64).eps
EPS32 = np.finfo(np.float32).eps
LAUNCHES = [(53, 2), (53, 4), (32, 2), (32, 4)]


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


def cover_counts(csr, wide, narrow):
    """(wide, narrow) blocks of the paired cover, restated row by row."""
    counts = [0, 0]
    ext, cols = csr.row_extents, csr.column_indices
    for row in range(csr.rows):
        columns = cols[ext[row]:ext[row + 1]]
        limit = -1
        for i, column in enumerate(columns):
            if column < limit:
                continue
            following = columns[i + 1:i + wide]
            is_wide = bool(np.any((following >= column + narrow) & (following < column + wide)))
            width = wide if is_wide else narrow
            start = min(int(column), csr.columns - width)
            limit = start + width
            counts[0 if is_wide else 1] += 1
    return tuple(counts)


def _runs(rng, rows, cols, longest=6):
    """Rows of runs of 1..longest adjacent columns at random offsets, entries in the first and last columns,
    empty rows."""
    rows_columns = []
    for r in range(rows):
        if r % 97 == 13:
            rows_columns.append(np.array([], dtype=np.int64))
            continue
        starts = np.sort(rng.choice(cols // 8, rng.integers(1, 12), replace=False)) * 8
        run = np.unique(np.concatenate([s + np.arange(rng.integers(1, longest + 1)) for s in starts]))
        rows_columns.append(run[run < cols])
    rows_columns[5] = np.array([cols - 1])
    rows_columns[6] = np.array([0, cols - 4, cols - 2, cols - 1])
    rows_columns[7] = np.arange(cols - 7, cols)
    lengths = np.array([len(c) for c in rows_columns])
    ext = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int32)
    return systems.CSR(rows, cols, ext, np.concatenate(rows_columns).astype(np.int32),
                       rng.standard_normal(int(ext[-1])))


def _systems(rng):
    yield "runs of 1..6 columns", _runs(rng, 3000, 20000)
    yield "runs, 300000 columns", _runs(rng, 2000, 300000)
    yield "3D Q2, lognormal", systems.lagrange(3, 2, 7, conductivity=np.exp(systems.normal(4, 1.0, 7 ** 3)))
    yield "3D Q4", systems.lagrange(3, 4, 3)
    yield "2D Q4 Legendre", systems.lagrange(2, 4, 9, basis="legendre")
    yield "2D 5-point", systems.five_point_2d(64).matrix


def test_cover_counts_follow_the_rule(hip, torch_gpu):
    rng = np.random.default_rng(7)
    for name, csr in _systems(rng):
        matrix = hip.CSRMatrix(csr)
        for width in (53, 32):
            expected = cover_counts(csr, width // 10, width % 10)
            assert matrix.paired_blocks(width) == expected, (name, width)


def test_q2_cover_needs_no_padding(hip, torch_gpu):
    """3D Q2 rows hold runs of 5 and 3 adjacent columns (4 and 2 where a run meets an eliminated boundary):
    1 x 5 and 1 x 3 blocks store each run as one block, padding only the runs of 4 (7.3% of the entries on
    8^3 elements, 3.4% on 16^3, shrinking with the boundary's share)."""
    csr = systems.q2_laplace_3d(8).matrix
    wide, narrow = hip.CSRMatrix(csr).paired_blocks(53)
    assert 5 * wide + 3 * narrow >= csr.nnz
    padding = (5 * wide + 3 * narrow - csr.nnz) / csr.nnz
    assert padding < 0.08, padding  # 1 x 3 blocks alone pad a third of every run of 5
    blocks3, _ = hip.CSRMatrix(csr).column_blocks(3)
    assert (5 * 4 + 4) * wide + (3 * 4 + 4) * narrow < 0.88 * (3 * 4 + 4) * blocks3  # FP32 bytes


@pytest.mark.parametrize("width,unroll", LAUNCHES)
def test_paired_products_within_the_rounding_bound(hip, torch_gpu, width, unroll):
    rng = np.random.default_rng(width * 10 + unroll)
    for name, csr in _systems(rng):
        matrix = hip.CSRMatrix(csr)
        for precision in (64, 32):
            matrix.set_launch(precision, "paired", width, unroll)
        x = rng.standard_normal(csr.columns)
        y0 = rng.standard_normal(csr.rows)
        for in_scale, out_scale in [(0, 1), (1, -1), (0.5, 2.0), (-1, 0.25)]:
            expected = oracle.csr_product(csr, in_scale, x, out_scale, y0.copy())
            for dtype, eps, slack in ((torch_gpu.float64, EPS64, 1), (torch_gpu.float32, EPS32, 4)):
                got = _product(hip, torch_gpu, matrix, x, y0, in_scale, out_scale, dtype)
                bound = slack * _bound(csr, x, y0, in_scale, out_scale, eps)
                assert np.all(np.abs(got - expected) <= bound), (name, dtype, in_scale, out_scale)
        again = _product(hip, torch_gpu, matrix, x, y0, 0.5, 2.0, torch_gpu.float64)
        assert np.array_equal(again, _product(hip, torch_gpu, matrix, x, y0, 0.5, 2.0, torch_gpu.float64))


def test_paired_values_follow_entry_updates(hip, torch_gpu):
    rng = np.random.default_rng(3)
    csr = systems.lagrange(3, 2, 5)
    matrix = hip.CSRMatrix(csr)
    matrix.set_launch(64, "paired", 53, 2)
    matrix.set_launch(32, "paired", 53, 4)
    x = rng.standard_normal(csr.columns)
    zeros = np.zeros(csr.rows)
    for _ in range(2):
        entries = rng.standard_normal(csr.nnz)
        matrix.update_entries(_dev(torch_gpu, entries))
        changed = systems.CSR(csr.rows, csr.columns, csr.row_extents, csr.column_indices, entries)
        expected = oracle.csr_product(changed, 0.0, x, 1.0, zeros.copy())
        for dtype, eps, slack in ((torch_gpu.float64, EPS64, 1), (torch_gpu.float32, EPS32, 4)):
            got = _product(hip, torch_gpu, matrix, x, zeros, 0, 1, dtype)
            assert np.all(np.abs(got - expected) <= slack * _bound(changed, x, zeros, 0, 1, eps))


def test_rows_longer_than_a_tile_are_refused(hip, torch_gpu):
    rng = np.random.default_rng(5)
    cols = 40000
    long_row = np.sort(rng.choice(cols // 4, 1500, replace=False)) * 4  # 1500 isolated entries: narrow blocks
    rows_columns = [np.arange(r * 3, r * 3 + 3) for r in range(200)] + [long_row]
    ext = np.concatenate([[0], np.cumsum([len(c) for c in rows_columns])]).astype(np.int32)
    csr = systems.CSR(len(rows_columns), cols, ext, np.concatenate(rows_columns).astype(np.int32),
                      rng.standard_normal(int(ext[-1])))
    x = rng.standard_normal(cols)
    for unroll in (2, 4):  # 512 and 1024 blocks of one width per tile
        matrix = hip.CSRMatrix(csr)
        matrix.set_launch(64, "paired", 53, unroll)
        with pytest.raises(InvalidArgument):
            _product(hip, torch_gpu, matrix, x, np.zeros(csr.rows), 0, 1, torch_gpu.float64)
    matrix = hip.CSRMatrix(csr)
    with pytest.raises(InvalidArgument):
        matrix.autotune(64, 1, kernels=("paired",))
    matrix.autotune(64, 1, kernels=("paired", "tiled"))
    assert matrix.launch(64)[0] == "tiled"
    with pytest.raises(InvalidArgument):
        hip.CSRMatrix(systems.CSR(1, 4, np.array([0, 1], np.int32), np.array([0], np.int32),
                                  np.ones(1))).set_launch(64, "paired", 53, 2)  # fewer columns than the wide block


@pytest.mark.parametrize("width", [53, 32])
def test_fused_and_mixed_solves_on_paired_tiles(hip, torch_gpu, width):
    w = systems.q2_laplace_3d(10)
    x_ref, s_ref, _, _ = oracle.pcg(w.matrix, w.rhs, relative_tolerance=1e-10)
    matrix = hip.CSRMatrix(w.matrix)
    matrix.set_launch(64, "paired", width, 2)
    matrix.set_launch(32, "paired", width, 4)
    b = _dev(torch_gpu, w.rhs)
    fp64 = hip.FusedPCG(matrix, Statistics(10000, 0.0, 1e-10))
    x = torch_gpu.zeros_like(b)
    fp64.product(0, b, 1, x)
    assert fp64.info.converged == 1
    assert abs(int(fp64.getStats().iterationCount) - int(s_ref.iterationCount)) <= 1
    assert np.linalg.norm(x.cpu().numpy() - x_ref) / np.linalg.norm(x_ref) < 1e-10
    mixed = hip.MixedPCG(matrix, Statistics(10000, 0.0, 1e-10))
    xm = torch_gpu.zeros_like(b)
    mixed.product(0, b, 1, xm)
    assert mixed.info.converged == 1 and mixed.info.trueRelativeResidual < 1e-10
    assert np.linalg.norm(xm.cpu().numpy() - x_ref) / np.linalg.norm(x_ref) < 1e-8


def test_paired_refused_where_a_tile_spans_too_many_columns(hip, torch_gpu):
    """Two rows in one tile whose 1 x 3 blocks lie 1.2M columns apart: the key's 20-bit column offset cannot
    hold them, so the paired launch fails with an error and the autotuner keeps another kernel."""
    cols = 1_200_000
    csr = systems.CSR(2, cols, np.array([0, 3, 6], np.int32),
                      np.array([0, 1, 2, cols - 3, cols - 2, cols - 1], np.int32), np.arange(1.0, 7.0))
    x = np.ones(cols)
    matrix = hip.CSRMatrix(csr)
    matrix.set_launch(64, "paired", 53, 2)
    with pytest.raises(InvalidArgument):
        _product(hip, torch_gpu, matrix, x, np.zeros(2), 0, 1, torch_gpu.float64)
    matrix = hip.CSRMatrix(csr)
    matrix.autotune(64, 1, kernels=("paired", "tiled"))
    assert matrix.launch(64)[0] == "tiled"
    assert np.array_equal(_product(hip, torch_gpu, matrix, x, np.zeros(2), 0, 1, torch_gpu.float64),
                          np.array([6.0, 15.0]))

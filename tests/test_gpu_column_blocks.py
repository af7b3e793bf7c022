"""The column-block SpMV (1 x w blocks of adjacent columns, CIE_HIP_SPMV_BLOCKS): its derived block values
follow every change of the entries, it is refused where the matrix is narrower than a block, and its
padding never contributes to a row, also where x is not finite."""
import numpy as np
import pytest

import oracle
from cie_amd import systems
from cie_amd._lib import InvalidArgument

pytestmark = pytest.mark.gpu


def _dev(torch, array, dtype=None):
    tensor = torch.from_numpy(np.ascontiguousarray(array)).cuda()
    return tensor if dtype is None else tensor.to(dtype)


def _product(hip, torch, matrix, x, dtype=None):
    space = hip.HIPSpace(dtype or torch.float64)
    y = space.makeVector(matrix.rows)
    hip.HIPCSROperator(matrix, space).product(0, _dev(torch, x, dtype), 1, y)
    return y.cpu().numpy()


def test_block_values_follow_entry_updates(hip, torch_gpu):
    """update_entries, device assembly: the next block product uses the new values (2 K x = 2 (K x) exactly)."""
    csr = systems.lagrange(3, 2, 6)
    matrix = hip.CSRMatrix(csr)
    matrix.set_launch(64, "blocks", 3, 8)
    matrix.set_launch(32, "blocks", 3, 8)
    x = systems.uniform(3, csr.rows)
    first = _product(hip, torch_gpu, matrix, x)
    first32 = _product(hip, torch_gpu, matrix, x, torch_gpu.float32)
    matrix.update_entries(_dev(torch_gpu, 2.0 * csr.entries))
    assert np.array_equal(_product(hip, torch_gpu, matrix, x), 2.0 * first)
    assert np.array_equal(_product(hip, torch_gpu, matrix, x, torch_gpu.float32), 2.0 * first32)
    dofs = systems.element_dofs(3, 2, 6)
    element = _dev(torch_gpu, systems.element_matrix(3, 2, 6))
    assembly = hip.Assembly(matrix, dofs)
    assembly.assemble(element, _dev(torch_gpu, np.full(dofs.shape[0], 4.0)))
    assert np.array_equal(_product(hip, torch_gpu, matrix, x), 4.0 * first)
    assembly.assemble(element)
    assert np.array_equal(_product(hip, torch_gpu, matrix, x), first)


def test_blocks_refused_for_narrow_matrices(hip, torch_gpu):
    csr = systems.CSR(3, 2, np.array([0, 1, 2, 2], np.int32), np.array([0, 1], np.int32), np.array([1.0, 2.0]))
    matrix = hip.CSRMatrix(csr)
    with pytest.raises(InvalidArgument):
        matrix.set_launch(64, "blocks", 3, 8)
    matrix.set_launch(64, "blocks", 2, 8)
    assert np.array_equal(_product(hip, torch_gpu, matrix, np.array([3.0, 5.0])), np.array([3.0, 10.0, 0.0]))
    with pytest.raises(InvalidArgument):
        matrix.set_launch(64, "blocks", 5, 8)
    with pytest.raises(InvalidArgument):
        matrix.set_launch(64, "blocks", 3, 32)


def test_padding_ignores_non_finite_x(hip, torch_gpu):
    """Row 0 holds columns 0 and 2: its 1 x 3 block pads column 1, where x is inf or NaN. The CSR product is
    finite there, and so is the block product. A stored zero is skipped as well (the CSR loop would give
    0 * inf = NaN in row 1)."""
    rows = 4
    ext = np.array([0, 2, 4, 5, 6], np.int32)
    cols = np.array([0, 2, 1, 3, 3, 5], np.int32)
    vals = np.array([2.0, -1.0, 0.0, 4.0, 1.5, -3.0])
    csr = systems.CSR(rows, 6, ext, cols, vals)
    matrix = hip.CSRMatrix(csr)
    matrix.set_launch(64, "blocks", 3, 4)
    matrix.set_launch(32, "blocks", 3, 4)
    for bad in (np.inf, np.nan):
        x = np.array([1.0, bad, 2.0, 3.0, bad, 0.5])
        expected = oracle.csr_product(csr, 0.0, x, 1.0, np.zeros(rows))
        got = _product(hip, torch_gpu, matrix, x)
        assert got[0] == expected[0] == 0.0 and got[2] == expected[2] == 4.5 and got[3] == expected[3] == -1.5
        assert np.isnan(expected[1]) and got[1] == 12.0
        got32 = _product(hip, torch_gpu, matrix, x, torch_gpu.float32)
        assert np.array_equal(got32[[0, 2, 3]], expected[[0, 2, 3]]) and got32[1] == 12.0


def _host_block_count(csr, width):
    """Blocks of the covering rule, restated on the host: each block starts at the row's first uncovered
    column (at columns - width when it would pass the last column) and covers width adjacent columns."""
    count = 0
    ext, cols = csr.row_extents, csr.column_indices
    for row in range(csr.rows):
        k, end = int(ext[row]), int(ext[row + 1])
        while k < end:
            start = min(int(cols[k]), csr.columns - width)
            count += 1
            while k < end and cols[k] < start + width:
                k += 1
    return count


@pytest.mark.parametrize("dim,order,n", [(3, 2, 5), (2, 4, 7), (3, 1, 6)])
def test_block_counts_match_the_covering_rule(hip, torch_gpu, dim, order, n):
    csr = systems.lagrange(dim, order, n)
    matrix = hip.CSRMatrix(csr)
    for width in (2, 3, 4):
        blocks, column_bytes = matrix.column_blocks(width)
        assert blocks == _host_block_count(csr, width), width
        assert width * blocks >= csr.nnz and column_bytes == 2
    # scattered rows of 0-6 columns, some at the last columns (blocks shifted to columns - width)
    rng = np.random.default_rng(4)
    rows = [np.unique(rng.integers(0, 40, rng.integers(0, 7))) for _ in range(50)]
    rows[7] = np.array([38, 39])
    ext = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    scattered = systems.CSR(50, 40, ext, np.concatenate(rows).astype(np.int32), np.ones(int(ext[-1])))
    matrix = hip.CSRMatrix(scattered)
    assert matrix.column_blocks(3)[0] == _host_block_count(scattered, 3)
    with pytest.raises(InvalidArgument):
        matrix.column_blocks(5)

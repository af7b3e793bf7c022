"""Empty inputs through the C ABI: matrices without entries or without
columns, assembly of no elements, element stiffness of no cells, reordering
of patterns without couplings, and a mask that drops every row. Each result
is checked against the oracle or its closed form (an empty row sums to +0)."""
import numpy as np
import pytest

import oracle
from cie_amd import systems
from cie_amd._lib import InvalidArgument, library

pytestmark = pytest.mark.gpu

SCALES = [(0.0, 1.0), (1.0, 1.0), (-1.0, 2.0), (2.0, -1.0), (0.5, 0.0)]


def _dev(torch, array):
    return torch.from_numpy(np.ascontiguousarray(array)).cuda()


def _empty(rows, columns):
    return systems.CSR(rows, columns, np.zeros(rows + 1, dtype=np.int32), np.zeros(0, dtype=np.int32),
                       np.zeros(0, dtype=np.float64))


@pytest.mark.parametrize("rows", [1, 7, 300, 5000])
def test_products_of_a_matrix_without_entries(hip, torch_gpu, rows):
    """Every launch family on rows x rows with nnz = 0: out = inScale out (CSROperator.cpp:33-111)."""
    csr = _empty(rows, rows)
    matrix = hip.CSRMatrix(csr)
    rng = np.random.default_rng(rows)
    x, y0 = rng.standard_normal(rows), rng.standard_normal(rows)
    launches = [("tiled", 256, 4), ("tiled", 256, 16), ("staged", 512, 4), ("vector", 4, 4)]
    block_launches = [("blocks", 3, 4), ("sorted", 3, 8), ("paired", 53, 2), ("paired", 32, 4)]
    if rows >= 5:  # block forms need at least as many columns as their (wide) block
        launches += block_launches
    else:
        for kernel, width, unroll in block_launches:
            with pytest.raises(InvalidArgument):
                matrix.set_launch(64, kernel, width, unroll)
    for kernel, width, unroll in launches:
        matrix.set_launch(64, kernel, width, unroll)
        op = hip.HIPCSROperator(matrix, hip.HIPSpace())
        for in_scale, out_scale in SCALES:
            y = _dev(torch_gpu, y0)
            op.product(in_scale, _dev(torch_gpu, x), out_scale, y)
            expected = oracle.csr_product(csr, in_scale, x, out_scale, y0.copy())
            assert np.array_equal(y.cpu().numpy(), expected), (kernel, unroll, in_scale, out_scale)


def test_product_of_a_matrix_without_columns(hip, torch_gpu):
    """rows x 0: nothing to gather, so only the output scaling remains."""
    rows = 2000
    matrix = hip.CSRMatrix(_empty(rows, 0))
    y0 = np.random.default_rng(1).standard_normal(rows)
    y = _dev(torch_gpu, y0)
    spare = torch_gpu.zeros(1, dtype=torch_gpu.float64, device="cuda")  # a valid pointer for the empty input
    context = hip.Context.default()
    status = library().cie_hip_csr_product_f64(context.handle, matrix.handle, 2.0, spare.data_ptr(), 0, 3.0,
                                               y.data_ptr(), rows)
    assert status == 0
    torch_gpu.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), 2.0 * y0)


def test_assembly_of_no_elements_is_zero(hip, torch_gpu):
    csr = systems.lagrange(2, 2, 4)
    matrix = hip.CSRMatrix(csr)
    npe = 9
    assembly = hip.Assembly(matrix, np.zeros((0, npe), dtype=np.int32))
    element = _dev(torch_gpu, np.ones((npe, npe)))
    assembly.assemble(element)
    out = matrix.copy_out()
    assert np.array_equal(out.row_extents, csr.row_extents)
    assert np.array_equal(out.column_indices, csr.column_indices)
    assert np.all(out.entries == 0) and not np.signbit(out.entries).any()
    vector = assembly.assemble_vector(_dev(torch_gpu, np.ones(npe)))
    assert np.array_equal(vector.cpu().numpy(), np.zeros(csr.rows))


def test_element_stiffness_of_no_cells(hip, torch_gpu):
    moments = _dev(torch_gpu, systems.gradient_moments(2, 2))
    jacobians = torch_gpu.zeros((0, 2, 2), dtype=torch_gpu.float64, device="cuda")
    out = hip.elementStiffness(moments, jacobians)
    assert tuple(out.shape) == (0, 9, 9)


def test_reordering_of_a_pattern_without_couplings(hip, torch_gpu):
    """No entries: every strategy keeps the natural order (reorder.cpp:33-55), and the permuted matrix is empty."""
    csr = _empty(50, 50)
    for strategy in ("none", "cuthill-mckee", "reverse-cuthill-mckee"):
        mapping = hip.makeReordering(csr, strategy)
        assert np.array_equal(mapping, oracle.make_reordering(csr, strategy))
        assert np.array_equal(mapping, np.arange(50))
    matrix = hip.CSRMatrix(csr)
    shuffle = np.random.default_rng(2).permutation(50).astype(np.int32)
    out = matrix.reorder(shuffle).copy_out()
    assert out.nnz == 0 and np.array_equal(out.row_extents, np.zeros(51, dtype=np.int32))
    assert hip.makeReordering(_empty(0, 0), "reverse-cuthill-mckee").size == 0


def test_mask_that_drops_every_row(hip, torch_gpu):
    """threshold 0 masks everything: the masked matrix is empty and its product leaves inScale out."""
    csr = systems.lagrange(2, 3, 5, basis="legendre")
    mask_values = systems.ansatz_mask(2, 3, 5)
    matrix = hip.CSRMatrix(csr)
    mask = hip.Mask(mask_values)
    masked = hip.maskedMatrix(matrix, mask, 0)
    assert masked.nnz == 0 and masked.rows == csr.rows
    rng = np.random.default_rng(3)
    x, y0 = rng.standard_normal(csr.rows), rng.standard_normal(csr.rows)
    op = hip.MaskedCSROperator(matrix, mask, 0, hip.HIPSpace())
    for in_scale, out_scale in SCALES:
        y = _dev(torch_gpu, y0)
        op.product(in_scale, _dev(torch_gpu, x), out_scale, y)
        expected = oracle.masked_csr_product(csr, mask_values, 0, in_scale, x, out_scale, y0.copy())
        assert np.array_equal(y.cpu().numpy(), expected), (in_scale, out_scale)

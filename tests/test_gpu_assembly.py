"""Device assembly of K = sum_e k_e P_e^T K_e0 P_e (the element scatter of
fem/packages/graph/impl/Assembler_impl.hpp:156-194) against the structured
generator, which sums every slot in element order on the CPU.

The assembled values are compared through the staged SpMV, whose rows are
bit-identical to the sequential CPU row loop: A x for random x must match the
oracle's product over the generator's values exactly."""
import time

import numpy as np
import pytest

import oracle
from cie_amd import adjoint, systems
from cie_amd._lib import ShapeError, Statistics

pytestmark = pytest.mark.gpu


def _dev(torch, array, dtype=None):
    tensor = torch.from_numpy(np.ascontiguousarray(array)).cuda()
    return tensor if dtype is None else tensor.to(dtype)


def _products_match(hip, torch, matrix, expected, seeds=(3, 4), threads=1):
    matrix.set_launch(64, "staged", 1024, 8)
    matrix.set_launch(32, "staged", 1024, 8)
    space, space32 = hip.HIPSpace(), hip.HIPSpace(torch.float32)
    for seed in seeds:
        x = systems.uniform(seed, expected.columns) - 0.5
        reference = oracle.csr_product(expected, 0.0, x, 1.0, np.zeros(expected.rows), threads=threads)
        y = space.makeVector(expected.rows)
        hip.HIPCSROperator(matrix, space).product(0, _dev(torch, x), 1, y)
        assert np.array_equal(y.cpu().numpy(), reference)
        # the FP32 plane is the rounded FP64 values
        y32 = space32.makeVector(expected.rows)
        hip.HIPCSROperator(matrix, space32).product(0, _dev(torch, x, torch.float32), 1, y32)
        assert np.linalg.norm(y32.cpu().numpy() - reference) <= 1e-5 * np.linalg.norm(reference)


@pytest.mark.parametrize("dim,order,n", [(3, 2, 6), (2, 4, 9), (3, 4, 3), (2, 1, 17), (3, 1, 7)])
def test_assembly_matches_generator_bitwise(hip, torch_gpu, dim, order, n):
    count = systems.lagrange_size(dim, order, n)[2]
    k = np.exp(systems.normal(11, 1.0, count))
    expected = systems.lagrange(dim, order, n, conductivity=k)
    matrix = hip.CSRMatrix(systems.lagrange(dim, order, n))  # k = 1 values on the same pattern
    assembly = hip.Assembly(matrix, systems.element_dofs(dim, order, n))
    element = _dev(torch_gpu, systems.element_matrix(dim, order, n))
    assembly.assemble(element, _dev(torch_gpu, k))
    _products_match(hip, torch_gpu, matrix, expected)
    # scale NULL means k = 1; assembling twice gives the same values (every slot starts from zero)
    assembly.assemble(element)
    assembly.assemble(element)
    _products_match(hip, torch_gpu, matrix, systems.lagrange(dim, order, n), seeds=(5,))


@pytest.mark.parametrize("dim,order,n", [(2, 2, 7), (3, 2, 4), (3, 4, 2)])
def test_assembly_of_per_element_matrices_matches_oracle(hip, torch_gpu, dim, order, n):
    """One random matrix per element (the general addContribution input) against the oracle's scatter."""
    pattern = systems.lagrange(dim, order, n)
    dofs = systems.element_dofs(dim, order, n)
    count, nodes = dofs.shape
    rng = np.random.default_rng(dim * 10 + order)
    matrices = rng.standard_normal((count, nodes, nodes))
    scale = rng.uniform(0.5, 2.0, count)
    expected = systems.CSR(pattern.rows, pattern.rows, pattern.row_extents, pattern.column_indices,
                           oracle.assemble(pattern, dofs, matrices, scale))
    matrix = hip.CSRMatrix(pattern)
    assembly = hip.Assembly(matrix, dofs)
    assembly.assemble(_dev(torch_gpu, matrices), _dev(torch_gpu, scale))
    _products_match(hip, torch_gpu, matrix, expected)
    unscaled = systems.CSR(pattern.rows, pattern.rows, pattern.row_extents, pattern.column_indices,
                           oracle.assemble(pattern, dofs, matrices))
    assembly.assemble(_dev(torch_gpu, matrices))
    _products_match(hip, torch_gpu, matrix, unscaled, seeds=(7,))


@pytest.mark.parametrize("dim,order,n", [(2, 1, 9), (2, 4, 5), (3, 2, 5)])
def test_vector_assembly_matches_oracle_bitwise(hip, torch_gpu, dim, order, n):
    pattern = systems.lagrange(dim, order, n)
    dofs = systems.element_dofs(dim, order, n)
    count, nodes = dofs.shape
    rng = np.random.default_rng(5)
    vectors, shared, scale = rng.standard_normal((count, nodes)), rng.standard_normal(nodes), rng.uniform(0.5, 2, count)
    assembly = hip.Assembly(hip.CSRMatrix(pattern), dofs)
    got = assembly.assemble_vector(_dev(torch_gpu, vectors), _dev(torch_gpu, scale)).cpu().numpy()
    assert np.array_equal(got, oracle.assemble_vector(pattern.rows, dofs, vectors, scale))
    got = assembly.assemble_vector(_dev(torch_gpu, shared)).cpu().numpy()
    assert np.array_equal(got, oracle.assemble_vector(pattern.rows, dofs, shared))


def test_assembly_then_solve_uses_new_values(hip, torch_gpu):
    dim, order, n = 3, 2, 8
    count = systems.lagrange_size(dim, order, n)[2]
    matrix = hip.CSRMatrix(systems.lagrange(dim, order, n))
    assembly = hip.Assembly(matrix, systems.element_dofs(dim, order, n))
    element = _dev(torch_gpu, systems.element_matrix(dim, order, n))
    b = systems.uniform(1, matrix.rows)
    solver = hip.FusedPCG(matrix, Statistics(10000, 0.0, 1e-11))
    for seed in (5, 6):  # the cached Jacobi diagonal must follow the new values
        k = np.exp(systems.normal(seed, 1.0, count))
        assembly.assemble(element, _dev(torch_gpu, k))
        x = torch_gpu.zeros(matrix.rows, dtype=torch_gpu.float64, device="cuda:0")
        solver.product(0, _dev(torch_gpu, b), 1, x)
        x_ref, _, _, _ = oracle.pcg(systems.lagrange(dim, order, n, conductivity=k), b, relative_tolerance=1e-11)
        assert np.linalg.norm(x.cpu().numpy() - x_ref) <= 1e-8 * np.linalg.norm(x_ref)


def test_assembly_rejects_a_pattern_without_a_coupling(hip, torch_gpu):
    dim, order, n = 2, 2, 4
    full = systems.lagrange(dim, order, n)
    drop = int(full.row_extents[1]) - 1  # the last (off-diagonal) entry of row 0
    assert full.column_indices[drop] != 0
    keep = np.ones(full.nnz, dtype=bool)
    keep[drop] = False
    extents = full.row_extents.copy()
    extents[1:] -= 1
    pruned = systems.CSR(full.rows, full.columns, extents, full.column_indices[keep], full.entries[keep])
    matrix = hip.CSRMatrix(pruned)
    with pytest.raises(ShapeError):
        hip.Assembly(matrix, systems.element_dofs(dim, order, n))
    # a row index beyond the matrix is a shape error too
    dofs = systems.element_dofs(dim, order, n)
    dofs[0, 0] = full.rows
    with pytest.raises(ShapeError):
        hip.Assembly(hip.CSRMatrix(full), dofs)


def test_adjoint_problem_assembles_on_the_device(hip, torch_gpu):
    dim, order, n = 3, 2, 5
    rows, _, count, _ = systems.lagrange_size(dim, order, n)
    problem = adjoint.LagrangeProblem(dim, order, n, np.zeros(rows), np.zeros(rows))
    g = systems.normal(7, 0.7, count)
    matrix = problem.assemble(g)
    _products_match(hip, torch_gpu, matrix, systems.lagrange(dim, order, n, conductivity=np.exp(g)), seeds=(8,))
    assert problem.assemble(g + 0.25) is matrix  # the pattern and plan are kept, only the values change
    _products_match(hip, torch_gpu, matrix, systems.lagrange(dim, order, n, conductivity=np.exp(g + 0.25)),
                    seeds=(9,))


def test_full_size_assembly_q2_64(hip, torch_gpu):
    """3D Q2 64^3 (2,048,383 rows, 127,263,527 nonzeros, 262,144 elements), lognormal k."""
    dim, order, n = 3, 2, 64
    count = systems.lagrange_size(dim, order, n)[2]
    k = np.exp(systems.normal(2, 1.0, count))
    t0 = time.perf_counter()
    expected = systems.lagrange(dim, order, n, conductivity=k, threads=16)
    host_seconds = time.perf_counter() - t0
    assert expected.nnz == 127263527
    matrix = hip.CSRMatrix(systems.lagrange(dim, order, n, threads=16))
    assembly = hip.Assembly(matrix, systems.element_dofs(dim, order, n))
    element = _dev(torch_gpu, systems.element_matrix(dim, order, n))
    k_dev = _dev(torch_gpu, k)
    assembly.assemble(element, k_dev)
    t0 = time.perf_counter()
    for _ in range(3):
        assembly.assemble(element, k_dev)
    device_seconds = (time.perf_counter() - t0) / 3
    _products_match(hip, torch_gpu, matrix, expected, seeds=(9,), threads=16)
    print(f"\nassembly of {expected.nnz} nonzeros: device {device_seconds * 1e3:.2f} ms "
          f"(FP64 + FP32 planes), host generator {host_seconds * 1e3:.0f} ms (16 threads)")

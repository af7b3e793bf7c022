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


def _entries_match(matrix, expected):
    """The stored FP64 values, read back, equal the expected CSR's bit for bit."""
    got = matrix.copy_out()
    assert np.array_equal(got.row_extents, expected.row_extents)
    assert np.array_equal(got.column_indices, expected.column_indices)
    assert np.array_equal(got.entries, expected.entries)


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
    _entries_match(matrix, expected)
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
    _entries_match(matrix, expected)
    _products_match(hip, torch_gpu, matrix, expected)
    unscaled = systems.CSR(pattern.rows, pattern.rows, pattern.row_extents, pattern.column_indices,
                           oracle.assemble(pattern, dofs, matrices))
    assembly.assemble(_dev(torch_gpu, matrices))
    _products_match(hip, torch_gpu, matrix, unscaled, seeds=(7,))


def test_assembly_with_a_row_listed_twice_in_one_element(hip, torch_gpu):
    """A dof map that names one row at two local nodes of an element (the scatter adds both)."""
    dim, order, n = 2, 2, 5
    pattern = systems.lagrange(dim, order, n)
    dofs = systems.element_dofs(dim, order, n)
    e = 7
    kept = np.flatnonzero(dofs[e] >= 0)
    dofs[e, kept[1]] = dofs[e, kept[0]]
    rng = np.random.default_rng(6)
    matrices = rng.standard_normal((dofs.shape[0], dofs.shape[1], dofs.shape[1]))
    expected = systems.CSR(pattern.rows, pattern.rows, pattern.row_extents, pattern.column_indices,
                           oracle.assemble(pattern, dofs, matrices))
    matrix = hip.CSRMatrix(pattern)
    hip.Assembly(matrix, dofs).assemble(_dev(torch_gpu, matrices))
    _entries_match(matrix, expected)


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


@pytest.mark.parametrize("dim,order,basis", [(2, 2, "lagrange"), (2, 4, "legendre"), (3, 1, "lagrange"),
                                             (3, 2, "lagrange"), (3, 4, "lagrange")])
def test_element_stiffness_matches_oracle_bitwise(hip, torch_gpu, dim, order, basis):
    rng = np.random.default_rng(dim * 10 + order)
    moments = systems.gradient_moments(dim, order, basis)
    count = 700
    jac = rng.standard_normal((count, dim, dim)) * 0.3 + np.eye(dim) * rng.uniform(0.5, 2.0, (count, 1, 1))
    k = np.exp(rng.standard_normal(count))
    got = hip.elementStiffness(_dev(torch_gpu, moments), _dev(torch_gpu, jac), _dev(torch_gpu, k))
    assert np.array_equal(got.cpu().numpy(), oracle.element_stiffness(moments, jac, k))
    got = hip.elementStiffness(_dev(torch_gpu, moments), _dev(torch_gpu, jac))
    assert np.array_equal(got.cpu().numpy(), oracle.element_stiffness(moments, jac))


def test_elements_of_more_than_256_nodes(hip, torch_gpu):
    """3D Q6 (343 nodes): local pairs a * 343 + b exceed 16 bits, so the slot lists store int32 locals, and one
    element matrix (117,649 entries) is more than one pass of the stiffness kernel's block."""
    dim, order, n = 3, 6, 2
    rng = np.random.default_rng(36)
    moments = systems.gradient_moments(dim, order)
    jac = rng.standard_normal((20, dim, dim)) * 0.2 + np.eye(dim) * rng.uniform(0.5, 2.0, (20, 1, 1))
    k = np.exp(rng.standard_normal(20))
    got = hip.elementStiffness(_dev(torch_gpu, moments), _dev(torch_gpu, jac), _dev(torch_gpu, k))
    assert np.array_equal(got.cpu().numpy(), oracle.element_stiffness(moments, jac, k))
    pattern = systems.lagrange(dim, order, n)
    dofs = systems.element_dofs(dim, order, n)
    matrices = rng.standard_normal((dofs.shape[0], dofs.shape[1], dofs.shape[1]))
    scale = rng.uniform(0.5, 2.0, dofs.shape[0])
    expected = systems.CSR(pattern.rows, pattern.rows, pattern.row_extents, pattern.column_indices,
                           oracle.assemble(pattern, dofs, matrices, scale))
    matrix = hip.CSRMatrix(pattern)
    assembly = hip.Assembly(matrix, dofs)
    assembly.assemble(_dev(torch_gpu, matrices), _dev(torch_gpu, scale))
    _entries_match(matrix, expected)
    vectors = rng.standard_normal((dofs.shape[0], dofs.shape[1]))
    got = assembly.assemble_vector(_dev(torch_gpu, vectors), _dev(torch_gpu, scale))
    assert np.array_equal(got.cpu().numpy(), oracle.assemble_vector(pattern.rows, dofs, vectors, scale))


@pytest.mark.parametrize("dim,order", [(2, 3), (3, 2)])
def test_stiffness_from_cell_geometry_assembles_the_generator_system(hip, torch_gpu, dim, order):
    """Uniform cells x = x0 + (h/2) xi with k_e: element stiffness on the device, then the per-element
    assembly, equals the generator's system (which sums tensor-product factors in another order)."""
    n = 6
    count = systems.lagrange_size(dim, order, n)[2]
    k = np.exp(systems.normal(3, 1.0, count))
    expected = systems.lagrange(dim, order, n, conductivity=k)
    moments = _dev(torch_gpu, systems.gradient_moments(dim, order))
    jac = _dev(torch_gpu, np.broadcast_to(np.eye(dim) * (0.5 / n), (count, dim, dim)).copy())
    stiffness = hip.elementStiffness(moments, jac, _dev(torch_gpu, k))
    matrix = hip.CSRMatrix(systems.lagrange(dim, order, n))
    assembly = hip.Assembly(matrix, systems.element_dofs(dim, order, n))
    assembly.assemble(stiffness)
    got = matrix.copy_out()
    assert np.allclose(got.entries, expected.entries, rtol=1e-12, atol=1e-12 * np.abs(expected.entries).max())


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
    _entries_match(matrix, expected)
    # element stiffness of the 262,144 cells from their geometry, then the per-element assembly
    moments = _dev(torch_gpu, systems.gradient_moments(dim, order))
    jac = _dev(torch_gpu, np.broadcast_to(np.eye(dim) * (0.5 / n), (count, dim, dim)).copy())
    stiffness = hip.elementStiffness(moments, jac, k_dev)
    torch_gpu.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        hip.elementStiffness(moments, jac, k_dev, out=stiffness)
    stiffness_seconds = (time.perf_counter() - t0) / 3
    t0 = time.perf_counter()
    for _ in range(3):
        assembly.assemble(stiffness)
    per_element_seconds = (time.perf_counter() - t0) / 3
    got = matrix.copy_out()
    assert np.allclose(got.entries, expected.entries, rtol=1e-12, atol=1e-12 * np.abs(expected.entries).max())
    print(f"\nassembly of {expected.nnz} nonzeros: device {device_seconds * 1e3:.2f} ms (shared K_e0, FP64 + FP32 "
          f"planes), host generator {host_seconds * 1e3:.0f} ms (16 threads); element stiffness of {count} cells "
          f"{stiffness_seconds * 1e3:.2f} ms, per-element assembly {per_element_seconds * 1e3:.2f} ms")

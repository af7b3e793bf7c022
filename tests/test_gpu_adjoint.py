"""Adjoint gradient dJ/dg_e = -k_e lambda_e^T K_e0 u_e against the oracle and
central finite differences (h = 1e-6)."""
import numpy as np
import pytest

import oracle
from cie_amd import adjoint, systems

pytestmark = pytest.mark.gpu


def test_element_gradient_kernel_bitwise(torch_gpu, hip):
    dim, order, n = 3, 2, 6
    problem = adjoint.LagrangeProblem(dim, order, n, np.zeros(systems.lagrange_size(dim, order, n)[0]),
                                      np.zeros(systems.lagrange_size(dim, order, n)[0]))
    rng = np.random.default_rng(4)
    lam = rng.standard_normal(problem.rows)
    u = rng.standard_normal(problem.rows)
    expected = oracle.element_gradient(problem.dofs_host, problem.element_matrix_host, lam, u)
    got = problem.element_gradient(torch_gpu.from_numpy(lam).cuda(), torch_gpu.from_numpy(u).cuda()).cpu().numpy()
    assert np.array_equal(got, expected)
    for dim, order, n in [(2, 4, 5), (3, 4, 2)]:  # element matrices of 25 and 125 nodes (LDS / global paths)
        rows = systems.lagrange_size(dim, order, n)[0]
        p = adjoint.LagrangeProblem(dim, order, n, np.zeros(rows), np.zeros(rows))
        lam, u = rng.standard_normal(rows), rng.standard_normal(rows)
        expected = oracle.element_gradient(p.dofs_host, p.element_matrix_host, lam, u)
        got = p.element_gradient(torch_gpu.from_numpy(lam).cuda(), torch_gpu.from_numpy(u).cuda()).cpu().numpy()
        assert np.array_equal(got, expected)


def cpu_reference_gradient(dim, order, n, g, g_target, rhs, tolerance=1e-13):
    k, kt = np.exp(g), np.exp(g_target)
    a = systems.lagrange(dim, order, n, conductivity=k)
    at = systems.lagrange(dim, order, n, conductivity=kt)
    u, _, _, _ = oracle.pcg(a, rhs, relative_tolerance=tolerance, max_iterations=100000)
    ut, _, _, _ = oracle.pcg(at, rhs, relative_tolerance=tolerance, max_iterations=100000)
    lam, _, _, _ = oracle.pcg(a, u - ut, relative_tolerance=tolerance, max_iterations=100000)
    grad = k * oracle.element_gradient(systems.element_dofs(dim, order, n), systems.element_matrix(dim, order, n),
                                       lam, u)
    return grad, 0.5 * np.dot(u - ut, u - ut), ut


@pytest.mark.parametrize("blocks", [False, True])
def test_autograd_matches_cpu_reference(torch_gpu, hip, blocks):
    """With blocks the solves walk 1 x 3 column blocks, whose values follow every re-assembly."""
    dim, order, n = 3, 2, 10
    _, _, count, _ = systems.lagrange_size(dim, order, n)
    g = systems.normal(4, np.sqrt(0.5), count)
    gt = systems.normal(5, np.sqrt(0.5), count)
    rhs = systems.uniform(1, systems.lagrange_size(dim, order, n)[0])
    grad_ref, j_ref, ut = cpu_reference_gradient(dim, order, n, g, gt, rhs)
    problem = adjoint.LagrangeProblem(dim, order, n, rhs, ut, tolerance=1e-13)
    if blocks:
        problem.assemble(gt)  # creates the matrix; the objective re-assembles it for g
        problem.matrix.set_launch(64, "blocks", 3, 8)
    g_t = torch_gpu.tensor(g, dtype=torch_gpu.float64, requires_grad=True)
    j = adjoint.Objective.apply(g_t, problem)
    j.backward()
    assert abs(j.item() - j_ref) <= 1e-9 * j_ref
    grad = g_t.grad.cpu().numpy()
    assert np.linalg.norm(grad - grad_ref) <= 1e-8 * np.linalg.norm(grad_ref)


def test_mixed_solves_give_the_fp64_gradient(torch_gpu, hip):
    """Forward and adjoint solves by the mixed-refined PCG stop on the FP64 true residual, so the gradient agrees
    with the FP64-solve gradient to the error both leave at that tolerance, and with the CPU reference."""
    dim, order, n = 3, 2, 10
    _, _, count, _ = systems.lagrange_size(dim, order, n)
    g = systems.normal(4, np.sqrt(0.5), count)
    gt = systems.normal(5, np.sqrt(0.5), count)
    rhs = systems.uniform(1, systems.lagrange_size(dim, order, n)[0])
    grad_ref, j_ref, ut = cpu_reference_gradient(dim, order, n, g, gt, rhs)
    grads = {}
    for mixed in (False, True):
        problem = adjoint.LagrangeProblem(dim, order, n, rhs, ut, tolerance=1e-13, mixed=mixed)
        g_t = torch_gpu.tensor(g, dtype=torch_gpu.float64, requires_grad=True)
        j = adjoint.Objective.apply(g_t, problem)
        j.backward()
        assert abs(j.item() - j_ref) <= 1e-9 * j_ref
        grads[mixed] = g_t.grad.cpu().numpy()
    assert np.linalg.norm(grads[True] - grads[False]) <= 1e-8 * np.linalg.norm(grads[False])
    assert np.linalg.norm(grads[True] - grad_ref) <= 1e-8 * np.linalg.norm(grad_ref)


def test_full_size_gradient_against_finite_differences(torch_gpu, hip):
    """3D Q2, 48^3 hexahedra (857,375 DOFs, 110,592 elements): forward and adjoint FP64 PCG to 1e-11,
    gradient over all elements, central FD with h = 1e-6 on 64 random elements."""
    w = systems.q2_adjoint_3d(48)
    g, gt = w.extra["log_conductivity"], w.extra["target_log_conductivity"]
    problem = adjoint.LagrangeProblem(3, 2, 48, w.rhs, np.zeros(w.matrix.rows), tolerance=1e-11)
    problem.assemble(gt)
    ut, info_t = problem.solve(problem.rhs)
    problem.target = ut
    g_t = torch_gpu.tensor(g, dtype=torch_gpu.float64, requires_grad=True)
    j = adjoint.Objective.apply(g_t, problem)
    j.backward()
    grad = g_t.grad.cpu().numpy()
    assert grad.shape == (110592,) and np.all(np.isfinite(grad))
    u, _ = problem.solve(problem.rhs)
    chosen = np.random.default_rng(64).choice(110592, size=64, replace=False)
    fd = adjoint.finite_difference(problem, g, u, chosen, h=1e-6)
    rel = np.abs(fd - grad[chosen]) / np.abs(grad[chosen])
    print(f"FD check: max rel {rel.max():.3e}, median {np.median(rel):.3e}; |grad| range "
          f"{np.abs(grad[chosen]).min():.3e}..{np.abs(grad[chosen]).max():.3e}")
    assert np.all(rel <= 1e-6), rel

"""Adjoint gradients of PCG solutions with respect to per-element conductivities.

For the Laplace stiffness K(k) = sum_e k_e P_e^T K_e0 P_e with k = exp(g)
(element integrand linear in the cell modulus,
fem/packages/integrands/impl/LinearIsotropicStiffnessIntegrand_impl.hpp:74)
and J(g) = 0.5 ||u(g) - u_target||^2 with K(k) u = f:

    K lambda = u - u_target            (adjoint; K is symmetric)
    dJ/dg_e  = -k_e lambda_e^T K_e0 u_e

K(exp(g)) is assembled on the device (cie_hip_assemble), both solves run the
fused FP64 PCG through the C ABI, and the element gather runs in
cie_hip_element_gradient. ``Objective`` wraps this as a
torch.autograd.Function; ``finite_difference`` is the central-difference check
used by the tests.
"""
import ctypes
from ctypes import c_void_p

import numpy as np

from . import linalg, systems
from ._lib import Statistics, check, library


def _torch():
    import torch

    return torch


def _dptr(tensor):
    return c_void_p(tensor.data_ptr())


class LagrangeProblem:
    """A structured Lagrange Laplace problem K(exp(g)) u = f, device resident."""

    def __init__(self, dim, order, elements, rhs, target, tolerance=1e-11, context=None, threads=0, mixed=False):
        """mixed: the forward and adjoint solves by the mixed-refined PCG (FP32 inner iterations, stopped on the
        FP64 true residual at the same tolerance) instead of the FP64 PCG."""
        torch = _torch()
        self.dim, self.order, self.elements = dim, order, elements
        self.context = context or linalg.Context.default()
        self.threads = threads
        rows, _, count, npe = systems.lagrange_size(dim, order, elements)
        self.rows, self.count, self.nodes = rows, count, npe
        self.device = torch.device("cuda", self.context.device)
        self.dofs_host = systems.element_dofs(dim, order, elements)
        self.element_matrix_host = systems.element_matrix(dim, order, elements)
        self.dofs = torch.from_numpy(self.dofs_host).to(self.device)
        self.element_matrix = torch.from_numpy(self.element_matrix_host).to(self.device)
        self.rhs = torch.as_tensor(np.asarray(rhs, dtype=np.float64), device=self.device)
        self.target = torch.as_tensor(target, dtype=torch.float64, device=self.device)
        self.tolerance = tolerance
        self.mixed = mixed
        self.matrix = None
        self.host_matrix = None
        self.assembly = None

    def assemble(self, g):
        """K(exp(g)) on the device: the pattern once from the structured generator, the values by
        the device assembly (bitwise the generator's values for the same k)."""
        torch = _torch()
        if torch.is_tensor(g):
            g = g.detach().cpu().numpy()
        k = np.exp(np.asarray(g, dtype=np.float64))
        if k.shape != (self.count,):
            raise ValueError(f"expected {self.count} element values, got {k.shape}")
        if self.matrix is None:
            self.host_matrix = systems.lagrange(self.dim, self.order, self.elements, threads=self.threads)
            self.matrix = linalg.CSRMatrix(self.host_matrix, self.context)
            self.assembly = linalg.Assembly(self.matrix, self.dofs_host)
        self.assembly.assemble(self.element_matrix, torch.from_numpy(k).to(self.device))
        return self.matrix

    def solve(self, b, x0=None, tolerance=None, absolute=0.0, mixed=None):
        """K x = b from x0 (zero), by the FP64 PCG or, with `mixed` (default: the problem's), the mixed one."""
        torch = _torch()
        x = torch.zeros_like(b) if x0 is None else x0.clone()
        solver_class = linalg.MixedPCG if (self.mixed if mixed is None else mixed) else linalg.FusedPCG
        solver = solver_class(self.matrix, Statistics(100000, absolute, tolerance or self.tolerance))
        solver.product(0, b, 1, x)
        return x, solver.info

    def element_gradient(self, lam, u, scale=None):
        """gradient[e] = scale[e] * (-lambda_e^T K_e0 u_e) (scale None: 1)."""
        torch = _torch()
        out = torch.empty(self.count, dtype=torch.float64, device=self.device)
        torch.cuda.synchronize()
        check(library().cie_hip_element_gradient(self.context.handle, self.count, self.nodes, _dptr(self.dofs),
                                                 _dptr(self.element_matrix), None if scale is None else _dptr(scale),
                                                 _dptr(lam), _dptr(u), _dptr(out)), "cie_hip_element_gradient")
        return out


class Objective:
    """J(g) = 0.5 ||u(g) - u_target||^2 as a torch.autograd.Function (adjoint backward)."""

    _function = None

    @classmethod
    def function(cls):
        if cls._function is None:
            torch = _torch()

            class _Objective(torch.autograd.Function):
                @staticmethod
                def forward(ctx, g, problem):
                    problem.assemble(g.detach().cpu().numpy())
                    u, info = problem.solve(problem.rhs)
                    ctx.problem = problem
                    ctx.save_for_backward(g, u)
                    ctx.forward_info = info
                    residual = u - problem.target
                    return 0.5 * torch.dot(residual, residual)

                @staticmethod
                def backward(ctx, grad_output):
                    g, u = ctx.saved_tensors
                    problem = ctx.problem
                    lam, _ = problem.solve(u - problem.target)
                    k = torch.exp(g.detach().to(problem.device, torch.float64))
                    grad = problem.element_gradient(lam, u, scale=k)
                    return (grad_output * grad).to(g.device, g.dtype), None

            cls._function = _Objective
        return cls._function

    @classmethod
    def apply(cls, g, problem):
        return cls.function().apply(g, problem)


def finite_difference(problem, g, u, elements, h=1e-6, tolerance=1e-13):
    """Central differences (J(g + h e_i) - J(g - h e_i)) / 2h for the listed elements.

    K(g +- h e_i) differs from K(g) by c+- K_e0 on element i's slots,
    c+- = k_i (exp(+-h) - 1). The perturbed solutions are written u+- = u + d+-
    with K+- d+- = r0 - c+- P_i^T K_e0 u_i, r0 = f - K u (shared bitwise by
    both), and J(+) - J(-) = sum (d+ - d-) (u - u_t + (d+ + d-) / 2), which is
    the same difference without forming J(+-) or u+- in floating point.
    """
    torch = _torch()
    matrix = problem.matrix
    device = problem.device
    k = np.exp(np.asarray(g, dtype=np.float64))
    r0 = problem.rhs.clone()
    # The clone runs on torch's stream and the product on the library's: wait for the clone (an unordered
    # read of r0 would start every correction system from whatever that memory held before).
    torch.cuda.synchronize()
    check(library().cie_hip_csr_product_f64(problem.context.handle, matrix.handle, 1.0, _dptr(u), problem.rows, -1.0,
                                            _dptr(r0), problem.rows), "r0")
    problem.context.synchronize()
    u_host = u.cpu().numpy()
    csr = problem.host_matrix  # the pattern
    results = []
    for element in elements:
        d = problem.dofs_host[element]
        keep = d >= 0
        pairs_r = np.repeat(d, problem.nodes)
        pairs_c = np.tile(d, problem.nodes)
        slots = systems.csr_slots(csr, np.where(pairs_r >= 0, pairs_r, -1), np.where(pairs_c >= 0, pairs_c, -1))
        slots[(pairs_r < 0) | (pairs_c < 0)] = -1
        slots_dev = torch.from_numpy(slots).to(device)
        ke = problem.element_matrix
        local = problem.element_matrix_host @ np.where(keep, u_host[np.maximum(d, 0)], 0.0)
        deltas = {}
        for sign in (+1, -1):
            c = k[element] * (np.exp(sign * h) - 1.0)
            factor = torch.tensor([c], dtype=torch.float64, device=device)
            torch.cuda.synchronize()
            check(library().cie_hip_csr_add_element_matrices(matrix.handle, 1, problem.nodes, _dptr(slots_dev),
                                                             _dptr(ke), _dptr(factor)), "perturb")
            rhs = r0.clone()
            rows = torch.from_numpy(d[keep].astype(np.int64)).to(device)
            rhs[rows] -= torch.from_numpy(c * local[keep]).to(device)
            delta, _ = problem.solve(rhs, tolerance=tolerance, mixed=False)
            deltas[sign] = delta
            problem.assemble(g)  # restore K(g) exactly (the assembly is deterministic)
        plus, minus = deltas[+1], deltas[-1]
        difference = torch.dot(plus - minus, u - problem.target + 0.5 * (plus + minus)).item()
        results.append(difference / (2.0 * h))
    return np.array(results)

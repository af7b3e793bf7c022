"""Contexts (cie_hip_context: one device, one stream, its own reduction scratch and solver state).

Solves on different contexts of one device share no state: the fused FP64 and the mixed solver and the
SpMV give bitwise the same results on the default context, on a second context with a stream of its own and
on a context that borrows a torch stream, also when the calls alternate between the contexts, and the FP64
solution agrees with the oracle's Jacobi-PCG (ConjugateGradients_impl.hpp:49-201). A matrix belongs to the
context it was created on: a call that pairs it with another context is refused. Closing a context closes
the objects still open on it first, so the garbage collector may finalize them in any order.
"""
import gc
from ctypes import c_void_p

import numpy as np
import pytest

import oracle
from cie_amd import systems
from cie_amd import _lib
from cie_amd._lib import InvalidArgument, Statistics

pytestmark = pytest.mark.gpu


def _dev(torch, array):
    return torch.from_numpy(np.ascontiguousarray(array, dtype=np.float64)).to("cuda:0")


def test_solves_on_separate_contexts_agree_bitwise(hip, torch_gpu):
    w = systems.q2_laplace_3d(8)
    x_ref, s_ref, _, _ = oracle.pcg(w.matrix, w.rhs, relative_tolerance=1e-10)
    borrowed = torch_gpu.cuda.Stream()
    contexts = [hip.Context.default(), hip.Context(0), hip.Context(0, c_void_p(borrowed.cuda_stream))]
    matrices = [hip.CSRMatrix(w.matrix, context=c) for c in contexts]
    rng = np.random.default_rng(4)
    x_in = rng.standard_normal(w.matrix.columns)
    b = _dev(torch_gpu, w.rhs)

    def solve(kind, matrix):
        solver = (hip.FusedPCG if kind == "fp64" else hip.MixedPCG)(matrix, Statistics(10000, 0.0, 1e-10))
        x = torch_gpu.zeros_like(b)
        solver.product(0, b, 1, x)
        assert solver.info.converged == 1
        return x.cpu().numpy(), int(solver.getStats().iterationCount)

    def spmv(matrix):
        space = hip.HIPSpace(context=matrix.context)
        y = torch_gpu.zeros(w.matrix.rows, dtype=torch_gpu.float64, device="cuda:0")
        hip.HIPCSROperator(matrix, space).product(0, _dev(torch_gpu, x_in), 1, y)
        return y.cpu().numpy()

    first = {kind: solve(kind, matrices[0]) for kind in ("fp64", "mixed")}
    product = spmv(matrices[0])
    assert abs(first["fp64"][1] - int(s_ref.iterationCount)) <= 1
    assert np.linalg.norm(first["fp64"][0] - x_ref) / np.linalg.norm(x_ref) < 1e-10
    # alternate between the contexts, twice round
    for _ in range(2):
        for matrix in matrices[1:] + matrices[:1]:
            for kind in ("fp64", "mixed"):
                x, iterations = solve(kind, matrix)
                assert iterations == first[kind][1] and np.array_equal(x, first[kind][0]), kind
            assert np.array_equal(spmv(matrix), product)
    del matrices
    gc.collect()
    torch_gpu.cuda.synchronize()
    for context in contexts[1:]:
        context.close()


def test_a_matrix_is_used_through_its_own_context(hip, torch_gpu):
    w = systems.five_point_2d(32)
    own, other = hip.Context.default(), hip.Context(0)
    matrix = hip.CSRMatrix(w.matrix, context=own)
    lib = _lib.library()
    n = w.matrix.rows
    x = torch_gpu.ones(n, dtype=torch_gpu.float64, device="cuda:0")
    y = torch_gpu.zeros_like(x)
    b = _dev(torch_gpu, w.rhs)
    torch_gpu.cuda.synchronize()
    calls = {
        "product": lambda c: lib.cie_hip_csr_product_f64(c.handle, matrix.handle, 0.0, c_void_p(x.data_ptr()), n, 1.0,
                                                         c_void_p(y.data_ptr()), n),
        "inverse diagonal": lambda c: lib.cie_hip_inverse_diagonal_f64(c.handle, matrix.handle,
                                                                       c_void_p(y.data_ptr()), None),
        "pcg": lambda c: lib.cie_hip_pcg_f64(c.handle, matrix.handle, 0.0, c_void_p(b.data_ptr()), 1.0,
                                             c_void_p(y.data_ptr()), n, Statistics(100, 0.0, 1e-8), 0, None),
    }
    for name, call in calls.items():
        with pytest.raises(InvalidArgument):
            _lib.check(call(other), name)
        _lib.check(call(own), name)  # the matrix's own context
    with pytest.raises(InvalidArgument):
        hip.HIPCSROperator(matrix, hip.HIPSpace(context=other)).product(0, x, 1, y)
    del matrix
    gc.collect()
    other.close()


def test_closing_a_context_closes_its_objects_first(hip, torch_gpu):
    w = systems.lagrange(2, 2, 6, basis="legendre")
    context = hip.Context(0)
    matrix = hip.CSRMatrix(w, context=context)
    mask = hip.Mask(systems.ansatz_mask(2, 2, 6), context=context)
    masked = hip.maskedMatrix(matrix, mask, 2)
    batch = hip.Batch(matrix, np.array([0, w.rows], dtype=np.int64))
    objects = [matrix, mask, masked, batch]
    context.close()
    assert not context.handle
    assert all(not obj.handle for obj in objects)
    for obj in objects:
        obj.close()  # a second close does nothing
    # a context and its matrix in a reference cycle: the collector finalizes them in its own order
    cycle_context = hip.Context(0)
    cycle_matrix = hip.CSRMatrix(w, context=cycle_context)
    cycle_context.keep = cycle_matrix
    del cycle_context, cycle_matrix
    gc.collect()
    # the default context still works
    y = torch_gpu.zeros(w.rows, dtype=torch_gpu.float64, device="cuda:0")
    hip.HIPCSROperator(hip.CSRMatrix(w), hip.HIPSpace()).product(0, _dev(torch_gpu, np.ones(w.columns)), 1, y)
    assert np.allclose(y.cpu().numpy(), oracle.csr_product(w, 0.0, np.ones(w.columns), 1.0, np.zeros(w.rows)))


def test_objects_may_be_destroyed_in_any_order(hip, torch_gpu):
    """Through the C ABI: a context destroyed before its matrix, and a matrix and a mask destroyed before the
    batch, assembly and multigrid plans that hold them, are released with their last holder."""
    lib = _lib.library()
    w = systems.lagrange(2, 2, 6, basis="legendre")
    mask_values = systems.ansatz_mask(2, 2, 6)
    for order in ("context, matrix, mask, plans", "matrix, mask, context, plans", "plans, matrix, mask, context"):
        context = hip.Context(0)
        matrix = hip.CSRMatrix(w, context=context)
        mask = hip.Mask(mask_values, context=context)
        plans = [hip.Batch(matrix, np.array([0, w.rows], dtype=np.int64)), hip.PMultigrid(matrix, mask, 2)]
        handles = {"context": context.handle, "matrix": matrix.handle, "mask": mask.handle}
        destroy = {"context": lib.cie_hip_context_destroy, "matrix": lib.cie_hip_csr_destroy,
                   "mask": lib.cie_hip_mask_destroy}
        for name in order.split(", "):
            if name == "plans":
                for plan in plans:
                    plan.close()
            else:
                assert destroy[name](handles[name]) == 0, (order, name)
        # the Python objects no longer own their handles
        for obj in (matrix, mask, context):
            obj.handle = c_void_p()
        context._dependents = []
        torch_gpu.cuda.synchronize()

"""The no-allocation dot product (cie_hip_inner_product_f64_static) against cie_hip_inner_product_f64.

Same kernels and summation order, partials in the library's static device array instead of the
context's scratch: the results must be bitwise equal, on two calls in a row (the second smaller,
so it writes fewer partials into the array the first one filled), through the C ABI.
"""
import ctypes

import numpy as np
import pytest

from cie_amd import _lib, linalg

pytestmark = pytest.mark.gpu


def dot_reference(lib, context, left, right):
    value = ctypes.c_double()
    _lib.check(lib.cie_hip_inner_product_f64(context.handle, ctypes.c_void_p(left.data_ptr()),
                                             ctypes.c_void_p(right.data_ptr()), left.numel(), ctypes.byref(value)),
               "cie_hip_inner_product_f64")
    return value.value


def dot_static(lib, context, torch, left, right, size=None):
    result = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    size = left.numel() if size is None else size
    _lib.check(lib.cie_hip_inner_product_f64_static(context.handle, ctypes.c_void_p(left.data_ptr()),
                                                    ctypes.c_void_p(right.data_ptr()), size,
                                                    ctypes.c_void_p(result.data_ptr())),
               "cie_hip_inner_product_f64_static")
    context.synchronize()  # the entry point itself never waits
    return result.item()


def test_static_dot_matches_scratch_dot_twice(torch_gpu, hip):
    lib = _lib.library()
    context = linalg.Context.default()
    rng = np.random.default_rng(17)
    for n in (1_000_003, 12_345):  # the second call writes fewer partials over the first call's
        a = torch_gpu.from_numpy(rng.standard_normal(n)).to("cuda:0")
        b = torch_gpu.from_numpy(rng.standard_normal(n)).to("cuda:0")
        expected = dot_reference(lib, context, a, b)
        got = dot_static(lib, context, torch_gpu, a, b)
        assert got == expected, (n, got, expected)
        assert abs(got - float(a.cpu().numpy() @ b.cpu().numpy())) <= 1e-12 * n
    # the empty sum is +0, and invalid arguments are refused before anything runs
    assert dot_static(lib, context, torch_gpu, a[:1], b[:1], size=0) == 0.0
    status = lib.cie_hip_inner_product_f64_static(context.handle, ctypes.c_void_p(a.data_ptr()),
                                                  ctypes.c_void_p(b.data_ptr()), -1, ctypes.c_void_p(a.data_ptr()))
    assert status == _lib.ERROR_INVALID_ARGUMENT

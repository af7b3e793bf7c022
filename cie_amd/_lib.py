"""ctypes binding of libcie_hip.so (the C ABI declared in include/cie_hip.h).

The shared library is built in-tree by ``cie_amd._build`` and loaded from
``cie_amd/libcie_hip.so``. There is no fallback: if the library is missing,
importing a compute entry point raises.
"""
import ctypes
from ctypes import POINTER, byref, c_char_p, c_double, c_float, c_int, c_int32, c_int64, c_uint16, c_uint64, c_void_p
from pathlib import Path

import os

# CIE_HIP_LIBRARY selects another build of the same C ABI (e.g. one compiled
# with different load or launch options); by default the in-tree build.
LIBRARY_PATH = Path(os.environ.get("CIE_HIP_LIBRARY") or Path(__file__).resolve().parent / "libcie_hip.so")

OK = 0
ERROR_INVALID_ARGUMENT = 1
ERROR_SHAPE = 2
ERROR_NOT_SPD = 3
ERROR_BREAKDOWN = 4
ERROR_ZERO_DIAGONAL = 5
ERROR_MISSING_DIAGONAL = 6
ERROR_HIP = 7
ERROR_OUT_OF_MEMORY = 8
ERROR_NO_DEVICE = 9
ERROR_NOT_IMPLEMENTED = 10

HOST = 0
DEVICE = 1
SPMV_VECTOR = 0
SPMV_STAGED = 1
SPMV_TILED = 2
SPMV_BLOCKS = 3
SPMV_SORTED = 4
SPMV_PAIRED = 5
REORDER_NONE, REORDER_CUTHILL_MCKEE, REORDER_REVERSE_CUTHILL_MCKEE = 0, 1, 2
PLANE_F64 = 1
PLANE_F32 = 2


class Statistics(ctypes.Structure):
    """IterativeSolver<TSpace>::Statistics (linalg/packages/solvers/inc/IterativeSolver.hpp:21-25)."""

    _fields_ = [("iterationCount", c_uint64), ("absoluteResidual", c_double), ("relativeResidual", c_double)]

    def __repr__(self):
        return (f"Statistics(iterationCount={self.iterationCount}, absoluteResidual={self.absoluteResidual:.6e}, "
                f"relativeResidual={self.relativeResidual:.6e})")


class SolveInfo(ctypes.Structure):
    _fields_ = [
        ("stats", Statistics),
        ("converged", c_int32),
        ("status", c_int32),
        ("trueRelativeResidual", c_double),
        ("seconds", c_double),
        ("refinements", c_uint64),
        ("spmvCount", c_uint64),
        ("productSeconds", c_double),
        ("productSamples", c_uint64),
        ("initialResidual", c_double),
        ("reportedIterations", c_uint64),
    ]


class MixedSettings(ctypes.Structure):
    _fields_ = [("settings", Statistics), ("refinementFactor", c_double), ("restart", c_int32), ("flags", c_int32)]


REPORT_TERMINATION = 0  # CIE_HIP_REPORT_TERMINATION (IterativeSolver::ReportType::Termination)
REPORT_ITERATION = 1    # CIE_HIP_REPORT_ITERATION (IterativeSolver::ReportType::Iteration)
MIXED_PLAIN_FP32 = 1  # CIE_HIP_MIXED_PLAIN_FP32: no row-sum correction in the inner FP32 product
DEFAULT_REFINEMENT_FACTOR = 0.03  # CIE_HIP_DEFAULT_REFINEMENT_FACTOR


class PmgInfo(ctypes.Structure):
    _fields_ = [("cycles", c_uint64), ("coarseIterations", c_uint64), ("initialResidual", c_double),
                ("residual", c_double), ("converged", c_int32), ("status", c_int32), ("seconds", c_double)]


class LagrangeSpec(ctypes.Structure):
    _fields_ = [("dim", c_int32), ("order", c_int32), ("elements", c_int32), ("eliminate", c_int32),
                ("basis", c_int32)]


class CieHipError(RuntimeError):
    """A failed C-ABI call; the counterpart of cie::Exception thrown by CIE_CHECK."""

    def __init__(self, status, message):
        super().__init__(f"[{status}] {message}")
        self.status = status


class NotPositiveDefinite(CieHipError):
    pass


class Breakdown(CieHipError):
    pass


class ZeroDiagonal(CieHipError):
    pass


class MissingDiagonal(CieHipError):
    pass


class NotImplementedForArguments(CieHipError):
    """NotImplementedException (MaskedJacobiOperator.cpp:64-65, ApproximateSubstitution.cpp:40-41)."""


class ShapeError(CieHipError):
    pass


class InvalidArgument(CieHipError):
    pass


_ERRORS = {
    ERROR_INVALID_ARGUMENT: InvalidArgument,
    ERROR_SHAPE: ShapeError,
    ERROR_NOT_SPD: NotPositiveDefinite,
    ERROR_BREAKDOWN: Breakdown,
    ERROR_ZERO_DIAGONAL: ZeroDiagonal,
    ERROR_MISSING_DIAGONAL: MissingDiagonal,
    ERROR_NOT_IMPLEMENTED: NotImplementedForArguments,
}

_P = c_void_p
_SIGNATURES = {
    # errors, devices, contexts
    "cie_hip_last_error": (c_char_p, []),
    "cie_hip_status_string": (c_char_p, [c_int]),
    "cie_hip_device_count": (c_int, [POINTER(c_int)]),
    "cie_hip_context_create": (c_int, [c_int, _P, POINTER(_P)]),
    "cie_hip_context_destroy": (c_int, [_P]),
    "cie_hip_context_synchronize": (c_int, [_P]),
    "cie_hip_context_stream": (c_int, [_P, POINTER(_P)]),
    "cie_hip_context_set_profiling": (c_int, [_P, c_int]),
    "cie_hip_context_kernel_seconds": (c_int, [_P, POINTER(c_double)]),
    "cie_hip_context_set_residual_history": (c_int, [_P, _P, c_int64]),
    "cie_hip_iteration_report": (c_int, [c_int, c_int, POINTER(Statistics), POINTER(Statistics), ctypes.c_char_p,
                                         ctypes.c_size_t, POINTER(ctypes.c_size_t)]),
    "cie_hip_malloc": (c_int, [_P, ctypes.c_size_t, POINTER(_P)]),
    "cie_hip_free": (c_int, [_P, _P]),
    "cie_hip_memcpy": (c_int, [_P, _P, _P, ctypes.c_size_t, c_int]),
    # CSR
    "cie_hip_csr_create": (c_int, [_P, c_int64, c_int64, c_int64, _P, _P, _P, c_int, c_int, POINTER(_P)]),
    "cie_hip_csr_destroy": (c_int, [_P]),
    "cie_hip_csr_shape": (c_int, [_P, POINTER(c_int64), POINTER(c_int64), POINTER(c_int64)]),
    "cie_hip_csr_update_entries": (c_int, [_P, _P]),
    "cie_hip_csr_column_bytes": (c_int, [_P, POINTER(c_int)]),
    "cie_hip_csr_column_blocks": (c_int, [_P, c_int, POINTER(c_int64), POINTER(c_int)]),
    "cie_hip_csr_paired_blocks": (c_int, [_P, c_int, POINTER(c_int64), POINTER(c_int64)]),
    "cie_hip_csr_set_launch": (c_int, [_P, c_int, c_int, c_int, c_int]),
    "cie_hip_csr_get_launch": (c_int, [_P, c_int, POINTER(c_int), POINTER(c_int), POINTER(c_int)]),
    "cie_hip_csr_autotune": (c_int, [_P, c_int, c_int, c_int, POINTER(c_double)]),
    "cie_hip_csr_product_f64": (c_int, [_P, _P, c_double, _P, c_int64, c_double, _P, c_int64]),
    "cie_hip_csr_product_f32": (c_int, [_P, _P, c_float, _P, c_int64, c_float, _P, c_int64]),
    # spaces
    "cie_hip_inner_product_f64": (c_int, [_P, _P, _P, c_int64, POINTER(c_double)]),
    "cie_hip_inner_product_f64_static": (c_int, [_P, _P, _P, c_int64, _P]),
    "cie_hip_scale_f64": (c_int, [_P, _P, _P, c_int64, c_double]),
    "cie_hip_scale_uniform_f64": (c_int, [_P, _P, c_int64, c_double]),
    "cie_hip_add_f64": (c_int, [_P, _P, _P, c_int64, c_double]),
    "cie_hip_assign_f64": (c_int, [_P, _P, _P, c_int64]),
    "cie_hip_fill_f64": (c_int, [_P, _P, c_int64, c_double]),
    "cie_hip_inner_product_f32": (c_int, [_P, _P, _P, c_int64, POINTER(c_float)]),
    "cie_hip_scale_f32": (c_int, [_P, _P, _P, c_int64, c_float]),
    "cie_hip_scale_uniform_f32": (c_int, [_P, _P, c_int64, c_float]),
    "cie_hip_add_f32": (c_int, [_P, _P, _P, c_int64, c_float]),
    "cie_hip_assign_f32": (c_int, [_P, _P, _P, c_int64]),
    "cie_hip_fill_f32": (c_int, [_P, _P, c_int64, c_float]),
    # Jacobi
    "cie_hip_inverse_diagonal_f64": (c_int, [_P, _P, _P, POINTER(c_int64)]),
    "cie_hip_inverse_diagonal_f32": (c_int, [_P, _P, _P, POINTER(c_int64)]),
    "cie_hip_csr_row_sum_correction": (c_int, [_P, _P, _P]),
    "cie_hip_diagonal_product_f64": (c_int, [_P, _P, c_double, _P, c_double, _P, c_int64]),
    "cie_hip_diagonal_product_f32": (c_int, [_P, _P, c_float, _P, c_float, _P, c_int64]),
    # solvers
    "cie_hip_pcg_f64": (c_int, [_P, _P, c_double, _P, c_double, _P, c_int64, POINTER(Statistics), c_int,
                                POINTER(SolveInfo)]),
    "cie_hip_pcg_mixed": (c_int, [_P, _P, c_double, _P, c_double, _P, c_int64, POINTER(MixedSettings),
                                  POINTER(SolveInfo)]),
    # batched
    "cie_hip_batch_create": (c_int, [_P, _P, c_int32, _P, POINTER(_P)]),
    "cie_hip_batch_destroy": (c_int, [_P]),
    "cie_hip_batch_set_tile_rows": (c_int, [_P, _P]),
    "cie_hip_batch_get_tile_rows": (c_int, [_P, _P]),
    "cie_hip_batch_autotune": (c_int, [_P, c_int, c_int]),
    "cie_hip_batch_set_column_blocks": (c_int, [_P, c_int]),
    "cie_hip_batch_get_column_blocks": (c_int, [_P, POINTER(c_int)]),
    "cie_hip_batch_set_tile_kernel": (c_int, [_P, c_int]),
    "cie_hip_batch_get_tile_kernel": (c_int, [_P, POINTER(c_int)]),
    "cie_hip_pcg_batched": (c_int, [_P, _P, _P, _P, POINTER(MixedSettings), c_int, c_int, _P, _P, _P,
                                    POINTER(SolveInfo)]),
    # adjoint
    "cie_hip_element_gradient": (c_int, [_P, c_int64, c_int32, _P, _P, _P, _P, _P, _P]),
    "cie_hip_csr_add_element_matrices": (c_int, [_P, c_int64, c_int32, _P, _P, _P]),
    # parameterised stiffness assembly
    "cie_hip_assembly_create": (c_int, [_P, c_int64, c_int32, _P, POINTER(_P)]),
    "cie_hip_assembly_destroy": (c_int, [_P]),
    "cie_hip_assemble": (c_int, [_P, _P, c_int, _P]),
    "cie_hip_assemble_vector": (c_int, [_P, _P, c_int, _P, _P]),
    "cie_hip_element_stiffness": (c_int, [_P, c_int64, c_int32, c_int32, _P, _P, _P, _P]),
    # reordering
    "cie_hip_make_reordering": (c_int, [c_int64, _P, _P, c_int, _P]),
    "cie_hip_csr_reorder": (c_int, [_P, _P, c_int, POINTER(_P)]),
    "cie_hip_reorder_f64": (c_int, [_P, _P, c_int, _P, _P, c_int64]),
    "cie_hip_reorder_f32": (c_int, [_P, _P, c_int, _P, _P, c_int64]),
    "cie_hip_csr_copy_out": (c_int, [_P, _P, _P, _P]),
    # masked operators, Jacobi smoothers, p-multigrid
    "cie_hip_mask_create": (c_int, [_P, _P, c_int64, POINTER(_P)]),
    "cie_hip_mask_destroy": (c_int, [_P]),
    "cie_hip_csr_masked": (c_int, [_P, _P, c_uint16, POINTER(_P)]),
    "cie_hip_masked_identity_f64": (c_int, [_P, _P, c_uint16, c_double, _P, c_double, _P, c_int64]),
    "cie_hip_jacobi_f64": (c_int, [_P, _P, _P, c_uint64, c_double, c_double, _P, c_double, _P, c_int64]),
    "cie_hip_masked_jacobi_f64": (c_int, [_P, _P, _P, c_uint16, c_uint64, c_double, c_double, _P, c_double, _P,
                                          c_int64]),
    "cie_hip_approximate_substitution_f64": (c_int, [_P, _P, c_uint64, c_double, _P, c_double, _P, c_int64]),
    "cie_hip_pmg_create": (c_int, [_P, _P, c_int, POINTER(_P)]),
    "cie_hip_pmg_destroy": (c_int, [_P]),
    "cie_hip_pmg_solve": (c_int, [_P, _P, _P, c_int64, c_double, c_uint64, POINTER(PmgInfo)]),
    "cie_hip_pmg_grid": (c_int, [_P, c_int, POINTER(_P)]),
    # MatrixMarket files
    "cie_hip_mm_write_csr_f64": (c_int, [c_char_p, c_int64, c_int64, c_int64, _P, _P, _P, c_int]),
    "cie_hip_mm_write_csr_f32": (c_int, [c_char_p, c_int64, c_int64, c_int64, _P, _P, _P, c_int]),
    "cie_hip_mm_write_array_f64": (c_int, [c_char_p, _P, c_int64, c_int]),
    "cie_hip_mm_write_array_f32": (c_int, [c_char_p, _P, c_int64, c_int]),
    "cie_hip_mm_read": (c_int, [c_char_p, POINTER(_P)]),
    "cie_hip_mm_shape": (c_int, [_P, POINTER(c_int64), POINTER(c_int64), POINTER(c_int64), POINTER(c_int)]),
    "cie_hip_mm_arrays": (c_int, [_P, POINTER(_P), POINTER(_P), POINTER(_P)]),
    "cie_hip_mm_free": (c_int, [_P]),
    # structured systems (include/cie_hip_systems.h)
    "cie_sys_last_error": (c_char_p, []),
    "cie_sys_splitmix64": (c_int, [c_uint64, c_int64, _P]),
    "cie_sys_uniform": (c_int, [c_uint64, c_int64, _P]),
    "cie_sys_normal": (c_int, [c_uint64, c_double, c_int64, _P]),
    "cie_sys_poisson5_size": (c_int, [c_int, POINTER(c_int64), POINTER(c_int64)]),
    "cie_sys_poisson5_fill": (c_int, [c_int, _P, _P, _P]),
    "cie_sys_lagrange_size": (c_int, [POINTER(LagrangeSpec), POINTER(c_int64), POINTER(c_int64), POINTER(c_int64),
                                      POINTER(c_int32)]),
    "cie_sys_lagrange_fill": (c_int, [POINTER(LagrangeSpec), _P, _P, _P, _P, c_int]),
    "cie_sys_lagrange_element_matrix": (c_int, [POINTER(LagrangeSpec), _P]),
    "cie_sys_lagrange_element_dofs": (c_int, [POINTER(LagrangeSpec), _P]),
    "cie_sys_csr_slots": (c_int, [c_int64, _P, _P, c_int64, _P, _P, _P]),
    "cie_sys_lagrange_ansatz_mask": (c_int, [POINTER(LagrangeSpec), _P]),
    "cie_sys_lagrange_gradient_moments": (c_int, [POINTER(LagrangeSpec), _P]),
}

_library = None


def library():
    """The loaded libcie_hip.so; raises if it has not been built."""
    global _library
    if _library is None:
        if not LIBRARY_PATH.exists():
            raise ImportError(f"{LIBRARY_PATH} is missing: build it with `python -m cie_amd._build` first")
        lib = ctypes.CDLL(str(LIBRARY_PATH))
        for name, (restype, argtypes) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _library = lib
    return _library


def check(status, what=""):
    if status != OK:
        message = library().cie_hip_last_error().decode(errors="replace")
        raise _ERRORS.get(status, CieHipError)(status, f"{what}: {message}" if what else message)
    return status


def check_sys(status, what=""):
    if status != 0:
        raise RuntimeError(f"{what}: {library().cie_sys_last_error().decode(errors='replace')}")


def exported_names():
    return list(_SIGNATURES)


__all__ = [
    "library", "check", "check_sys", "Statistics", "SolveInfo", "MixedSettings", "LagrangeSpec", "CieHipError",
    "NotPositiveDefinite", "Breakdown", "ZeroDiagonal", "MissingDiagonal", "ShapeError", "byref",
]

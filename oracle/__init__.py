"""CPU oracle of the sparse PCG path -- TEST INFRASTRUCTURE ONLY.

ctypes wrappers over oracle/liboracle.so (a plain-C restatement of the
reference's threaded CG path; see oracle.c for the file:line each function
follows). Only the tests, the smoke check and the benchmark's CPU-baseline leg
use this package, and only as the checker / baseline -- never as the product.
"""
import ctypes
from ctypes import POINTER, byref, c_double, c_int, c_int32, c_int64, c_uint64, c_void_p
from pathlib import Path

import numpy as np

LIBRARY = Path(__file__).resolve().parent / "liboracle.so"

OP_SCALE, OP_SCALE_UNIFORM, OP_ADD, OP_ASSIGN, OP_FILL = range(5)
STATUS = {0: "ok", 2: "shape", 3: "not spd", 4: "breakdown", 5: "zero diagonal", 6: "missing diagonal",
          10: "not implemented"}


class Statistics(ctypes.Structure):
    _fields_ = [("iterationCount", c_uint64), ("absoluteResidual", c_double), ("relativeResidual", c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIBRARY.exists():
            from .build import build

            build()
        l = ctypes.CDLL(str(LIBRARY))
        P = c_void_p
        l.oracle_csr_product.argtypes = [c_int64, P, P, P, c_double, P, c_double, P, c_int]
        l.oracle_inner_product.argtypes = [c_int64, P, P, c_int]
        l.oracle_inner_product.restype = c_double
        l.oracle_vector_op.argtypes = [c_int, P, P, c_double, c_int64, c_int]
        l.oracle_inverse_diagonal.argtypes = [c_int64, P, P, P, P, POINTER(c_int64)]
        l.oracle_diagonal_product.argtypes = [P, c_double, P, c_double, P, c_int64]
        l.oracle_pcg.argtypes = [c_int64, P, P, P, c_double, P, c_double, P, c_uint64, c_double, c_double, c_int,
                                 c_int, POINTER(Statistics), P, c_int64, POINTER(c_double)]
        l.oracle_element_gradient.argtypes = [c_int64, c_int32, P, P, P, P, P]
        l.oracle_assemble.argtypes = [c_int64, P, P, c_int64, c_int32, P, P, c_int64, P, P]
        l.oracle_assemble_vector.argtypes = [c_int64, c_int64, c_int32, P, P, c_int64, P, P]
        l.oracle_degree_sort.argtypes = [c_int64, P, P]
        l.oracle_element_stiffness.argtypes = [c_int64, c_int, c_int, P, P, P, P]
        l.oracle_heap_sort.argtypes = [c_int64, P, P]
        l.oracle_make_reordering.argtypes = [c_int64, P, P, c_int, P]
        l.oracle_reorder_csr.argtypes = [c_int64, P, c_int, P, P, P, P, P, P]
        l.oracle_reorder_vector.argtypes = [c_int64, P, c_int, P, P]
        l.oracle_masked_csr_product.argtypes = [c_int64, P, P, P, P, ctypes.c_uint16, c_double, P, c_double, P, c_int]
        l.oracle_masked_identity.argtypes = [c_int64, P, ctypes.c_uint16, c_double, P, c_double, P]
        l.oracle_jacobi.argtypes = [c_int64, P, P, P, P, ctypes.c_uint16, c_uint64, c_double, c_double, P, c_double,
                                    P, c_int]
        l.oracle_masked_jacobi.argtypes = [c_int64, P, P, P, P, ctypes.c_uint16, c_uint64, c_double, c_double, P,
                                           c_double, P]
        l.oracle_approximate_substitution.argtypes = [c_int64, P, P, P, c_uint64, c_double, P, c_double, P]
        l.oracle_pmg.argtypes = [c_int64, P, P, P, P, c_int, P, P, c_double, c_int64, P, P, c_int64,
                                 POINTER(c_int64), POINTER(c_uint64), c_int]
        _lib = l
    return _lib


def _p(a):
    return a.ctypes.data_as(c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class OracleError(RuntimeError):
    def __init__(self, status, message=""):
        super().__init__(f"[{status}] {STATUS.get(status, 'error')} {message}".strip())
        self.status = status


def csr_product(csr, in_scale, x, out_scale, y, threads=1):
    """y = in_scale * y + out_scale * A x (in place on y, float64)."""
    ext, cols, vals = _i32(csr.row_extents), _i32(csr.column_indices), _f64(csr.entries)
    assert y.dtype == np.float64 and y.flags.c_contiguous
    x = _f64(x)
    lib().oracle_csr_product(csr.rows, _p(ext), _p(cols), _p(vals), in_scale, _p(x), out_scale, _p(y), threads)
    return y


def inner_product(left, right, threads=1):
    left, right = _f64(left), _f64(right)
    return lib().oracle_inner_product(left.size, _p(left), _p(right), threads)


def vector_op(op, target, source=None, value=0.0, threads=1):
    assert target.dtype == np.float64 and target.flags.c_contiguous
    src = None if source is None else _f64(source)
    lib().oracle_vector_op(op, _p(target), None if src is None else _p(src), value, target.size, threads)
    return target


def inverse_diagonal(csr):
    out = np.empty(csr.rows, dtype=np.float64)
    failed = c_int64(-1)
    ext, cols, vals = _i32(csr.row_extents), _i32(csr.column_indices), _f64(csr.entries)
    status = lib().oracle_inverse_diagonal(csr.rows, _p(ext), _p(cols), _p(vals), _p(out), byref(failed))
    if status:
        raise OracleError(status, f"row {failed.value}")
    return out


def diagonal_product(inverse, in_scale, x, out_scale, y):
    inverse, x = _f64(inverse), _f64(x)
    lib().oracle_diagonal_product(_p(inverse), in_scale, _p(x), out_scale, _p(y), y.size)
    return y


def pcg(csr, b, x0=None, max_iterations=5000, absolute_tolerance=0.0, relative_tolerance=1e-10, precondition=True,
        threads=1, in_scale=0.0, out_scale=1.0, history=False):
    """ConjugateGradients::product(in_scale, b, out_scale, x) with Jacobi; returns (x, stats, history, seconds)."""
    ext, cols, vals = _i32(csr.row_extents), _i32(csr.column_indices), _f64(csr.entries)
    b = _f64(b)
    x = np.zeros(csr.rows) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    stats = Statistics()
    capacity = max_iterations if history else 0
    hist = np.zeros(max(1, capacity))
    seconds = c_double()
    status = lib().oracle_pcg(csr.rows, _p(ext), _p(cols), _p(vals), in_scale, _p(b), out_scale, _p(x),
                              max_iterations, absolute_tolerance, relative_tolerance, 1 if precondition else 0,
                              threads, byref(stats), _p(hist), capacity, byref(seconds))
    if status:
        raise OracleError(status)
    count = min(capacity, stats.iterationCount + 1)
    return x, stats, hist[:count] if history else None, seconds.value


def element_gradient(dofs, element_matrix, lam, u):
    """dJ/dk_e = -lambda_e^T K_e0 u_e for every element (dofs: E x npe, -1 = eliminated)."""
    dofs = np.ascontiguousarray(dofs, dtype=np.int32)
    ke = _f64(element_matrix)
    lam, u = _f64(lam), _f64(u)
    out = np.empty(dofs.shape[0])
    lib().oracle_element_gradient(dofs.shape[0], dofs.shape[1], _p(dofs), _p(ke), _p(lam), _p(u), _p(out))
    return out


def assemble(csr, dofs, matrices, scale=None):
    """Entries of sum_e scale_e P_e^T K_e P_e on the pattern of `csr`, elements in order.

    matrices: one nodes x nodes matrix shared by all elements, or elements x nodes x nodes."""
    dofs = np.ascontiguousarray(dofs, dtype=np.int32)
    elements, nodes = dofs.shape
    matrices = _f64(matrices)
    stride = 0 if matrices.size == nodes * nodes and matrices.ndim <= 2 else nodes * nodes
    assert matrices.size == (nodes * nodes if stride == 0 else elements * nodes * nodes)
    scale = None if scale is None else _f64(scale)
    ext, cols = _i32(csr.row_extents), _i32(csr.column_indices)
    out = np.empty(int(ext[-1]))
    status = lib().oracle_assemble(csr.rows, _p(ext), _p(cols), elements, nodes, _p(dofs), _p(matrices), stride,
                                   None if scale is None else _p(scale), _p(out))
    if status:
        raise OracleError(status, "a coupling is missing from the pattern")
    return out


def assemble_vector(rows, dofs, vectors, scale=None):
    """sum_e scale_e P_e^T v_e; vectors: one shared nodes vector or elements x nodes."""
    dofs = np.ascontiguousarray(dofs, dtype=np.int32)
    elements, nodes = dofs.shape
    vectors = _f64(vectors)
    stride = 0 if vectors.size == nodes and vectors.ndim == 1 else nodes
    assert vectors.size == (nodes if stride == 0 else elements * nodes)
    scale = None if scale is None else _f64(scale)
    out = np.empty(rows)
    status = lib().oracle_assemble_vector(rows, elements, nodes, _p(dofs), _p(vectors), stride,
                                          None if scale is None else _p(scale), _p(out))
    if status:
        raise OracleError(status, "row index out of range")
    return out


REORDERINGS = {"none": 0, "cuthill-mckee": 1, "reverse-cuthill-mckee": 2}


def degree_sort(items, key):
    """items sorted by key[item] in the order GCC's std::sort leaves them (ties included)."""
    items, key = _i32(items).copy(), _i32(key)
    lib().oracle_degree_sort(items.size, _p(items), _p(key))
    return items


def heap_sort(items, key):
    """The heapsort step of that order alone (std::partial_sort over the whole range)."""
    items, key = _i32(items).copy(), _i32(key)
    lib().oracle_heap_sort(items.size, _p(items), _p(key))
    return items


def make_reordering(csr, strategy="reverse-cuthill-mckee"):
    """map[new] = old (makeReordering, reorder.cpp:23-173)."""
    ext, cols = _i32(csr.row_extents), _i32(csr.column_indices)
    out = np.empty(csr.rows, dtype=np.int32)
    status = lib().oracle_make_reordering(csr.rows, _p(ext), _p(cols), REORDERINGS[strategy], _p(out))
    if status:
        raise OracleError(status, "make_reordering")
    return out


def reorder_csr(csr, mapping, reverse=False):
    """The reordered CSR arrays (reorder.cpp:215-365), as (row_extents, column_indices, entries)."""
    mapping = _i32(mapping)
    ext, cols, vals = _i32(csr.row_extents), _i32(csr.column_indices), _f64(csr.entries)
    new_ext = np.empty(csr.rows + 1, dtype=np.int32)
    new_cols = np.empty(max(1, csr.nnz), dtype=np.int32)
    new_vals = np.empty(max(1, csr.nnz), dtype=np.float64)
    status = lib().oracle_reorder_csr(csr.rows, _p(mapping), 1 if reverse else 0, _p(ext), _p(cols), _p(vals),
                                      _p(new_ext), _p(new_cols), _p(new_vals))
    if status:
        raise OracleError(status, "reorder_csr: the map is not a permutation")
    return new_ext, new_cols[:csr.nnz], new_vals[:csr.nnz]


def reorder_vector(mapping, values, reverse=False):
    mapping, values = _i32(mapping), _f64(values)
    out = np.empty_like(values)
    status = lib().oracle_reorder_vector(values.size, _p(mapping), 1 if reverse else 0, _p(values), _p(out))
    if status:
        raise OracleError(status, "reorder_vector")
    return out


_ref = None


def ref_lib():
    """oracle/_ref/libref_matrix_market.so: the reference's own MatrixMarket writer (None if not built and the
    reference tree is absent)."""
    global _ref
    if _ref is None:
        from .build import REF_OUTPUT, build_reference

        if not REF_OUTPUT.exists() and build_reference() is None:
            return None
        l = ctypes.CDLL(str(REF_OUTPUT))
        P = c_void_p
        for suffix in ("f64", "f32"):
            getattr(l, f"ref_mm_write_csr_{suffix}").argtypes = [ctypes.c_char_p, c_int, c_int, c_int, P, P, P]
            getattr(l, f"ref_mm_write_array_{suffix}").argtypes = [ctypes.c_char_p, P, c_int64]
        _ref = l
    return _ref


def ref_write_csr(path, csr, dtype=np.float64):
    """Writes `csr` with the reference's MatrixMarket::Output (coordinate real general)."""
    suffix = "f32" if dtype == np.float32 else "f64"
    entries = np.ascontiguousarray(csr.entries, dtype=dtype)
    ext, cols = _i32(csr.row_extents), _i32(csr.column_indices)
    status = getattr(ref_lib(), f"ref_mm_write_csr_{suffix}")(str(path).encode(), csr.rows, csr.columns, csr.nnz,
                                                               _p(ext), _p(cols), _p(entries))
    assert status == 0


def ref_write_array(path, values, dtype=np.float64):
    suffix = "f32" if dtype == np.float32 else "f64"
    values = np.ascontiguousarray(values, dtype=dtype)
    status = getattr(ref_lib(), f"ref_mm_write_array_{suffix}")(str(path).encode(), _p(values), values.size)
    assert status == 0


def _u16(a):
    return np.ascontiguousarray(a, dtype=np.uint16)


def _csr(csr):
    return _i32(csr.row_extents), _i32(csr.column_indices), _f64(csr.entries)


def _check(status, what):
    if status:
        raise OracleError(status, what)


def masked_csr_product(csr, mask, threshold, in_scale, x, out_scale, y, threads=1):
    """MaskedCSROperator::product, in place on y."""
    ext, cols, vals = _csr(csr)
    mask, x = _u16(mask), _f64(x)
    assert y.dtype == np.float64 and y.flags.c_contiguous
    lib().oracle_masked_csr_product(csr.rows, _p(ext), _p(cols), _p(vals), _p(mask), threshold, in_scale, _p(x),
                                    out_scale, _p(y), threads)
    return y


def masked_identity(mask, threshold, in_scale, x, out_scale, y):
    mask, x = _u16(mask), _f64(x)
    lib().oracle_masked_identity(y.size, _p(mask), threshold, in_scale, _p(x), out_scale, _p(y))
    return y


def jacobi(csr, iterations, relaxation, in_scale, x, out_scale, y, mask=None, threshold=0, threads=1):
    """JacobiOperator over csr (or the masked csr) with its full inverse diagonal, in place on y."""
    ext, cols, vals = _csr(csr)
    m = None if mask is None else _u16(mask)
    x = _f64(x)
    _check(lib().oracle_jacobi(csr.rows, _p(ext), _p(cols), _p(vals), None if m is None else _p(m), threshold,
                               iterations, relaxation, in_scale, _p(x), out_scale, _p(y), threads), "jacobi")
    return y


def masked_jacobi(csr, mask, threshold, iterations, relaxation, in_scale, x, out_scale, y):
    ext, cols, vals = _csr(csr)
    mask, x = _u16(mask), _f64(x)
    _check(lib().oracle_masked_jacobi(csr.rows, _p(ext), _p(cols), _p(vals), _p(mask), threshold, iterations,
                                      relaxation, in_scale, _p(x), out_scale, _p(y)), "masked_jacobi")
    return y


def approximate_substitution(csr, iterations, in_scale, x, out_scale, y):
    ext, cols, vals = _csr(csr)
    x = _f64(x)
    _check(lib().oracle_approximate_substitution(csr.rows, _p(ext), _p(cols), _p(vals), iterations, in_scale, _p(x),
                                                 out_scale, _p(y)), "approximate_substitution")
    return y


def pmg(csr, mask, order, b, relative_tolerance=1e-6, max_cycles=500, threads=1):
    """solveMultigrid (poisson2D/solver.cpp:111-232) from x = 0.

    Returns (x, residual norm after each cycle, coarse CG iterations of each cycle)."""
    ext, cols, vals = _csr(csr)
    mask, b = _u16(mask), _f64(b)
    x = np.zeros(csr.rows)
    history = np.zeros(max(1, max_cycles))
    coarse_history = np.zeros(max(1, max_cycles), dtype=np.uint64)
    cycles, coarse = c_int64(), c_uint64()
    _check(lib().oracle_pmg(csr.rows, _p(ext), _p(cols), _p(vals), _p(mask), order, _p(b), _p(x),
                            relative_tolerance, max_cycles, _p(history), _p(coarse_history), max_cycles,
                            byref(cycles), byref(coarse), threads), "pmg")
    return x, history[:cycles.value].copy(), coarse_history[:cycles.value].astype(np.int64)


def element_stiffness(moments, jacobians, modulus=None):
    """K_e for affine cells from the reference gradient moments (dim, dim, npe, npe); jacobians (E, dim, dim)."""
    moments, jacobians = _f64(moments), _f64(jacobians)
    dim, npe, elements = moments.shape[0], moments.shape[2], jacobians.shape[0]
    modulus = None if modulus is None else _f64(modulus)
    out = np.empty((elements, npe, npe))
    _check(lib().oracle_element_stiffness(elements, dim, npe, _p(moments), _p(jacobians),
                                          None if modulus is None else _p(modulus), _p(out)), "element_stiffness")
    return out

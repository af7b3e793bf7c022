"""MatrixMarket files of CSR matrices and vectors (cieutils/packages/io/inc/MatrixMarket.hpp).

``MatrixMarket.Output(path)`` mirrors the reference's writer: called with
(rows, columns, nnz, row_extents, column_indices, entries) it writes a
coordinate file, called with one vector an array file (MatrixMarket.cpp:36-78),
in the reference's number format unless ``digits`` asks for more.
``MatrixMarket.Input(path).read()`` has no reference counterpart; it returns a
host ``systems.CSR`` for coordinate files and a numpy array for array files.
Both run in libcie_hip (host code).
"""
from ctypes import POINTER, byref, c_double, c_int, c_int32, c_int64, c_void_p, cast

import numpy as np

from . import systems
from ._lib import check, library


def _path(path):
    return str(path).encode()


def _copy(pointer, ctype, count):
    """A numpy copy of `count` values behind a library pointer."""
    dtype = np.float64 if ctype is c_double else np.int32
    if count == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(cast(pointer, POINTER(ctype)), shape=(count,)).copy()


class MatrixMarket:
    class Output:
        """Writes to `path`; digits 0 = std::ostream's default 6 significant digits, as the reference."""

        def __init__(self, path, digits=0):
            self.path = path
            self.digits = int(digits)

        def __call__(self, *arguments):
            if len(arguments) == 1:
                values = np.asarray(arguments[0])
                suffix = "f32" if values.dtype == np.float32 else "f64"
                values = np.ascontiguousarray(values, dtype=np.float32 if suffix == "f32" else np.float64).ravel()
                check(getattr(library(), f"cie_hip_mm_write_array_{suffix}")(
                    _path(self.path), values.ctypes.data_as(c_void_p), values.size, self.digits),
                    "MatrixMarket.Output")
                return self
            rows, columns, nnz, extents, indices, entries = arguments
            extents = np.ascontiguousarray(extents, dtype=np.int32)
            indices = np.ascontiguousarray(indices, dtype=np.int32)
            entries = np.asarray(entries)
            suffix = "f32" if entries.dtype == np.float32 else "f64"
            entries = np.ascontiguousarray(entries, dtype=np.float32 if suffix == "f32" else np.float64)
            if extents.size != int(rows) + 1 or indices.size < int(nnz) or entries.size < int(nnz):
                raise ValueError("array sizes do not match rows and nnz")
            check(getattr(library(), f"cie_hip_mm_write_csr_{suffix}")(
                _path(self.path), int(rows), int(columns), int(nnz), extents.ctypes.data_as(c_void_p),
                indices.ctypes.data_as(c_void_p), entries.ctypes.data_as(c_void_p), self.digits),
                "MatrixMarket.Output")
            return self

        def write_csr(self, csr):
            return self(csr.rows, csr.columns, csr.nnz, csr.row_extents, csr.column_indices, csr.entries)

    class Input:
        def __init__(self, path):
            self.path = path

        def read(self):
            handle = c_void_p()
            check(library().cie_hip_mm_read(_path(self.path), byref(handle)), "MatrixMarket.Input")
            try:
                rows, columns, nnz, dense = c_int64(), c_int64(), c_int64(), c_int()
                check(library().cie_hip_mm_shape(handle, byref(rows), byref(columns), byref(nnz), byref(dense)))
                ext_p, col_p, val_p = c_void_p(), c_void_p(), c_void_p()
                check(library().cie_hip_mm_arrays(handle, byref(ext_p), byref(col_p), byref(val_p)))
                count = nnz.value
                values = _copy(val_p, c_double, count)
                if dense.value:
                    return values if columns.value == 1 else values.reshape(columns.value, rows.value).T.copy()
                extents = _copy(ext_p, c_int32, rows.value + 1)
                indices = _copy(col_p, c_int32, count)
                return systems.CSR(rows.value, columns.value, extents, indices, values)
            finally:
                library().cie_hip_mm_free(handle)

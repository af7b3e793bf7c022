"""Python mirror of the reference's solver interface, backed by the HIP C ABI.

Class and method names follow linalg/packages/solvers (paths relative to the
reference repository):

    HIPSpace              LinalgSpaceLike (inc/LinalgSpace.hpp:10-35), like
                          DefaultSpace / SYCLSpace (src/SYCLSpace.cpp:114-343)
    LinearOperator        inc/LinearOperator.hpp:10-34
    HIPCSROperator        CSROperator / SYCLCSROperator (src/SYCLCSROperator.cpp:13-89)
    DiagonalOperator,
    makeDiagonalOperator  src/DiagonalOperator.cpp:17-162
    ConjugateGradients    impl/ConjugateGradients_impl.hpp:22-201 (generic, over
                          any space; every step is one space/operator call)
    FusedPCG, MixedPCG    the fused device solvers (cie_hip_pcg_f64 / _mixed)

Vectors are torch tensors on a HIP device; torch only provides the memory.
Every compute call goes through libcie_hip.so.
"""
import ctypes
import logging
import math
import sys
import weakref
from ctypes import byref, c_double, c_float, c_int, c_int64, c_void_p

import numpy as np

from . import _lib
from ._lib import MixedSettings, SolveInfo, Statistics, check, library

_DBL_MIN = sys.float_info.min
_DBL_EPS = sys.float_info.epsilon


def _torch():
    import torch

    return torch


_KERNEL_NAMES = {_lib.SPMV_VECTOR: "vector", _lib.SPMV_STAGED: "staged", _lib.SPMV_TILED: "tiled",
                 _lib.SPMV_BLOCKS: "blocks", _lib.SPMV_SORTED: "sorted", _lib.SPMV_PAIRED: "paired"}


def _kernel_code(name):
    """C-ABI code of an SpMV kernel name ("vector", "staged", "tiled", "blocks", "sorted")."""
    for code, known in _KERNEL_NAMES.items():
        if known == name:
            return code
    raise ValueError(f"unknown SpMV kernel {name!r}")


class Context:
    """One device + one HIP stream (cie_hip_context).

    The objects created on a context (matrices, masks, batches, assembly and multigrid plans) hold it; closing
    the context first closes those still open, newest first, so the order in which the garbage collector
    finalizes them does not matter."""

    _defaults = {}

    def __init__(self, device=0, stream=None):
        self.device = device
        self.handle = c_void_p()
        self._dependents = []  # weak references, in creation order
        check(library().cie_hip_context_create(device, stream, byref(self.handle)), "cie_hip_context_create")

    def _register(self, obj):
        """Records an object whose library handle lives on this context."""
        if len(self._dependents) >= 256 and len(self._dependents) % 256 == 0:
            self._dependents = [ref for ref in self._dependents if ref() is not None]
        self._dependents.append(weakref.ref(obj))

    @classmethod
    def default(cls, device=0):
        if device not in cls._defaults:
            cls._defaults[device] = cls(device)
        return cls._defaults[device]

    def synchronize(self):
        check(library().cie_hip_context_synchronize(self.handle))

    def close(self):
        if self.handle:
            for ref in reversed(getattr(self, "_dependents", [])):
                obj = ref()
                if obj is not None:
                    obj.close()
            self._dependents = []
            library().cie_hip_context_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _sync_torch():
    torch = _torch()
    if torch.cuda.is_available():
        torch.cuda.synchronize()


def _dptr(tensor):
    return c_void_p(tensor.data_ptr())


class CSRMatrix:
    """Device-resident CSR (int32 extents/columns, FP64 and optional FP32 planes).

    Built from host arrays (uploaded) -- the device counterpart of a
    CSRView<const double, const int> (utilities/inc/CSRView.hpp:11-86).
    """

    def __init__(self, csr, context=None, planes=_lib.PLANE_F64 | _lib.PLANE_F32):
        self.context = context or Context.default()
        self.rows, self.columns, self.nnz = int(csr.rows), int(csr.columns), int(csr.row_extents[-1])
        extents = np.ascontiguousarray(csr.row_extents, dtype=np.int32)
        cols = np.ascontiguousarray(csr.column_indices, dtype=np.int32)
        vals = np.ascontiguousarray(csr.entries, dtype=np.float64)
        self.handle = c_void_p()
        check(library().cie_hip_csr_create(self.context.handle, self.rows, self.columns, self.nnz,
                                           extents.ctypes.data_as(c_void_p), cols.ctypes.data_as(c_void_p),
                                           vals.ctypes.data_as(c_void_p), _lib.HOST, planes, byref(self.handle)),
              "cie_hip_csr_create")
        self.context._register(self)

    def rowCount(self):
        return self.rows

    def columnCount(self):
        return self.columns

    def launch(self, precision=64):
        """(kernel, width, unroll) of the SpMV for `precision` (64 or 32)."""
        kernel, width, unroll = c_int(), c_int(), c_int()
        check(library().cie_hip_csr_get_launch(self.handle, precision, byref(kernel), byref(width), byref(unroll)))
        return _KERNEL_NAMES[kernel.value], width.value, unroll.value

    def column_bytes(self):
        """Bytes per column index the tiled SpMV reads: 2 (narrow 16-bit plane) or 4."""
        value = c_int()
        check(library().cie_hip_csr_column_bytes(self.handle, byref(value)), "cie_hip_csr_column_bytes")
        return value.value

    def column_blocks(self, width):
        """(blocks, bytes per block column) of the 1 x width column-block form the "blocks" kernel reads."""
        blocks, value = c_int64(), c_int()
        check(library().cie_hip_csr_column_blocks(self.handle, width, byref(blocks), byref(value)),
              "cie_hip_csr_column_blocks")
        return blocks.value, value.value

    def paired_blocks(self, width):
        """(wide blocks, narrow blocks) of the paired form with width 10 * wide + narrow (53 or 32)."""
        wide, narrow = c_int64(), c_int64()
        check(library().cie_hip_csr_paired_blocks(self.handle, width, byref(wide), byref(narrow)),
              "cie_hip_csr_paired_blocks")
        return wide.value, narrow.value

    def set_launch(self, precision, kernel, width, unroll):
        code = _kernel_code(kernel)
        check(library().cie_hip_csr_set_launch(self.handle, precision, code, width, unroll), "cie_hip_csr_set_launch")

    def autotune(self, precision=64, repetitions=10,
                 kernels=("tiled", "blocks", "sorted", "paired", "staged", "vector")):
        """Time the candidate launches of `kernels` ("staged", "vector", "tiled", "blocks", "sorted", "paired") and
        keep the fastest."""
        mask = sum(1 << _kernel_code(k) for k in kernels)
        best = c_double()
        check(library().cie_hip_csr_autotune(self.handle, precision, mask, repetitions, byref(best)),
              "cie_hip_csr_autotune")
        return best.value

    @classmethod
    def _adopt(cls, handle, context, rows, columns, nnz):
        matrix = cls.__new__(cls)
        matrix.context, matrix.handle = context, handle
        matrix.rows, matrix.columns, matrix.nnz = rows, columns, nnz
        context._register(matrix)
        return matrix

    def copy_out(self):
        """The device arrays as a host CSR (row extents, column indices, FP64 entries)."""
        from .systems import CSR

        extents = np.empty(self.rows + 1, dtype=np.int32)
        columns = np.empty(self.nnz, dtype=np.int32)
        entries = np.empty(self.nnz, dtype=np.float64)
        _sync_torch()
        check(library().cie_hip_csr_copy_out(self.handle, extents.ctypes.data_as(c_void_p),
                                             columns.ctypes.data_as(c_void_p), entries.ctypes.data_as(c_void_p)),
              "cie_hip_csr_copy_out")
        return CSR(self.rows, self.columns, extents, columns, entries)

    def reorder(self, mapping, reverse=False):
        """A new device matrix with rows and columns permuted (reorder / reverseReorder, reorder.cpp:215-365).

        mapping[new] = old (reverse False) or mapping[old] = new (reverse True), a host permutation."""
        mapping = np.ascontiguousarray(mapping, dtype=np.int32)
        if mapping.shape != (self.rows,):
            raise ValueError("the map needs one entry per row")
        handle = c_void_p()
        _sync_torch()
        check(library().cie_hip_csr_reorder(self.handle, mapping.ctypes.data_as(c_void_p), 1 if reverse else 0,
                                            byref(handle)), "cie_hip_csr_reorder")
        return CSRMatrix._adopt(handle, self.context, self.rows, self.columns, self.nnz)

    def row_sum_correction(self):
        """The FP32 plane's row-sum correction that the mixed solvers' inner product adds on the diagonal:
        fl32(sum_j (a_ij - fl32(a_ij))) per row (a device float32 tensor)."""
        out = _torch().empty(self.rows, dtype=_torch().float32, device=f"cuda:{self.context.device}")
        _sync_torch()
        check(library().cie_hip_csr_row_sum_correction(self.context.handle, self.handle, _dptr(out)),
              "cie_hip_csr_row_sum_correction")
        return out

    def update_entries(self, entries):
        """Replace the FP64 entries (device tensor, same pattern); refreshes the FP32 plane."""
        _sync_torch()
        check(library().cie_hip_csr_update_entries(self.handle, _dptr(entries)), "cie_hip_csr_update_entries")

    @classmethod
    def _borrow(cls, handle, context, owner):
        """A view of a matrix that `owner` holds (kept alive by the view, never destroyed by it)."""
        rows, columns, nnz = c_int64(), c_int64(), c_int64()
        check(library().cie_hip_csr_shape(handle, byref(rows), byref(columns), byref(nnz)), "cie_hip_csr_shape")
        matrix = cls._adopt(handle, context, rows.value, columns.value, nnz.value)
        matrix._owner = owner
        return matrix

    def close(self):
        if self.handle and getattr(self, "_owner", None) is None:
            library().cie_hip_csr_destroy(self.handle)
        self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HIPSpace:
    """LinalgSpaceLike over device vectors of one value type (float64 or float32)."""

    def __init__(self, dtype=None, context=None):
        torch = _torch()
        self.dtype = dtype or torch.float64
        if self.dtype not in (torch.float64, torch.float32):
            raise TypeError("HIPSpace supports float64 and float32")
        self.context = context or Context.default()
        self.device = torch.device("cuda", self.context.device)
        self._suffix = "f64" if self.dtype == torch.float64 else "f32"
        self._scalar = c_double if self.dtype == torch.float64 else c_float

    # --- LinalgSpace.hpp:17-22 ---
    @staticmethod
    def view(vector):
        return vector

    @staticmethod
    def size(view):
        return int(view.numel())

    def makeVector(self, size):
        return _torch().empty(int(size), dtype=self.dtype, device=self.device)

    def _call(self, name, *args):
        # Synchronous like SYCLSpace (every op ends in wait_and_throw, SYCLSpace.cpp:166-311):
        # torch's stream is drained before and the library stream after the call.
        _sync_torch()
        check(getattr(library(), f"cie_hip_{name}_{self._suffix}")(self.context.handle, *args), name)
        self.context.synchronize()

    def _check(self, *vectors):
        for v in vectors:
            if v.dtype != self.dtype or v.device != self.device or not v.is_contiguous():
                raise TypeError(f"expected a contiguous {self.dtype} vector on {self.device}")

    # --- LinalgSpace.hpp:24-29 ---
    def innerProduct(self, left, right):
        self._check(left, right)
        if left.numel() != right.numel():
            raise _lib.ShapeError(_lib.ERROR_SHAPE, "innerProduct: sizes differ")
        result = self._scalar()
        self._call("inner_product", _dptr(left), _dptr(right), left.numel(), byref(result))
        return result.value

    def scale(self, target, source_or_value, value=None):
        """scale(t, s, a): t = a * t .* s;  scale(v, a): v *= a (DefaultSpace.cpp:80-122)."""
        if value is None:
            self._check(target)
            self._call("scale_uniform", _dptr(target), target.numel(), self._scalar(source_or_value))
        else:
            self._check(target, source_or_value)
            self._call("scale", _dptr(target), _dptr(source_or_value), target.numel(), self._scalar(value))

    def add(self, target, source, value):
        self._check(target, source)
        self._call("add", _dptr(target), _dptr(source), target.numel(), self._scalar(value))

    def assign(self, target, source):
        """Device<-device copy, or host<->device when one side is a numpy array (SYCLSpace.hpp:116-122)."""
        if isinstance(source, np.ndarray):
            self._check(target)
            target.copy_(_torch().from_numpy(np.ascontiguousarray(source, dtype=np.float64 if self._suffix == "f64"
                                                                  else np.float32)))
            _sync_torch()
        elif isinstance(target, np.ndarray):
            self._check(source)
            _sync_torch()
            target[...] = source.cpu().numpy()
        else:
            self._check(target, source)
            self._call("assign", _dptr(target), _dptr(source), target.numel())

    def fill(self, target, value):
        self._check(target)
        self._call("fill", _dptr(target), target.numel(), self._scalar(value))

    def getContext(self):
        return self.context


class LinearOperator:
    """out = inScale * out + outScale * A * in (LinearOperator.hpp:15-33)."""

    def product(self, inScale, in_, outScale, out):
        raise NotImplementedError


class HIPCSROperator(LinearOperator):
    def __init__(self, matrix, space):
        self.matrix = matrix
        self.space = space

    def product(self, inScale, in_, outScale, out):
        self.space._check(in_, out)
        suffix = self.space._suffix
        _sync_torch()
        fn = getattr(library(), f"cie_hip_csr_product_{suffix}")
        check(fn(self.space.context.handle, self.matrix.handle, inScale, _dptr(in_), in_.numel(), outScale,
                 _dptr(out), out.numel()), "product")
        self.space.context.synchronize()


class DiagonalOperator(LinearOperator):
    """Jacobi preconditioner holding the inverse diagonal (DiagonalOperator.cpp:37-53)."""

    def __init__(self, inverseDiagonal, space):
        self.inverseDiagonal = inverseDiagonal
        self.space = space

    def product(self, inScale, in_, outScale, out):
        self.space._check(in_, out)
        _sync_torch()
        fn = getattr(library(), f"cie_hip_diagonal_product_{self.space._suffix}")
        check(fn(self.space.context.handle, _dptr(self.inverseDiagonal), inScale, _dptr(in_), outScale, _dptr(out),
                 out.numel()), "DiagonalOperator.product")
        self.space.context.synchronize()


def makeDiagonalOperator(matrix, space):
    """inverse_i = 1 / a_ii; raises ZeroDiagonal / MissingDiagonal like DiagonalOperator.cpp:76-82."""
    inverse = space.makeVector(matrix.rows)
    failed = c_int64(-1)
    _sync_torch()
    fn = getattr(library(), f"cie_hip_inverse_diagonal_{space._suffix}")
    check(fn(space.context.handle, matrix.handle, _dptr(inverse), byref(failed)), "makeDiagonalOperator")
    return DiagonalOperator(inverse, space)


def _float_equal(left, right, value=float):
    """cieutils/packages/maths/impl/Comparison_impl.hpp:49-56 with the default tolerances of the value type
    (numeric_limits<T>::min() and epsilon(), :36-40); `value` is float (double) or numpy.float32."""
    if value is float:
        norm = min(abs(left) + abs(right), sys.float_info.max)
        return abs(left - right) < max(_DBL_MIN, _DBL_EPS * norm)
    info = np.finfo(np.float32)
    left, right = np.float32(left), np.float32(right)
    norm = min(np.float32(abs(left) + abs(right)), info.max)
    return bool(abs(np.float32(left - right)) < max(info.tiny, np.float32(info.eps * norm)))


def _float_less(left, right, value=float):
    return bool(value(left) < value(right)) and not _float_equal(left, right, value)


def makeIterationReport(verbosity, reportType, status, settings):
    """IterativeSolver::makeIterationReport (linalg/packages/solvers/src/IterativeSolver.cpp:17-58), formatted
    by cie_hip_iteration_report: the table header at iteration 0, a coloured row on termination, on convergence
    or (verbosity >= 3) on every iteration, the closing rule once converged; empty below verbosity 2."""
    lib = library()
    length = ctypes.c_size_t(0)
    check(lib.cie_hip_iteration_report(verbosity, reportType, byref(status), byref(settings), None, 0, byref(length)),
          "cie_hip_iteration_report")
    buffer = ctypes.create_string_buffer(length.value + 1)
    check(lib.cie_hip_iteration_report(verbosity, reportType, byref(status), byref(settings), buffer, len(buffer),
                                       byref(length)), "cie_hip_iteration_report")
    return buffer.value.decode()


# The reports go to this logger, one record per non-empty report, as the
# reference logs them into its "ConjugateGradients" log block
# (ConjugateGradients_impl.hpp:72-74); each solver also keeps them in `reports`.
REPORT_LOG = logging.getLogger("cie_amd.ConjugateGradients")

# Loop trips whose residual the fused solvers record for their reports; the
# report of the final trip always comes from the solver's statistics.
HISTORY_CAPACITY = 1 << 20


class IterativeSolver(LinearOperator):
    def __init__(self, settings=None):
        self._settings = settings or Statistics()
        self._results = None
        self.reports = []

    def _log_report(self, verbosity, reportType, stats, settings):
        text = makeIterationReport(verbosity, reportType, stats, settings)
        if text:
            self.reports.append(text)
            REPORT_LOG.info(text)

    def getConfiguration(self):
        return self._settings

    def configure(self, settings):
        self._settings = settings

    def getStats(self):
        return self._results

    def report(self, stats):
        self._results = stats


class ConjugateGradients(IterativeSolver):
    """Generic (P)CG over any LinalgSpace; mirrors ConjugateGradients_impl.hpp:49-201 step by step.

    Every operation is one HIPSpace / operator call, so this path exercises the
    backend exactly as the reference's solver template would.
    """

    def __init__(self, lhs, space, preconditioner=None, settings=None, verbosity=0):
        super().__init__(settings)
        self._lhs, self._space, self._preconditioner, self._verbosity = lhs, space, preconditioner, verbosity
        self.history = []

    def product(self, inScale, in_, outScale, out):
        space = self._space
        size = space.size(out)
        if space.size(in_) != size:
            raise _lib.ShapeError(_lib.ERROR_SHAPE, f"incompatible vector sizes {space.size(in_)} != {size}")
        settings = self.getConfiguration()
        # The scalars are TS::Value: double, or float for ConjugateGradients<...<float>>
        # (ConjugateGradients.cpp:10,16), whose sqrt, quotients and Comparison<float> round in float.
        V = float if getattr(space, "_suffix", "f64") == "f64" else np.float32
        sqrt = math.sqrt if V is float else (lambda value: np.sqrt(np.float32(value)))
        less = lambda left, right: _float_less(left, right, V)  # noqa: E731
        if V is not float:
            settings = Statistics(settings.iterationCount, float(np.float32(settings.absoluteResidual)),
                                  float(np.float32(settings.relativeResidual)))
        maximum = sys.float_info.max if V is float else float(np.finfo(np.float32).max)
        stats = Statistics(0, maximum, maximum)
        self.history = []
        self.reports = []
        solution = space.makeVector(size)
        space.assign(solution, out)
        if settings.iterationCount:
            search = space.makeVector(size)
            searchProduct = space.makeVector(size)
            space.fill(searchProduct, 0)
            preconditioned = None
            residual = space.makeVector(size)
            space.assign(residual, in_)
            self._lhs.product(1, solution, -1, residual)
            residualNorm = V(space.innerProduct(residual, residual))
            initialResidualNorm = V(sqrt(residualNorm))
            stats.absoluteResidual = float(initialResidualNorm)
            stats.relativeResidual = 1.0
            if less(stats.absoluteResidual, settings.absoluteResidual) or less(stats.relativeResidual,
                                                                              settings.relativeResidual):
                # :100-111 return before the output step of :190-198: `out` keeps its value.
                self._log_report(self._verbosity, _lib.REPORT_TERMINATION, stats, settings)
                self.report(stats)
                return
            if self._preconditioner is not None:
                preconditioned = space.makeVector(size)
                self._preconditioner.product(0, residual, 1, preconditioned)
                preconditionedNorm = V(space.innerProduct(residual, preconditioned))
                space.assign(search, preconditioned)
                if not 0 < preconditionedNorm:
                    raise _lib.Breakdown(_lib.ERROR_BREAKDOWN, "preconditioned norm vanished in iteration 0 (r^T @ z = 0)")
            else:
                preconditionedNorm = residualNorm
                space.assign(search, residual)
                if not 0 < preconditionedNorm:
                    raise _lib.Breakdown(_lib.ERROR_BREAKDOWN, "residual norm vanished in iteration 0 (r^T @ r = 0)")
            while stats.iterationCount < settings.iterationCount:
                self._lhs.product(0, search, 1, searchProduct)
                denominator = V(space.innerProduct(search, searchProduct))
                if not 0 < denominator:
                    raise _lib.NotPositiveDefinite(_lib.ERROR_NOT_SPD,
                                                   f"p^T @ A @ p vanished in iteration {stats.iterationCount}")
                searchScale = V(preconditionedNorm / denominator)
                space.add(solution, search, searchScale)
                space.add(residual, searchProduct, -searchScale)
                residualNorm = V(space.innerProduct(residual, residual))
                previous = preconditionedNorm
                if preconditioned is not None:
                    self._preconditioner.product(0, residual, 1, preconditioned)
                    preconditionedNorm = V(space.innerProduct(residual, preconditioned))
                else:
                    preconditionedNorm = residualNorm
                searchScale = V(preconditionedNorm / previous)
                space.scale(search, searchScale)
                space.add(search, preconditioned if preconditioned is not None else residual, 1)
                absolute = V(sqrt(residualNorm))
                stats.absoluteResidual = float(absolute)
                stats.relativeResidual = float(V(absolute / initialResidualNorm))
                self.history.append(stats.absoluteResidual)
                self._log_report(self._verbosity, _lib.REPORT_ITERATION, stats, settings)
                if less(stats.absoluteResidual, settings.absoluteResidual) or less(stats.relativeResidual,
                                                                                  settings.relativeResidual):
                    break
                stats.iterationCount += 1
        self.report(stats)
        return self._finish(inScale, outScale, out, solution)

    def _finish(self, inScale, outScale, out, solution):
        space = self._space
        if inScale == 1:
            space.add(out, solution, outScale)
        elif inScale == 0:
            space.assign(out, solution)
            space.scale(out, outScale)
        else:
            space.scale(out, inScale)
            space.add(out, solution, outScale)


class _FusedSolver(IterativeSolver):
    """A fused solver behind the IterativeSolver face; with verbosity >= 2 it records each loop trip's residual
    on the device (cie_hip_context_set_residual_history) and logs the reference's per-iteration reports from it
    after the solve (ConjugateGradients_impl.hpp:100-111 for the start check, :175-182 for each trip)."""

    def __init__(self, matrix, settings=None, verbosity=0):
        super().__init__(settings)
        self.matrix = matrix
        self.verbosity = verbosity
        self.info = None
        self.history = None

    def _solve(self, call, name):
        settings = self.getConfiguration()
        info = SolveInfo()
        lib = library()
        context = self.matrix.context.handle
        recorded = None
        if self.verbosity >= 2 and settings.iterationCount:
            recorded = _torch().empty(max(1, min(int(settings.iterationCount), HISTORY_CAPACITY)), dtype=_torch().float64,
                                      device=f"cuda:{self.matrix.context.device}")
            check(lib.cie_hip_context_set_residual_history(context, _dptr(recorded), recorded.numel()),
                  "cie_hip_context_set_residual_history")
        _sync_torch()
        try:
            status = call(lib, info)
        finally:
            if recorded is not None:
                lib.cie_hip_context_set_residual_history(context, None, 0)
        self.info = info
        self.reports = []
        rows = int(info.reportedIterations)
        self.history = recorded[:min(rows, recorded.numel())].cpu().numpy() if recorded is not None else None
        if recorded is not None:
            self._log_reports(info, settings)
        check(status, name)
        self.report(Statistics(info.stats.iterationCount, info.stats.absoluteResidual, info.stats.relativeResidual))

    def _log_reports(self, info, settings):
        rows = int(info.reportedIterations)
        if rows == 0:
            if info.converged:  # the start check ended the solve: one Termination report
                self._log_report(self.verbosity, _lib.REPORT_TERMINATION,
                                 Statistics(0, info.initialResidual, 1.0), settings)
            return
        for trip in range(rows):
            if trip == rows - 1:  # the solver's own final statistics (the mixed form's FP64 residual at a cap)
                status = Statistics(trip, info.stats.absoluteResidual, info.stats.relativeResidual)
            elif trip < len(self.history):
                absolute = float(self.history[trip])
                status = Statistics(trip, absolute, absolute / info.initialResidual)
            else:
                continue  # beyond the recorded trips
            self._log_report(self.verbosity, _lib.REPORT_ITERATION, status, settings)


class FusedPCG(_FusedSolver):
    """Jacobi-PCG in three fused kernels per iteration (cie_hip_pcg_f64)."""

    def __init__(self, matrix, settings=None, true_residual=True, verbosity=0):
        super().__init__(matrix, settings, verbosity)
        self.true_residual = true_residual

    def product(self, inScale, in_, outScale, out):
        settings = self.getConfiguration()
        self._solve(lambda lib, info: lib.cie_hip_pcg_f64(
            self.matrix.context.handle, self.matrix.handle, inScale, _dptr(in_), outScale, _dptr(out), out.numel(),
            byref(settings), 1 if self.true_residual else 0, byref(info)), "cie_hip_pcg_f64")


class MixedPCG(_FusedSolver):
    """FP32 inner PCG with FP64 solution accumulation and FP64 residual refinement (cie_hip_pcg_mixed)."""

    def __init__(self, matrix, settings=None, refinement_factor=_lib.DEFAULT_REFINEMENT_FACTOR, restart=False,
                 row_sum_correction=True, verbosity=0):
        super().__init__(matrix, settings, verbosity)
        self.refinement_factor = refinement_factor
        self.restart = restart
        self.row_sum_correction = row_sum_correction  # False: the inner product is the plain FP32 plane

    def product(self, inScale, in_, outScale, out):
        flags = 0 if self.row_sum_correction else _lib.MIXED_PLAIN_FP32
        mixed = MixedSettings(self.getConfiguration(), self.refinement_factor, 1 if self.restart else 0, flags)
        self._solve(lambda lib, info: lib.cie_hip_pcg_mixed(
            self.matrix.context.handle, self.matrix.handle, inScale, _dptr(in_), outScale, _dptr(out), out.numel(),
            byref(mixed), byref(info)), "cie_hip_pcg_mixed")


class Mask:
    """An ansatz mask on the device (one uint16 per row; makeAnsatzMask, fem/packages/graph/src/Assembler.cpp:194-213)."""

    def __init__(self, values, context=None):
        self.context = context or Context.default()
        self.values = np.ascontiguousarray(values, dtype=np.uint16)
        self.handle = c_void_p()
        check(library().cie_hip_mask_create(self.context.handle, self.values.ctypes.data_as(c_void_p),
                                            self.values.size, byref(self.handle)), "cie_hip_mask_create")
        self.context._register(self)

    def close(self):
        if self.handle:
            library().cie_hip_mask_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def maskedMatrix(matrix, mask, threshold):
    """The matrix of MaskedCSROperator(lhs, mask, threshold) (MaskedCSROperator.cpp:37-118): rows and columns with
    mask >= threshold compacted away, so its products are the masked products."""
    handle = c_void_p()
    _sync_torch()
    check(library().cie_hip_csr_masked(matrix.handle, mask.handle, int(threshold), byref(handle)), "cie_hip_csr_masked")
    rows, columns, nnz = c_int64(), c_int64(), c_int64()
    check(library().cie_hip_csr_shape(handle, byref(rows), byref(columns), byref(nnz)))
    masked = CSRMatrix._adopt(handle, matrix.context, rows.value, columns.value, nnz.value)
    masked.source = matrix  # keeps the source alive as long as the masked copy is used with it
    return masked


class MaskedCSROperator(HIPCSROperator):
    """MaskedCSROperator(lhs, mask, threshold) over a HIPSpace (MaskedCSROperator.hpp)."""

    def __init__(self, matrix, mask, threshold, space):
        super().__init__(maskedMatrix(matrix, mask, threshold), space)


class MaskedIdentityOperator(LinearOperator):
    """out = inScale out + outScale (mask < threshold ? in : 0) (MaskedIdentityOperator.cpp:14-70)."""

    def __init__(self, space, mask, threshold):
        self.space, self.mask, self.threshold = space, mask, int(threshold)

    def product(self, inScale, in_, outScale, out):
        self.space._check(in_, out)
        _sync_torch()
        check(library().cie_hip_masked_identity_f64(self.space.context.handle, self.mask.handle, self.threshold,
                                                    inScale, _dptr(in_), outScale, _dptr(out), out.numel()),
              "MaskedIdentityOperator.product")


class JacobiOperator(LinearOperator):
    """JacobiOperator(space, size, lhs, inverseDiagonal, iterations, relaxation) (JacobiOperator.cpp:15-62).

    lhs is a CSRMatrix (a masked one included), inverseDiagonal a DiagonalOperator."""

    def __init__(self, space, size, lhs, inverseDiagonal, iterations=1, relaxation=1.0):
        self.space, self.size, self.lhs = space, int(size), lhs
        self.inverseDiagonal, self.iterations, self.relaxation = inverseDiagonal, int(iterations), float(relaxation)

    def product(self, inScale, in_, outScale, out):
        self.space._check(in_, out)
        _sync_torch()
        check(library().cie_hip_jacobi_f64(self.space.context.handle, self.lhs.handle,
                                           _dptr(self.inverseDiagonal.inverseDiagonal), self.iterations,
                                           self.relaxation, inScale, _dptr(in_), outScale, _dptr(out), out.numel()),
              "JacobiOperator.product")


class MaskedJacobiOperator(LinearOperator):
    """MaskedJacobiOperator (MaskedJacobiOperator.cpp:46-161) on the device."""

    def __init__(self, matrix, iterations, relaxation, mask, threshold, space):
        self.matrix, self.iterations, self.relaxation = matrix, int(iterations), float(relaxation)
        self.mask, self.threshold, self.space = mask, int(threshold), space

    def product(self, inScale, in_, outScale, out):
        self.space._check(in_, out)
        _sync_torch()
        check(library().cie_hip_masked_jacobi_f64(self.space.context.handle, self.matrix.handle, self.mask.handle,
                                                  self.threshold, self.iterations, self.relaxation, inScale,
                                                  _dptr(in_), outScale, _dptr(out), out.numel()),
              "MaskedJacobiOperator.product")


class ApproximateSubstitution(LinearOperator):
    """ApproximateSubstitution (ApproximateSubstitution.cpp:27-111) on the device."""

    def __init__(self, matrix, iterations, space):
        self.matrix, self.iterations, self.space = matrix, int(iterations), space

    def product(self, inScale, in_, outScale, out):
        self.space._check(in_, out)
        _sync_torch()
        check(library().cie_hip_approximate_substitution_f64(self.space.context.handle, self.matrix.handle,
                                                             self.iterations, inScale, _dptr(in_), outScale,
                                                             _dptr(out), out.numel()),
              "ApproximateSubstitution.product")


class NestedProductOperator(LinearOperator):
    """lhs(rhs(in)) (NestedProductOperator_impl.hpp:24-32): buffer = rhs(in), then lhs.product(inScale, buffer, ...)."""

    def __init__(self, space, lhs, rhs, rowCount):
        self.space, self.lhs, self.rhs = space, lhs, rhs
        self.buffer = space.makeVector(rowCount)
        space.fill(self.buffer, 0)

    def product(self, inScale, in_, outScale, out):
        self.rhs.product(0, in_, 1, self.buffer)
        self.lhs.product(inScale, self.buffer, outScale, out)


class PMultigrid:
    """The p-multigrid of poisson2D/solver.cpp:111-232 (solveMultigrid) on the device (cie_hip_pmg_*)."""

    def __init__(self, matrix, mask, order):
        self.matrix, self.mask, self.order = matrix, mask, int(order)
        self.handle = c_void_p()
        check(library().cie_hip_pmg_create(matrix.handle, mask.handle, self.order, byref(self.handle)),
              "cie_hip_pmg_create")
        matrix.context._register(self)
        self.info = None

    def grid(self, index):
        """The matrix of grid `index` (-1: the renumbered coarse matrix; None if there is none), a view whose
        launch may be set or autotuned (cie_hip_pmg_grid). The masked grids are compacted to their kept rows and
        columns (grid 0 is the coarse matrix), and the finest is the plan's own matrix."""
        handle = c_void_p()
        check(library().cie_hip_pmg_grid(self.handle, int(index), byref(handle)), "cie_hip_pmg_grid")
        return None if not handle.value else CSRMatrix._borrow(handle, self.matrix.context, self)

    def solve(self, b, x, relative_tolerance=1e-6, max_cycles=10000):
        """Cycles from x (expected zero, as in the driver) until ||r|| <= relative_tolerance ||b||."""
        info = _lib.PmgInfo()
        _sync_torch()
        status = library().cie_hip_pmg_solve(self.handle, _dptr(b), _dptr(x), x.numel(), relative_tolerance,
                                             max_cycles, byref(info))
        self.info = info
        check(status, "cie_hip_pmg_solve")
        return info

    def close(self):
        if self.handle:
            library().cie_hip_pmg_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_REORDERINGS = {"none": _lib.REORDER_NONE, "cuthill-mckee": _lib.REORDER_CUTHILL_MCKEE,
                "reverse-cuthill-mckee": _lib.REORDER_REVERSE_CUTHILL_MCKEE}


def makeReordering(csr, strategy="cuthill-mckee"):
    """map[new] = old for a host CSR pattern (makeReordering, reorder.cpp:23-173; the strategy names are the
    driver's --reordering values, poisson2D/solver.cpp:490-493)."""
    if strategy not in _REORDERINGS:
        raise ValueError(f"unknown reordering strategy: {strategy}")
    extents = np.ascontiguousarray(csr.row_extents, dtype=np.int32)
    columns = np.ascontiguousarray(csr.column_indices, dtype=np.int32)
    mapping = np.empty(int(csr.rows), dtype=np.int32)
    check(library().cie_hip_make_reordering(int(csr.rows), extents.ctypes.data_as(c_void_p),
                                            columns.ctypes.data_as(c_void_p), _REORDERINGS[strategy],
                                            mapping.ctypes.data_as(c_void_p)), "cie_hip_make_reordering")
    return mapping


def reverseReorder(mapping):
    """The inverse permutation (reverseReorder(map, buffer), reorder.cpp:176-198), on the host."""
    mapping = np.ascontiguousarray(mapping, dtype=np.int32)
    inverse = np.empty_like(mapping)
    inverse[mapping] = np.arange(mapping.size, dtype=np.int32)
    return inverse


def reorder(mapping, vector, out=None, reverse=False, context=None):
    """out[i] = vector[mapping[i]] (reverse: out[mapping[i]] = vector[i]) on device vectors
    (reorder / reverseReorder of an array, reorder.cpp:368-426). mapping is a device int32 tensor."""
    torch = _torch()
    context = context or Context.default()
    if mapping.dtype != torch.int32 or mapping.numel() != vector.numel() or not mapping.is_contiguous():
        raise ValueError("mapping must be a contiguous int32 device tensor of the vector's size")
    if out is None:
        out = torch.empty_like(vector)
    suffix = {torch.float64: "f64", torch.float32: "f32"}.get(vector.dtype)
    if suffix is None or out.dtype != vector.dtype or out.numel() != vector.numel():
        raise TypeError("reorder works on float64 or float32 vectors of equal size")
    _sync_torch()
    check(getattr(library(), f"cie_hip_reorder_{suffix}")(context.handle, _dptr(mapping), 1 if reverse else 0,
                                                          _dptr(vector), _dptr(out), vector.numel()),
          f"cie_hip_reorder_{suffix}")
    return out


def elementStiffness(moments, jacobians, modulus=None, out=None, context=None):
    """K_e = |det J_e| k_e sum_ij (J_e^-1 J_e^-T)_ij moments[i, j] for affine cells (cie_hip_element_stiffness).

    moments: (dim, dim, npe, npe) device tensor; jacobians: (elements, dim, dim); modulus: (elements,) or None.
    Returns (elements, npe, npe)."""
    torch = _torch()
    context = context or Context.default()
    dim, npe = moments.shape[0], moments.shape[2]
    elements = jacobians.shape[0]
    for t in (moments, jacobians) + ((modulus,) if modulus is not None else ()):
        if t.dtype != torch.float64 or not t.is_contiguous() or t.device.type != "cuda":
            raise ValueError("inputs must be contiguous float64 device tensors")
    if tuple(moments.shape) != (dim, dim, npe, npe) or tuple(jacobians.shape) != (elements, dim, dim):
        raise ValueError("moments must be (dim, dim, npe, npe) and jacobians (elements, dim, dim)")
    if modulus is not None and modulus.numel() != elements:
        raise ValueError("modulus needs one value per element")
    if out is None:
        out = torch.empty((elements, npe, npe), dtype=torch.float64, device=jacobians.device)
    _sync_torch()
    check(library().cie_hip_element_stiffness(context.handle, elements, dim, npe, _dptr(moments), _dptr(jacobians),
                                              None if modulus is None else _dptr(modulus), _dptr(out)),
          "cie_hip_element_stiffness")
    return out


class Assembly:
    """K = sum_e scale[e] P_e^T K_e0 P_e on the pattern of `matrix`, assembled on the device
    (cie_hip_assembly; the element scatter of fem/packages/graph/impl/Assembler_impl.hpp:156-194)."""

    def __init__(self, matrix, dofs):
        self.matrix = matrix
        dofs = np.ascontiguousarray(dofs, dtype=np.int32)
        if dofs.ndim != 2:
            raise ValueError("dofs must be elements x nodes per element")
        self.elements, self.nodes = dofs.shape
        self.handle = c_void_p()
        check(library().cie_hip_assembly_create(matrix.handle, self.elements, self.nodes,
                                                dofs.ctypes.data_as(c_void_p), byref(self.handle)),
              "cie_hip_assembly_create")
        matrix.context._register(self)

    def assemble(self, element_matrices, scale=None):
        """Overwrite the matrix values with sum_e scale[e] K_e. element_matrices (device tensor) is one
        nodes x nodes matrix shared by all elements or elements x nodes x nodes; scale (elements) or None."""
        self._check_values(element_matrices)
        shared = element_matrices.numel() == self.nodes * self.nodes and element_matrices.dim() <= 2
        if not shared and element_matrices.numel() != self.elements * self.nodes * self.nodes:
            raise ValueError("element matrices have the wrong size")
        self._check_scale(scale)
        _sync_torch()
        check(library().cie_hip_assemble(self.handle, _dptr(element_matrices), 0 if shared else 1,
                                         None if scale is None else _dptr(scale)), "cie_hip_assemble")

    def assemble_vector(self, element_vectors, scale=None, out=None):
        """sum_e scale[e] P_e^T v_e as a device vector; element_vectors: one nodes vector or elements x nodes."""
        torch = _torch()
        self._check_values(element_vectors)
        shared = element_vectors.numel() == self.nodes and element_vectors.dim() == 1
        if not shared and element_vectors.numel() != self.elements * self.nodes:
            raise ValueError("element vectors have the wrong size")
        self._check_scale(scale)
        if out is None:
            out = torch.empty(self.matrix.rows, dtype=torch.float64, device=element_vectors.device)
        elif out.numel() != self.matrix.rows or out.dtype != torch.float64 or not out.is_contiguous():
            raise ValueError("out must be a contiguous float64 vector of one value per row")
        _sync_torch()
        check(library().cie_hip_assemble_vector(self.handle, _dptr(element_vectors), 0 if shared else 1,
                                                None if scale is None else _dptr(scale), _dptr(out)),
              "cie_hip_assemble_vector")
        return out

    @staticmethod
    def _check_values(values):
        if values.dtype != _torch().float64 or not values.is_contiguous() or values.device.type != "cuda":
            raise ValueError("element values must be a contiguous float64 device tensor")

    def _check_scale(self, scale):
        if scale is not None and (scale.numel() != self.elements or scale.dtype != _torch().float64
                                  or not scale.is_contiguous() or scale.device.type != "cuda"):
            raise ValueError("scale needs one contiguous float64 value per element")

    def close(self):
        if self.handle:
            library().cie_hip_assembly_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many independent systems on the diagonal blocks of one CSRMatrix (cie_hip_batch)."""

    def __init__(self, matrix, offsets):
        self.matrix = matrix
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.systems = len(self.offsets) - 1
        self.handle = c_void_p()
        check(library().cie_hip_batch_create(matrix.context.handle, matrix.handle, self.systems,
                                             self.offsets.ctypes.data_as(c_void_p), byref(self.handle)),
              "cie_hip_batch_create")
        matrix.context._register(self)

    def autotune(self, precision=32, repetitions=3):
        check(library().cie_hip_batch_autotune(self.handle, precision, repetitions), "cie_hip_batch_autotune")

    def column_blocks(self):
        """3 when the batch's tiles hold 1 x 3 column blocks, 0 when they hold matrix entries."""
        width = c_int()
        check(library().cie_hip_batch_get_column_blocks(self.handle, byref(width)), "cie_hip_batch_get_column_blocks")
        return width.value

    def set_column_blocks(self, width):
        check(library().cie_hip_batch_set_column_blocks(self.handle, int(width)), "cie_hip_batch_set_column_blocks")

    def tile_kernel(self):
        """Tile format of the batch's products: "tiled" (entries), "blocks" or "sorted" (1 x 3 column blocks),
        "paired" (1 x 5 and 1 x 3 column blocks)."""
        kernel = c_int()
        check(library().cie_hip_batch_get_tile_kernel(self.handle, byref(kernel)), "cie_hip_batch_get_tile_kernel")
        return _KERNEL_NAMES[kernel.value]

    def set_tile_kernel(self, kernel):
        check(library().cie_hip_batch_set_tile_kernel(self.handle, _kernel_code(kernel)), "cie_hip_batch_set_tile_kernel")

    def tile_rows(self):
        out = np.empty(self.systems, dtype=np.int32)
        check(library().cie_hip_batch_get_tile_rows(self.handle, out.ctypes.data_as(c_void_p)))
        return out

    def set_tile_rows(self, rows):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        check(library().cie_hip_batch_set_tile_rows(self.handle, rows.ctypes.data_as(c_void_p)))

    def solve(self, b, x, settings, mixed=True, graph=True, refinement_factor=_lib.DEFAULT_REFINEMENT_FACTOR,
              restart=False, row_sum_correction=True):
        """Solve every system in place on x; returns (stats, status, refinements, summary)."""
        stats = (Statistics * self.systems)()
        status = np.zeros(self.systems, dtype=np.int32)
        refinements = np.zeros(self.systems, dtype=np.uint64)
        summary = SolveInfo()
        mixed_settings = MixedSettings(settings, refinement_factor, 1 if restart else 0,
                                       0 if row_sum_correction else _lib.MIXED_PLAIN_FP32)
        _sync_torch()
        code = library().cie_hip_pcg_batched(self.matrix.context.handle, self.handle, _dptr(b), _dptr(x),
                                             byref(mixed_settings), 1 if mixed else 0, 1 if graph else 0, stats,
                                             status.ctypes.data_as(c_void_p), refinements.ctypes.data_as(c_void_p),
                                             byref(summary))
        check(code, "cie_hip_pcg_batched")
        return list(stats), status, refinements, summary

    def close(self):
        if self.handle:
            library().cie_hip_batch_destroy(self.handle)
            self.handle = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""Structured CSR systems and seeded inputs of the solver workloads (host side).

Thin numpy wrappers over include/cie_hip_systems.h. Node numbering is
lexicographic with x fastest and Dirichlet boundary rows/columns eliminated, so
each row's column indices are sorted (as the reference's assembler produces,
fem/packages/graph/src/Assembler.cpp:131-142).
"""
from ctypes import byref, c_int32, c_int64
from dataclasses import dataclass, field
import math

import numpy as np

from ._lib import LagrangeSpec, check_sys, library


@dataclass
class CSR:
    """Host CSR arrays in the layout of CSRView (linalg/packages/utilities/inc/CSRView.hpp:11-86)."""

    rows: int
    columns: int
    row_extents: np.ndarray
    column_indices: np.ndarray
    entries: np.ndarray

    @property
    def nnz(self):
        return int(self.row_extents[-1])

    def to_scipy(self):
        import scipy.sparse

        return scipy.sparse.csr_matrix((self.entries, self.column_indices, self.row_extents),
                                       shape=(self.rows, self.columns))


def _ptr(array):
    return array.ctypes.data_as(__import__("ctypes").c_void_p)


def splitmix64(seed, count):
    out = np.empty(count, dtype=np.uint64)
    check_sys(library().cie_sys_splitmix64(seed, count, _ptr(out)), "splitmix64")
    return out


def uniform(seed, count):
    """U(-1, 1): 2 * (x >> 11) * 2^-53 - 1 over the SplitMix64 stream of `seed`."""
    out = np.empty(count, dtype=np.float64)
    check_sys(library().cie_sys_uniform(seed, count, _ptr(out)), "uniform")
    return out


def normal(seed, sigma, count):
    """N(0, sigma^2) by Box-Muller, two SplitMix64 draws per value."""
    out = np.empty(count, dtype=np.float64)
    check_sys(library().cie_sys_normal(seed, float(sigma), count, _ptr(out)), "normal")
    return out


def poisson5(m):
    """2D 5-point Laplacian (4 on the diagonal, -1 off it) on an m x m interior grid."""
    rows, nnz = c_int64(), c_int64()
    lib = library()
    check_sys(lib.cie_sys_poisson5_size(m, byref(rows), byref(nnz)), "poisson5_size")
    extents = np.empty(rows.value + 1, dtype=np.int32)
    cols = np.empty(nnz.value, dtype=np.int32)
    vals = np.empty(nnz.value, dtype=np.float64)
    check_sys(lib.cie_sys_poisson5_fill(m, _ptr(extents), _ptr(cols), _ptr(vals)), "poisson5_fill")
    return CSR(rows.value, rows.value, extents, cols, vals)


BASES = {"lagrange": 0, "legendre": 1}


def _spec(dim, order, elements, eliminate=True, basis="lagrange"):
    return LagrangeSpec(dim, order, elements, 1 if eliminate else 0, BASES[basis])


def lagrange_size(dim, order, elements, eliminate=True):
    spec = _spec(dim, order, elements, eliminate)
    rows, nnz, count, npe = c_int64(), c_int64(), c_int64(), c_int32()
    check_sys(library().cie_sys_lagrange_size(byref(spec), byref(rows), byref(nnz), byref(count), byref(npe)),
              "lagrange_size")
    return rows.value, nnz.value, count.value, npe.value


def lagrange(dim, order, elements, conductivity=None, eliminate=True, threads=0, basis="lagrange"):
    """Laplace stiffness of order-`order` tensor-product elements on an elements^dim unit mesh
    (basis "lagrange": nodal, "legendre": integrated Legendre, hierarchical)."""
    spec = _spec(dim, order, elements, eliminate, basis)
    rows, nnz, count, _ = lagrange_size(dim, order, elements, eliminate)
    if conductivity is not None:
        conductivity = np.ascontiguousarray(conductivity, dtype=np.float64)
        if conductivity.shape != (count,):
            raise ValueError(f"expected {count} element conductivities, got {conductivity.shape}")
    extents = np.empty(rows + 1, dtype=np.int32)
    cols = np.empty(nnz, dtype=np.int32)
    vals = np.empty(nnz, dtype=np.float64)
    check_sys(library().cie_sys_lagrange_fill(byref(spec), None if conductivity is None else _ptr(conductivity),
                                              _ptr(extents), _ptr(cols), _ptr(vals), threads), "lagrange_fill")
    return CSR(rows, rows, extents, cols, vals)


def element_matrix(dim, order, elements, basis="lagrange"):
    """Element stiffness for k_e = 1 (nodes per element squared, row-major)."""
    _, _, _, npe = lagrange_size(dim, order, elements)
    out = np.empty((npe, npe), dtype=np.float64)
    check_sys(library().cie_sys_lagrange_element_matrix(byref(_spec(dim, order, elements, basis=basis)),
                                                        _ptr(out)), "element_matrix")
    return out


def gradient_moments(dim, order, basis="lagrange"):
    """moments[i, j] = sum_q w_q grad_i N grad_j N^T on the reference element [-1, 1]^dim (dim x dim x npe x npe)."""
    npe = (order + 1) ** dim
    out = np.empty((dim, dim, npe, npe), dtype=np.float64)
    check_sys(library().cie_sys_lagrange_gradient_moments(byref(_spec(dim, order, 2, basis=basis)), _ptr(out)),
              "gradient_moments")
    return out


def ansatz_mask(dim, order, elements, eliminate=True):
    """Per-row ansatz mask of the integrated Legendre basis (largest 1D degree, at least 1), uint16."""
    rows = lagrange_size(dim, order, elements, eliminate)[0]
    out = np.empty(rows, dtype=np.uint16)
    check_sys(library().cie_sys_lagrange_ansatz_mask(byref(_spec(dim, order, elements, eliminate, "legendre")),
                                                     _ptr(out)), "ansatz_mask")
    return out


def element_dofs(dim, order, elements, eliminate=True):
    """Row index of each element node, -1 for eliminated (Dirichlet) nodes."""
    _, _, count, npe = lagrange_size(dim, order, elements, eliminate)
    out = np.empty((count, npe), dtype=np.int32)
    check_sys(library().cie_sys_lagrange_element_dofs(byref(_spec(dim, order, elements, eliminate)), _ptr(out)),
              "element_dofs")
    return out


def csr_slots(csr, rows, cols):
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    out = np.empty(rows.shape, dtype=np.int64)
    check_sys(library().cie_sys_csr_slots(csr.rows, _ptr(csr.row_extents), _ptr(csr.column_indices), rows.size,
                                          _ptr(rows), _ptr(cols), _ptr(out)), "csr_slots")
    return out


# ---------------------------------------------------------------------------
# Workloads
# ---------------------------------------------------------------------------

@dataclass
class Workload:
    name: str
    matrix: CSR
    rhs: np.ndarray
    relative_tolerance: float
    description: str
    extra: dict = field(default_factory=dict)


def five_point_2d(m=256, seed=0, tolerance=1e-8):
    """2D 5-point Laplacian on an m x m interior grid, b ~ U(-1,1) from `seed`."""
    matrix = poisson5(m)
    return Workload(f"five_point_2d_{m}", matrix, uniform(seed, matrix.rows), tolerance,
                    f"2D 5-point Laplacian, {m}x{m} interior grid, b~U(-1,1) seed {seed}")


def q2_laplace_3d(elements=64, seed=1, tolerance=1e-10, threads=0):
    """Q2 hexahedra on an elements^3 unit-cube mesh, k_e = 1, b ~ U(-1,1) from `seed`."""
    matrix = lagrange(3, 2, elements, threads=threads)
    return Workload(f"q2_laplace_3d_{elements}", matrix, uniform(seed, matrix.rows), tolerance,
                    f"3D Q2 Laplace stiffness, {elements}^3 hexahedra, Dirichlet eliminated, b~U(-1,1) seed {seed}")


def q4_lognormal_3d(elements=32, seed=2, rhs_seed=1, tolerance=1e-10, threads=0):
    """Q4 hexahedra with k_e = exp(g_e), g_e ~ N(0,1) from `seed`; b ~ U(-1,1) from `rhs_seed`."""
    _, _, count, _ = lagrange_size(3, 4, elements)
    g = normal(seed, 1.0, count)
    matrix = lagrange(3, 4, elements, conductivity=np.exp(g), threads=threads)
    return Workload(f"q4_lognormal_3d_{elements}", matrix, uniform(rhs_seed, matrix.rows), tolerance,
                    f"3D Q4 Laplace, {elements}^3 hexahedra, k_e=exp(N(0,1)) seed {seed}, b~U(-1,1) seed {rhs_seed}",
                    {"log_conductivity": g})


def q2_adjoint_3d(elements=48, seed=4, target_seed=5, rhs_seed=1, variance=0.5, tolerance=1e-11, threads=0):
    """Q2 hexahedra with k_e = exp(g_e), g_e ~ N(0, variance) from `seed`; the target
    field uses `target_seed`; b ~ U(-1,1) from `rhs_seed`."""
    _, _, count, _ = lagrange_size(3, 2, elements)
    sigma = math.sqrt(variance)
    g = normal(seed, sigma, count)
    g_target = normal(target_seed, sigma, count)
    matrix = lagrange(3, 2, elements, conductivity=np.exp(g), threads=threads)
    return Workload(f"q2_adjoint_3d_{elements}", matrix, uniform(rhs_seed, matrix.rows), tolerance,
                    f"3D Q2 Laplace, {elements}^3 hexahedra, k_e=exp(N(0,{variance})) seed {seed}",
                    {"log_conductivity": g, "target_log_conductivity": g_target})


def q2_quad_batch(count=256, seed=3, smallest=16, largest=256, sigma=0.5, tolerance=1e-10, threads=0):
    """`count` independent 2D Q2 Laplace systems on n_i x n_i element grids.

    n_i = smallest + floor(u_i (largest - smallest + 1)) for the first `count`
    SplitMix64 uniforms u_i in [0, 1) of `seed`; system i has k_e = exp(g_e)
    with g_e ~ N(0, sigma^2) from seed 1000 * seed + i, and b ~ U(-1, 1) from
    seed 1000 * seed + 500 + i. Returned as one block-diagonal CSR (system i
    owns rows offsets[i]..offsets[i+1]) with the concatenated right-hand side.
    """
    raw = splitmix64(seed, count)
    unit = (raw >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    sizes = (smallest + np.floor(unit * (largest - smallest + 1))).astype(np.int64)
    blocks, rhs = [], []
    for i, n in enumerate(sizes):
        _, _, elements, _ = lagrange_size(2, 2, int(n))
        g = normal(1000 * seed + i, sigma, elements)
        blocks.append(lagrange(2, 2, int(n), conductivity=np.exp(g), threads=threads))
        rhs.append(uniform(1000 * seed + 500 + i, blocks[-1].rows))
    matrix, offsets = block_diagonal(blocks)
    return Workload(f"q2_quad_batch_{count}", matrix, np.concatenate(rhs), tolerance,
                    f"{count} 2D Q2 Laplace systems, n_i~U{{{smallest}..{largest}}} seed {seed}, "
                    f"k_e lognormal(0,{sigma})", {"offsets": offsets, "sizes": sizes})


def block_diagonal(blocks):
    """One CSR holding `blocks` on its diagonal; returns (csr, row offsets)."""
    offsets = np.zeros(len(blocks) + 1, dtype=np.int64)
    offsets[1:] = np.cumsum([b.rows for b in blocks])
    nnz = np.zeros(len(blocks) + 1, dtype=np.int64)
    nnz[1:] = np.cumsum([b.nnz for b in blocks])
    if nnz[-1] > np.iinfo(np.int32).max:
        raise ValueError("the batch exceeds int32 nonzero indexing")
    extents = np.empty(offsets[-1] + 1, dtype=np.int32)
    cols = np.empty(nnz[-1], dtype=np.int32)
    vals = np.empty(nnz[-1], dtype=np.float64)
    for i, b in enumerate(blocks):
        extents[offsets[i]:offsets[i + 1]] = b.row_extents[:-1] + nnz[i]
        cols[nnz[i]:nnz[i + 1]] = b.column_indices + offsets[i]
        vals[nnz[i]:nnz[i + 1]] = b.entries
    extents[-1] = nnz[-1]
    return CSR(int(offsets[-1]), int(offsets[-1]), extents, cols, vals), offsets

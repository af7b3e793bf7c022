"""The FP32 plane's row-sum correction of the mixed solvers.

Rounding the stiffness values to FP32 perturbs each row sum by about eps32 * a_ii; FEM rows away from the
boundary sum to zero, so the perturbation falls on the smooth modes CG resolves last, and every residual
replacement of the refinement re-injects it (on 3D Q2 64^3: 430 inner iterations against 388 in FP64). The
inner product therefore adds c_i p_i with c_i = fl32(sum_j (a_ij - fl32(a_ij))): the FP32 operator then keeps
the FP64 row sums to FP32 rounding of c itself, in FP32 arithmetic and 4 bytes per row.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from cie_amd import systems
from cie_amd._lib import Statistics

pytestmark = pytest.mark.gpu


def to_scipy(csr):
    return sp.csr_matrix((csr.entries, csr.column_indices, csr.row_extents), shape=(csr.rows, csr.columns))


def expected_correction(csr):
    residue = csr.entries - csr.entries.astype(np.float32).astype(np.float64)  # exact (Sterbenz)
    rows = np.repeat(np.arange(csr.rows), np.diff(csr.row_extents))
    return np.bincount(rows, weights=residue, minlength=csr.rows)


@pytest.mark.parametrize("dim,order,n", [(3, 2, 6), (2, 4, 9), (3, 4, 3)])
def test_row_sum_correction_plane(hip, torch_gpu, dim, order, n):
    k = np.exp(systems.normal(2, 1.0, n ** dim))
    csr = systems.lagrange(dim, order, n, conductivity=k)
    matrix = hip.CSRMatrix(csr)
    got = matrix.row_sum_correction().cpu().numpy()
    exact = expected_correction(csr)
    # the device sums the residues in another order: equal up to the final FP32 rounding
    assert np.all(np.abs(got - exact) <= np.spacing(np.abs(exact).astype(np.float32)) + 1e-30)
    # the property: the FP32 plane plus the correction keeps the FP64 row sums (sums taken exactly enough in
    # FP64: every FP32 value and residue is a short binary fraction)
    a = to_scipy(csr)
    s64 = a @ np.ones(csr.rows)
    s32 = a.astype(np.float32).astype(np.float64) @ np.ones(csr.rows)
    magnitude = abs(a) @ np.ones(csr.rows)
    plain = np.abs(s32 - s64)
    corrected = np.abs(s32 + got.astype(np.float64) - s64)
    assert np.all(corrected <= np.spacing(np.abs(got)) + 8 * np.finfo(np.float64).eps * magnitude)
    assert corrected.max() < 0.05 * plain.max()


def test_mixed_iterations_with_and_without_correction(hip, torch_gpu):
    """3D Q2 32^3 (250,047 rows), refinement factor 0.1: with the correction the mixed solve needs about the
    FP64 solve's iteration count (a CPU model of the same arithmetic gives 200 against 198 loop trips; measured
    196 FP64, 201 mixed), without it about 11% more (220)."""
    w = systems.q2_laplace_3d(32)
    matrix = hip.CSRMatrix(w.matrix)
    b = torch_gpu.from_numpy(w.rhs).to("cuda:0")
    runs = {}
    for label, solver in (("fp64", hip.FusedPCG(matrix, Statistics(20000, 0.0, 1e-10))),
                          ("corrected", hip.MixedPCG(matrix, Statistics(20000, 0.0, 1e-10), refinement_factor=0.1)),
                          ("plain", hip.MixedPCG(matrix, Statistics(20000, 0.0, 1e-10), refinement_factor=0.1,
                                                 row_sum_correction=False))):
        x = torch_gpu.zeros_like(b)
        solver.product(0, b, 1, x)
        assert solver.info.converged == 1 and solver.info.trueRelativeResidual < (2e-10 if label == "fp64" else 1e-10)
        runs[label] = int(solver.getStats().iterationCount)
    print(runs)
    assert runs["corrected"] <= runs["fp64"] + 6
    assert runs["plain"] > runs["corrected"] + 10

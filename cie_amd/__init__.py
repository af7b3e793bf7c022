"""MI355X (gfx950) HIP backend for CiE's sparse linear-solve path.

    cie_amd.linalg   -- HIPSpace, HIPCSROperator, DiagonalOperator, ConjugateGradients,
                        FusedPCG, MixedPCG (mirrors linalg/packages/solvers)
    cie_amd.systems  -- structured CSR systems and seeded inputs
    cie_amd._lib     -- ctypes binding of libcie_hip.so (include/cie_hip.h)
"""
from ._lib import LIBRARY_PATH, library  # noqa: F401

__all__ = ["library", "LIBRARY_PATH"]

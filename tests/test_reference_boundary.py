"""The drop-in boundary against the reference's own headers (CPU, compile-only).

include/cie_hip/linalg.hpp is compiled inside the reference tree with
CIE_HIP_WITH_REFERENCE_HEADERS: HIPSpace<double|float> must satisfy the
reference's LinalgSpaceLike (linalg/packages/solvers/inc/LinalgSpace.hpp:10-35),
the HIP operators must derive from its LinearOperator
(inc/LinearOperator.hpp:10-34) and override its product, and its own
ConjugateGradients template (inc/ConjugateGradients.hpp:15-46) must accept the
HIP space in the driver's construction sequence (poisson2D/solver.cpp:238-288).
Skipped where the reference tree is absent.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
REFERENCE = Path("/root/reference/libraries")
SOURCE = ROOT / "tests" / "cpp" / "reference_headers_check.cpp"

needs_reference = pytest.mark.skipif(not (REFERENCE / "linalg/packages/solvers/inc/LinalgSpace.hpp").exists(),
                                     reason="the reference tree is not present")


def compile_check(*defines):
    command = ["g++", "-std=c++20", "-fsyntax-only", *defines, f"-I{ROOT / 'include'}",
               f"-I{REFERENCE / 'linalg'}", f"-I{REFERENCE / 'cieutils'}", str(SOURCE)]
    return subprocess.run(command, capture_output=True, text=True, timeout=300)


@needs_reference
@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is not installed")
def test_hip_face_compiles_against_reference_headers():
    result = compile_check("-DCIE_HIP_WITH_REFERENCE_HEADERS")
    assert result.returncode == 0, result.stderr


@needs_reference
@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is not installed")
def test_check_refuses_the_standalone_build():
    """Without the define the check must not pass silently on the header's own restated interfaces."""
    result = compile_check()
    assert result.returncode != 0 and "CIE_HIP_WITH_REFERENCE_HEADERS" in result.stderr

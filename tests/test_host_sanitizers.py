"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer and under ThreadSanitizer (CPU only).

The reference's CI builds and runs its tests with -fsanitize=address, thread
and undefined (linalg/.github/workflows/github-actions.yml:21-55). The same is
done here for this build's threaded host code: the CPU oracle's pthread pool
(oracle/oracle.c), the structured-system generator (cie_amd/csrc/systems.cpp)
and the MatrixMarket writer and reader (cie_amd/csrc/matrix_market.cpp). They
are compiled from source with g++/gcc (host code only: no HIP file is built
with a sanitizer) into tests/cpp/host_sanitize.cpp, whose checks must all pass
and whose sanitizers must stay silent.
"""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
DRIVER = ROOT / "tests" / "cpp" / "host_sanitize.cpp"
HOST_SOURCES = [ROOT / "cie_amd" / "csrc" / "systems.cpp", ROOT / "cie_amd" / "csrc" / "matrix_market.cpp"]
ORACLE = ROOT / "oracle" / "oracle.c"

VARIANTS = {
    "address,undefined": ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"],
    "thread": ["-fsanitize=thread"],
}


def run(command, **kwargs):
    result = subprocess.run(command, capture_output=True, text=True, timeout=600, **kwargs)
    assert result.returncode == 0, " ".join(map(str, command)) + "\n" + result.stdout + result.stderr
    return result


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="gcc/g++ not installed")
@pytest.mark.parametrize("variant", list(VARIANTS))
def test_host_code_under_sanitizer(tmp_path, variant):
    flags = ["-O1", "-g", *VARIANTS[variant]]
    objects = [tmp_path / "oracle.o"]
    run(["gcc", *flags, "-std=c11", "-D_POSIX_C_SOURCE=200809L", "-ffp-contract=off", "-pthread", "-c", ORACLE,
         "-o", objects[0]])
    for source in HOST_SOURCES:
        obj = tmp_path / (source.stem + ".o")
        run(["g++", *flags, "-std=c++17", "-pthread", "-c", source, "-o", obj])
        objects.append(obj)
    binary = tmp_path / "host_sanitize"
    run(["g++", *flags, "-std=c++17", "-pthread", DRIVER, *objects, "-o", binary, "-lm"])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=66",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", TSAN_OPTIONS="halt_on_error=1:exitcode=66")
    result = run([binary, tmp_path], env=env)
    assert "all checks passed" in result.stdout
    assert "Sanitizer" not in result.stderr, result.stderr

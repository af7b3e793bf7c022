"""Build recipe of libcie_hip.so (HIP for gfx950) -- no CMake, plain hipcc.

Every HIP translation unit is compiled to an object in parallel and linked
into ``cie_amd/libcie_hip.so``. ``build()`` always recompiles.
"""
import concurrent.futures
import os
import subprocess
import sys
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
ROOT = PACKAGE.parent
SOURCES = PACKAGE / "csrc"
OUTPUT = PACKAGE / "libcie_hip.so"
OBJECTS = PACKAGE / "_objects"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # Packed FP32 instructions stay off (only ordinary scalar/vector FP32 code).
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-pthread", "-Wall"]


def _run(command):
    result = subprocess.run(command, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(command) + "\n" + result.stdout + result.stderr)
    return result


def _compile(source: Path) -> Path:
    obj = OBJECTS / (source.name + ".o")
    if source.suffix == ".hip":
        command = [HIPCC, *HIP_FLAGS, "-c", str(source), "-o", str(obj)]
    else:
        command = ["g++", *HOST_FLAGS, "-c", str(source), "-o", str(obj)]
    _run(command)
    return obj


def build(verbose: bool = False) -> Path:
    OBJECTS.mkdir(exist_ok=True)
    sources = sorted(SOURCES.glob("*.hip")) + sorted(SOURCES.glob("*.cpp"))
    workers = max(1, min(len(sources), os.cpu_count() or 1))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        objects = list(pool.map(_compile, sources))
    tmp = OUTPUT.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", *map(str, objects), "-o", str(tmp)])
    os.replace(tmp, OUTPUT)
    if verbose:
        print(f"built {OUTPUT}", file=sys.stderr)
    return OUTPUT


if __name__ == "__main__":
    build(verbose=True)

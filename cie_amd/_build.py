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
    # No implicit multiply-add contraction: every product and sum rounds like the
    # CPU loops it mirrors; kernels that want FMA call fma() explicitly.
    "-ffp-contract=off",
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


def _compile(source: Path, objects: Path = OBJECTS, extra=()) -> Path:
    obj = objects / (source.name + ".o")
    if source.suffix == ".hip":
        command = [HIPCC, *HIP_FLAGS, *extra, "-c", str(source), "-o", str(obj)]
    else:
        command = ["g++", *HOST_FLAGS, "-c", str(source), "-o", str(obj)]
    _run(command)
    return obj


def build(verbose: bool = False, output: Path = OUTPUT, extra_flags=(), sources_dir: Path = SOURCES) -> Path:
    """Compile every source of `sources_dir` and link `output`; extra_flags go to each hipcc compile."""
    output = Path(output)
    objects_dir = OBJECTS if output == OUTPUT else output.parent / (output.stem + "_objects")
    objects_dir.mkdir(parents=True, exist_ok=True)
    sources = sorted(Path(sources_dir).glob("*.hip")) + sorted(Path(sources_dir).glob("*.cpp"))
    workers = max(1, min(len(sources), os.cpu_count() or 1))
    with concurrent.futures.ThreadPoolExecutor(workers) as pool:
        objects = list(pool.map(lambda s: _compile(s, objects_dir, tuple(extra_flags)), sources))
    tmp = output.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", *map(str, objects), "-o", str(tmp)])
    os.replace(tmp, output)
    if verbose:
        print(f"built {output}", file=sys.stderr)
    return output


def build_facade_test(verbose: bool = False) -> Path:
    """Compile tests/cpp/facade_test.cpp (the C++ face over the C ABI) against the library."""
    source = ROOT / "tests" / "cpp" / "facade_test.cpp"
    output = ROOT / "tests" / "cpp" / "build" / "facade_test"
    output.parent.mkdir(parents=True, exist_ok=True)
    _run(["g++", "-std=c++20", "-O2", "-Wall", f"-I{ROOT / 'include'}", str(source), f"-L{PACKAGE}", "-lcie_hip",
          "-Wl,-rpath,$ORIGIN/../../../cie_amd", "-o", str(output)])
    if verbose:
        print(f"built {output}", file=sys.stderr)
    return output


if __name__ == "__main__":
    build(verbose=True)

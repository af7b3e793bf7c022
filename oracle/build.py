"""Build recipe of the CPU oracle (test infrastructure only).

Compiles oracle.c with gcc into oracle/liboracle.so. FMA contraction is off
so the arithmetic matches a plain x86-64 build of the reference's loops
(separate multiply and add roundings).

The reference's own solver sources cannot be compiled in this image: they
include C++20 <format> (CSROperator.cpp:9, ConjugateGradients_impl.hpp:14,
DiagonalOperator.cpp:10), which libstdc++ 11 lacks, and linalg requires Eigen3;
reorder.cpp needs <format> and the tsl::robin_set library as well. For those
paths the restatement in oracle.c is the oracle.

One reference source does build with g++ alone: the MatrixMarket writer
(libraries/cieutils/packages/io/src/MatrixMarket.cpp, standard library only).
build_reference() compiles it where it lies in the reference tree, with the
C entry points of ref_matrix_market.cpp, into oracle/_ref/ (kept out of git,
shipped with the tree like the other built libraries).
"""
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUTPUT = HERE / "liboracle.so"


def build() -> Path:
    command = ["gcc", "-O3", "-std=c11", "-D_POSIX_C_SOURCE=200809L", "-ffp-contract=off", "-fPIC", "-shared",
               "-pthread", str(HERE / "oracle.c"), "-o", str(OUTPUT) + ".tmp", "-lm"]
    result = subprocess.run(command, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + result.stdout + result.stderr)
    Path(str(OUTPUT) + ".tmp").replace(OUTPUT)
    return OUTPUT


REFERENCE = Path("/root/reference")
REF_OUTPUT = HERE / "_ref" / "libref_matrix_market.so"


def build_reference() -> Path:
    """Builds oracle/_ref/libref_matrix_market.so from the reference tree (None when it is absent)."""
    source = REFERENCE / "libraries/cieutils/packages/io/src/MatrixMarket.cpp"
    if not source.exists():
        return None
    REF_OUTPUT.parent.mkdir(exist_ok=True)
    command = ["g++", "-O2", "-std=c++20", "-fPIC", "-shared", "-I", str(REFERENCE / "libraries/cieutils"),
               str(source), str(HERE / "ref_matrix_market.cpp"), "-o", str(REF_OUTPUT) + ".tmp"]
    result = subprocess.run(command, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError("reference MatrixMarket build failed:\n" + result.stdout + result.stderr)
    Path(str(REF_OUTPUT) + ".tmp").replace(REF_OUTPUT)
    return REF_OUTPUT


if __name__ == "__main__":
    print(build())
    print(build_reference())

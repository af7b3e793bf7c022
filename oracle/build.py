"""Build recipe of the CPU oracle (test infrastructure only).

Compiles oracle.c with gcc into oracle/liboracle.so. FMA contraction is off
so the arithmetic matches a plain x86-64 build of the reference's loops
(separate multiply and add roundings).

The reference's own solver sources cannot be compiled in this image: they
include C++20 <format> (CSROperator.cpp:9, ConjugateGradients_impl.hpp:14,
DiagonalOperator.cpp:10), which libstdc++ 11 lacks, and linalg requires Eigen3.
So there is no oracle/_ref build; the restatement in oracle.c is the oracle.
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


if __name__ == "__main__":
    print(build())

"""Writes the committed fixtures of tests/golden/ (run from the repository root).

rng.json -- the first 16 SplitMix64 outputs of each workload seed (0..5) and
            the derived U(-1,1) values, computed here in pure Python from the
            published SplitMix64 definition (independent of the C++ generator).
kat.json -- known-answer vectors that the reference's own SYCL space test
            holds (linalg/packages/solvers/test/SYCLSpace_test.cpp), written
            out as data: the dot-product sizes, the 2x2 SpMV case and its
            expected output [233, 462].
"""
import json
from pathlib import Path

HERE = Path(__file__).resolve().parent
MASK = (1 << 64) - 1


def splitmix64(seed, count):
    state = seed & MASK
    out = []
    for _ in range(count):
        state = (state + 0x9E3779B97F4A7C15) & MASK
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        out.append(z ^ (z >> 31))
    return out


def main():
    rng = {}
    for seed in range(6):
        raw = splitmix64(seed, 16)
        rng[str(seed)] = {"raw": [str(v) for v in raw], "uniform": [2.0 * ((v >> 11) * 2.0**-53) - 1.0 for v in raw]}
    (HERE / "rng.json").write_text(json.dumps(rng, indent=1) + "\n")
    kat = {
        "source": "linalg/packages/solvers/test/SYCLSpace_test.cpp",
        "inner_product_sizes": [11, 102, 1003, 10004, 100005, 1000006, 10000007],
        "assign_add_scale_sizes": [10, 100, 1000, 10000, 100000],
        "add": {"target_factor": 10.0, "first_scale": 100.0, "second_scale": 1.0},
        "scale_uniform": 2.0,
        "scale_componentwise": {"target_factor": 10.0, "scale": 100.0},
        "spmv": {"entries": [1.0, 2.0, 3.0, 4.0], "column_indices": [0, 1, 0, 1], "row_extents": [0, 2, 4],
                 "columns": 2, "in": [5.0, 6.0], "out": [7.0, 8.0], "in_scale": 9.0, "out_scale": 10.0,
                 "expected": [233.0, 462.0]},
        "masked_spmv": {"entries": [1.0, 2.0, 100.0, 3.0, 4.0], "column_indices": [0, 2, 1, 0, 2],
                        "row_extents": [0, 2, 3, 5], "mask": [1, 2, 1], "threshold": 2, "in": [5.0, 100.0, 6.0],
                        "out": [7.0, 1000.0, 8.0], "in_scale": 9.0, "out_scale": 10.0,
                        "expected": [233.0, 9000.0, 462.0]},
    }
    (HERE / "kat.json").write_text(json.dumps(kat, indent=1) + "\n")


if __name__ == "__main__":
    main()

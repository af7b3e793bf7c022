"""The C++ face (include/cie_hip/linalg.hpp) driven like the reference's
poisson2D driver, compared with the oracle: generic ConjugateGradients over
HIPSpace<double>, the fused FP64 solver and the mixed solver."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from cie_amd import systems

pytestmark = pytest.mark.gpu
BINARY = Path(__file__).resolve().parent / "cpp" / "build" / "facade_test"


def test_cpp_facade_matches_oracle(tmp_path, torch_gpu):
    assert BINARY.exists(), f"{BINARY} is not built (run __graft_entry__.build())"
    w = systems.five_point_2d(256)
    x_ref, s_ref, _, _ = oracle.pcg(w.matrix, w.rhs, relative_tolerance=1e-8)
    path = tmp_path / "system.bin"
    with open(path, "wb") as f:
        np.array([w.matrix.rows, w.matrix.nnz], dtype=np.int64).tofile(f)
        np.array([1e-8]).tofile(f)
        w.matrix.row_extents.astype(np.int32).tofile(f)
        w.matrix.column_indices.astype(np.int32).tofile(f)
        w.matrix.entries.astype(np.float64).tofile(f)
        w.rhs.astype(np.float64).tofile(f)
    result = subprocess.run([str(BINARY), str(path), str(tmp_path / "x")], capture_output=True, text=True, timeout=120)
    assert result.returncode == 0, result.stderr
    lines = {line.split()[0]: line.split()[1:] for line in result.stdout.strip().splitlines()}
    for name in ("generic", "fused", "mixed"):
        iterations, absolute, relative = int(lines[name][0]), float(lines[name][1]), float(lines[name][2])
        x = np.fromfile(tmp_path / f"x_{name}.bin")
        error = np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)
        assert relative < 1e-8, name
        if name == "mixed":
            assert error < 1e-6, (name, error)
        else:
            assert abs(iterations - int(s_ref.iterationCount)) <= 1, (name, iterations, s_ref.iterationCount)
            assert error < 1e-10, (name, error)
    assert lines["missing-diagonal"] == [str(6)]

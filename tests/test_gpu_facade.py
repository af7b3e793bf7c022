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
    # ConjugateGradients<HIPSpace<float>> (ConjugateGradients.cpp:10,16) to relative 1e-6
    iterations, relative = int(lines["generic32"][0]), float(lines["generic32"][2])
    assert relative < 1e-6 and iterations > 0
    x32 = np.fromfile(tmp_path / "x_generic32.bin", dtype=np.float32).astype(np.float64)
    assert_fp32_solution_bound(w, x32, x_ref)
    early = [line.split() for line in result.stdout.splitlines() if line.startswith("early ")]
    assert len(early) == 6 and all(fields[3] == "1" for fields in early), early
    check_reports_match_python(w, tmp_path)


def check_reports_match_python(w, tmp_path):
    """The C++ face's reports at verbosity 3 are the Python mirror's, character for character: both run the same
    deterministic kernels on the same system and format through cie_hip_iteration_report."""
    import torch

    from cie_amd import linalg
    from cie_amd._lib import Statistics

    matrix = linalg.CSRMatrix(w.matrix)
    space = linalg.HIPSpace()
    b = torch.from_numpy(w.rhs).to("cuda:0")
    solvers = {
        "generic": linalg.ConjugateGradients(linalg.HIPCSROperator(matrix, space), space,
                                             linalg.makeDiagonalOperator(matrix, space), Statistics(20000, 0.0, 1e-8),
                                             verbosity=3),
        "fused": linalg.FusedPCG(matrix, Statistics(20000, 0.0, 1e-8), verbosity=3),
    }
    for name, solver in solvers.items():
        x = torch.zeros_like(b)
        solver.product(0, b, 1, x)
        text = (tmp_path / f"x_reports_{name}.txt").read_text()
        assert len(solver.reports) == int(solver.getStats().iterationCount) + 1, name
        assert text == "".join(report + "\n" for report in solver.reports), name


def five_point_condition(m):
    """2-norm condition number of the 5-point Dirichlet Laplacian on an m x m interior grid (closed form):
    eigenvalues 4 - 2cos(j pi/(m+1)) - 2cos(k pi/(m+1)), so cond = cot^2(pi / (2(m+1)))."""
    return 1.0 / np.tan(np.pi / (2 * (m + 1))) ** 2


def assert_fp32_solution_bound(w, x32, x_ref):
    """The FP32 solution is bounded against the FP64 oracle by its own FP64 true residual:
    ||x32 - x*|| / ||x*|| <= cond(A) * ||b - A x32|| / ||b||, and the oracle is within cond(A) * 1e-8."""
    r = oracle.csr_product(w.matrix, 1.0, x32, -1.0, w.rhs.copy())
    true32 = np.linalg.norm(r) / np.linalg.norm(w.rhs)
    cond = five_point_condition(256)
    assert true32 < cond * np.finfo(np.float32).eps, true32  # FP32 attainable accuracy is about eps32 * cond
    error = np.linalg.norm(x32 - x_ref) / np.linalg.norm(x_ref)
    assert error <= cond * (true32 + 1e-8), (error, true32)
    return error, true32

"""The driver's --write-linear-system files (poisson2D/solver.cpp:518-566) read
back into a device matrix and solved: with 17 digits the system, and so the
solve, is bitwise the one written."""
import time

import numpy as np
import pytest

from cie_amd import systems
from cie_amd._lib import Statistics
from cie_amd.matrix_market import MatrixMarket

pytestmark = pytest.mark.gpu


def test_written_system_solves_to_the_same_solution(hip, torch_gpu, tmp_path):
    workload = systems.q4_lognormal_3d(6)
    lhs, rhs = workload.matrix, workload.rhs
    t0 = time.perf_counter()
    MatrixMarket.Output(tmp_path / "lhs.mm", digits=17).write_csr(lhs)
    MatrixMarket.Output(tmp_path / "rhs.mm", digits=17)(rhs)
    write_seconds = time.perf_counter() - t0
    t0 = time.perf_counter()
    lhs_read = MatrixMarket.Input(tmp_path / "lhs.mm").read()
    rhs_read = MatrixMarket.Input(tmp_path / "rhs.mm").read()
    read_seconds = time.perf_counter() - t0
    assert np.array_equal(lhs_read.entries, lhs.entries) and np.array_equal(rhs_read, rhs)
    settings = Statistics(20000, 0.0, 1e-10)
    solutions = []
    for csr, b in ((lhs, rhs), (lhs_read, rhs_read)):
        matrix = hip.CSRMatrix(csr)
        x = torch_gpu.zeros(csr.rows, dtype=torch_gpu.float64, device="cuda:0")
        hip.FusedPCG(matrix, settings).product(0, torch_gpu.from_numpy(b).cuda(), 1, x)
        solutions.append(x.cpu().numpy())
    assert np.array_equal(solutions[0], solutions[1])
    # the solution file in the reference's format (6 significant digits)
    MatrixMarket.Output(tmp_path / "solution.mm")(solutions[1])
    back = MatrixMarket.Input(tmp_path / "solution.mm").read()
    assert np.allclose(back, solutions[1], rtol=5e-6, atol=0)
    print(f"\n{lhs.nnz} nonzeros: write {write_seconds * 1e3:.0f} ms, read {read_seconds * 1e3:.0f} ms (17 digits)")

"""Per-iteration reports of the HIP solvers (ConjugateGradients_impl.hpp:100-111 and :175-182 call
IterativeSolver::makeIterationReport, linalg/packages/solvers/src/IterativeSolver.cpp:17-58).

The fused solvers record each loop trip's residual on the device and log the reports after the solve; the
generic CG logs them as it goes. The recorded residuals follow the oracle's history (the reference CG's
residual after each trip), and the reports' layout follows the reference's rules: the header at iteration 0,
one row per trip at verbosity 3, only the converged row at verbosity 2, the closing rule once converged, a
single Termination report when the start check ends the solve.
"""
import ctypes
import re

import numpy as np
import pytest

import oracle
from cie_amd import _lib, systems
from cie_amd._lib import Statistics

pytestmark = pytest.mark.gpu

RULE = "+ --------- + ----------------- + ----------------- +"
HEADER = RULE + "\n| iteration | absolute residual | relative residual |\n" + RULE
ROW = re.compile(r"^\| \x1b\[[0-9;]*m +(\d+)\x1b\[0m \| \x1b\[[0-9;]*m +(\S+)\x1b\[0m \| \x1b\[[0-9;]*m +(\S+)\x1b\[0m \|$")


def rows_of(reports):
    """(iteration, absolute cell, relative cell) of every row line, in order."""
    found = []
    for text in reports:
        for line in text.splitlines():
            match = ROW.match(line)
            if match:
                found.append((int(match.group(1)), match.group(2), match.group(3)))
    return found


@pytest.fixture(scope="module")
def five_point(hip):
    w = systems.five_point_2d(256)
    x, stats, hist, _ = oracle.pcg(w.matrix, w.rhs, relative_tolerance=w.relative_tolerance, history=True)
    return w, hip.CSRMatrix(w.matrix), stats, hist


def to_dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda:0")


def check_table(solver, rows, converged=True, closings=None):
    """Header once, rows 0 .. rows-1 in order, the closing rule after every converged row (`closings`: their
    number, 1 by default for a converged solve)."""
    reports = solver.reports
    assert reports[0].startswith(HEADER)
    assert reports[-1].endswith(RULE) == converged
    closings = (1 if converged else 0) if closings is None else closings
    assert sum(text.count(RULE) for text in reports) == 2 + closings
    table = rows_of(reports)
    assert [row[0] for row in table] == list(range(rows))
    return table


def test_fused_reports_follow_the_oracle_history(hip, torch_gpu, five_point):
    w, matrix, s_ref, h_ref = five_point
    solver = hip.FusedPCG(matrix, Statistics(5000, 0.0, w.relative_tolerance), verbosity=3)
    x = torch_gpu.zeros(w.matrix.rows, dtype=torch_gpu.float64, device="cuda:0")
    solver.product(0, to_dev(torch_gpu, w.rhs), 1, x)
    info, stats = solver.info, solver.getStats()
    rows = int(stats.iterationCount) + 1
    assert info.reportedIterations == rows and len(solver.history) == rows
    assert abs(rows - len(h_ref)) <= 1
    n = min(rows, len(h_ref))
    assert np.allclose(solver.history[:n], h_ref[:n], rtol=1e-6)
    assert solver.history[-1] == stats.absoluteResidual
    assert info.initialResidual == pytest.approx(np.linalg.norm(w.rhs), rel=1e-12)  # x0 = 0
    table = check_table(solver, rows)
    for k, (_, absolute, relative) in enumerate(table):
        assert absolute == format(float(solver.history[k]), ".5E")
        assert relative == format(float(solver.history[k]) / info.initialResidual, ".5E")


def test_verbosity_two_reports_header_and_converged_row(hip, torch_gpu, five_point):
    w, matrix, _, _ = five_point
    solver = hip.FusedPCG(matrix, Statistics(5000, 0.0, w.relative_tolerance), verbosity=2)
    x = torch_gpu.zeros(w.matrix.rows, dtype=torch_gpu.float64, device="cuda:0")
    solver.product(0, to_dev(torch_gpu, w.rhs), 1, x)
    assert len(solver.reports) == 2 and solver.reports[0] == HEADER
    assert [row[0] for row in rows_of(solver.reports)] == [int(solver.getStats().iterationCount)]
    assert solver.reports[1].endswith(RULE)
    quiet = hip.FusedPCG(matrix, Statistics(5000, 0.0, w.relative_tolerance))
    quiet.product(0, to_dev(torch_gpu, w.rhs), 1, torch_gpu.zeros_like(x))
    assert quiet.reports == [] and quiet.history is None


def test_generic_cg_reports_match_the_fused_layout(hip, torch_gpu, five_point):
    w, matrix, s_ref, h_ref = five_point
    space = hip.HIPSpace()
    solver = hip.ConjugateGradients(hip.HIPCSROperator(matrix, space), space, hip.makeDiagonalOperator(matrix, space),
                                    Statistics(5000, 0.0, w.relative_tolerance), verbosity=3)
    x = space.makeVector(w.matrix.rows)
    space.fill(x, 0)
    solver.product(0, to_dev(torch_gpu, w.rhs), 1, x)
    rows = int(solver.getStats().iterationCount) + 1
    table = check_table(solver, rows)
    assert [absolute for _, absolute, _ in table] == [format(v, ".5E") for v in solver.history]
    n = min(rows, len(h_ref))
    assert np.allclose(solver.history[:n], h_ref[:n], rtol=1e-6)


def test_mixed_reports_end_on_the_fp64_residual(hip, torch_gpu):
    w = systems.q2_laplace_3d(8)
    matrix = hip.CSRMatrix(w.matrix)
    solver = hip.MixedPCG(matrix, Statistics(10000, 0.0, 1e-10), verbosity=3)
    x = torch_gpu.zeros(w.matrix.rows, dtype=torch_gpu.float64, device="cuda:0")
    solver.product(0, to_dev(torch_gpu, w.rhs), 1, x)
    info, stats = solver.info, solver.getStats()
    rows = int(stats.iterationCount) + 1
    assert info.refinements >= 1 and info.reportedIterations == rows
    # The last trip folded the correction: its entry is the FP64 true residual that passed the test.
    assert solver.history[-1] == stats.absoluteResidual
    assert stats.relativeResidual < 1e-10
    # An FP32 residual that passes the tolerance triggers a refinement on the next trip, so rows that look
    # converged are at most one per refinement besides the last; each closes the table once.
    looks_converged = np.count_nonzero(solver.history / info.initialResidual < 1e-10)
    assert 1 <= looks_converged <= info.refinements + 1
    check_table(solver, rows, closings=looks_converged)


def test_start_check_reports_one_termination(hip, torch_gpu, five_point):
    w, matrix, _, _ = five_point
    x0 = torch_gpu.ones(w.matrix.rows, dtype=torch_gpu.float64, device="cuda:0")
    for cls in (hip.FusedPCG, hip.MixedPCG):
        solver = cls(matrix, Statistics(100, 1e300, 0.0), verbosity=2)
        x = x0.clone()
        solver.product(1, to_dev(torch_gpu, w.rhs), 1, x)
        assert torch_gpu.equal(x, x0)  # the start check returns before the output step
        assert solver.info.reportedIterations == 0
        assert len(solver.reports) == 1
        text = solver.reports[0]
        assert text.startswith(HEADER) and text.endswith(RULE)
        assert rows_of([text]) == [(0, format(solver.info.initialResidual, ".5E"), "1.00000E+00")]


def test_iteration_cap_reports_every_trip_without_closing(hip, torch_gpu, five_point):
    w, matrix, _, h_ref = five_point
    solver = hip.FusedPCG(matrix, Statistics(5, 0.0, w.relative_tolerance), verbosity=3)
    x = torch_gpu.zeros(w.matrix.rows, dtype=torch_gpu.float64, device="cuda:0")
    solver.product(0, to_dev(torch_gpu, w.rhs), 1, x)
    assert solver.getStats().iterationCount == 5 and solver.info.reportedIterations == 5
    check_table(solver, 5, converged=False)
    assert np.allclose(solver.history, h_ref[:5], rtol=1e-10)


def test_history_respects_its_capacity(hip, torch_gpu, five_point):
    """Through the C ABI: trips at or past the capacity are not written."""
    w, matrix, _, h_ref = five_point
    lib = _lib.library()
    buffer = torch_gpu.full((16,), -1.0, dtype=torch_gpu.float64, device="cuda:0")
    handle = matrix.context.handle
    torch_gpu.cuda.synchronize()
    _lib.check(lib.cie_hip_context_set_residual_history(handle, ctypes.c_void_p(buffer.data_ptr()), 3), "history")
    try:
        info = _lib.SolveInfo()
        x = torch_gpu.zeros(w.matrix.rows, dtype=torch_gpu.float64, device="cuda:0")
        b = to_dev(torch_gpu, w.rhs)
        settings = Statistics(8, 0.0, w.relative_tolerance)
        _lib.check(lib.cie_hip_pcg_f64(handle, matrix.handle, 0.0, ctypes.c_void_p(b.data_ptr()), 1.0,
                                       ctypes.c_void_p(x.data_ptr()), w.matrix.rows, ctypes.byref(settings), 0,
                                       ctypes.byref(info)), "pcg")
    finally:
        lib.cie_hip_context_set_residual_history(handle, None, 0)
    got = buffer.cpu().numpy()
    assert np.allclose(got[:3], h_ref[:3], rtol=1e-10)
    assert np.all(got[3:] == -1.0)
    with pytest.raises(_lib.InvalidArgument):
        _lib.check(lib.cie_hip_context_set_residual_history(handle, ctypes.c_void_p(buffer.data_ptr()), -1), "history")

"""The per-iteration report (cie_hip_iteration_report), CPU only: pure host formatting.

Pinned against IterativeSolver::makeIterationReport
(linalg/packages/solvers/src/IterativeSolver.cpp:17-58) three ways:
- literal reports written out by hand from that code for the cases it
  distinguishes (header at iteration 0, row on termination / convergence /
  verbosity 3, closing rule once converged, nothing below verbosity 2);
- the number cells against Python's format mini-language ("{:>17.5E}" and
  "{:>9}"), whose grammar std::format's standard format specification follows,
  over random, tiny, huge, zero and non-finite values;
- the colours against the 24-bit ANSI escape of the reference's colour
  constants, TUMGreen (162, 173, 0) and TUMOrange (227, 114, 34)
  (cieutils/packages/types/src/Color.cpp:191-203, :255-258).
The library is loaded, no device is touched.
"""
import math
import re

import numpy as np
import pytest

from cie_amd import _lib
from cie_amd._lib import Statistics
from cie_amd.linalg import makeIterationReport

GREEN = "\x1b[38;2;162;173;0m"
ORANGE = "\x1b[38;2;227;114;34m"
RESET = "\x1b[0m"
RULE = "+ --------- + ----------------- + ----------------- +"
HEADER = RULE + "\n| iteration | absolute residual | relative residual |\n" + RULE
TERMINATION, ITERATION = _lib.REPORT_TERMINATION, _lib.REPORT_ITERATION


def report(verbosity, kind, iteration, absolute, relative, settings=(0.0, 1e-10)):
    return makeIterationReport(verbosity, kind, Statistics(iteration, absolute, relative),
                               Statistics(1000, settings[0], settings[1]))


def test_escape_codes_follow_the_colour_constants():
    assert GREEN == "\x1b[38;2;" + ";".join(map(str, (162, 173, 0))) + "m"
    assert ORANGE == "\x1b[38;2;" + ";".join(map(str, (227, 114, 34))) + "m"


def test_reports_below_verbosity_two_are_empty():
    for verbosity in (-1, 0, 1):
        for kind in (TERMINATION, ITERATION):
            assert report(verbosity, kind, 0, 1.0, 1.0) == ""
            assert report(verbosity, kind, 7, 1e-20, 1e-20) == ""


def test_first_iteration_at_verbosity_three():
    expected = (HEADER + "\n| " + ORANGE + "        0" + RESET + " | " + ORANGE + "      1.50000E+00" + RESET
                + " | " + ORANGE + "      5.00000E-01" + RESET + " |")
    assert report(3, ITERATION, 0, 1.5, 0.5) == expected


def test_first_iteration_at_verbosity_two_prints_the_header_only():
    assert report(2, ITERATION, 0, 1.5, 0.5) == HEADER


def test_later_unconverged_iteration():
    assert report(2, ITERATION, 12, 3.25e-3, 2.5e-6) == ""
    expected = ("| " + ORANGE + "       12" + RESET + " | " + ORANGE + "      3.25000E-03" + RESET + " | " + ORANGE
                + "      2.50000E-06" + RESET + " |")
    assert report(3, ITERATION, 12, 3.25e-3, 2.5e-6) == expected


def test_converged_iteration_closes_the_table():
    expected = ("| " + GREEN + "      388" + RESET + " | " + ORANGE + "      2.71828E-09" + RESET + " | " + GREEN
                + "      9.43524E-11" + RESET + " |\n" + RULE)
    for verbosity in (2, 3, 4):
        assert report(verbosity, ITERATION, 388, 2.718281828e-9, 9.43524349965e-11) == expected


def test_absolute_tolerance_alone_converges():
    text = report(2, ITERATION, 5, 1e-12, 0.25, settings=(1e-11, 1e-300))
    assert text == ("| " + GREEN + "        5" + RESET + " | " + GREEN + "      1.00000E-12" + RESET + " | " + ORANGE
                    + "      2.50000E-01" + RESET + " |\n" + RULE)


def test_termination_at_the_start_check():
    """ConjugateGradients_impl.hpp:100-111: iteration 0, relative residual 1, tolerance met by the absolute one."""
    text = report(2, TERMINATION, 0, 4.0e-3, 1.0, settings=(1.0, 0.0))
    assert text == (HEADER + "\n| " + GREEN + "        0" + RESET + " | " + GREEN + "      4.00000E-03" + RESET
                    + " | " + ORANGE + "      1.00000E+00" + RESET + " |\n" + RULE)
    # A termination report of an unconverged state (a cap) still prints its row.
    assert report(2, TERMINATION, 9, 1.0, 0.5) == ("| " + ORANGE + "        9" + RESET + " | " + ORANGE
                                                   + "      1.00000E+00" + RESET + " | " + ORANGE + "      5.00000E-01"
                                                   + RESET + " |")


def cells(text):
    row = re.search(r"\| (.*) \| (.*) \| (.*) \|", text.splitlines()[-2 if text.endswith(RULE) else -1])
    strip = lambda cell: cell.replace(GREEN, "").replace(ORANGE, "").replace(RESET, "")  # noqa: E731
    return [strip(c) for c in row.groups()]


def test_number_cells_follow_the_format_specification():
    rng = np.random.default_rng(11)
    values = list(10.0 ** rng.uniform(-320, 308, 400)) + list(rng.standard_normal(100))
    values += [0.0, -0.0, 5e-324, 2.2250738585072014e-308, 1.7976931348623157e308, 0.5, 9.999995e-5, 1.0000049999e3,
               float(np.float32(1e-6)), float(np.finfo(np.float32).max), math.inf]
    for k, value in enumerate(values):
        iteration = int(rng.integers(0, 10**9)) if k % 3 else k
        count, absolute, relative = cells(report(3, ITERATION, iteration, value, value / 3.0 if math.isfinite(value)
                                                 else value))
        assert count == format(iteration, ">9"), (iteration, count)
        assert absolute == format(value, ">17.5E"), (value, absolute)
        expected = format(value / 3.0 if math.isfinite(value) else value, ">17.5E")
        assert relative == expected, (value, relative)


def test_null_buffer_gives_the_length_and_a_short_buffer_truncates():
    import ctypes

    lib = _lib.library()
    status, settings = Statistics(0, 1.0, 1.0), Statistics(10, 0.0, 1e-10)
    length = ctypes.c_size_t()
    assert lib.cie_hip_iteration_report(3, ITERATION, ctypes.byref(status), ctypes.byref(settings), None, 0,
                                        ctypes.byref(length)) == 0
    full = report(3, ITERATION, 0, 1.0, 1.0)
    assert length.value == len(full)
    buffer = ctypes.create_string_buffer(8)
    assert lib.cie_hip_iteration_report(3, ITERATION, ctypes.byref(status), ctypes.byref(settings), buffer, 8,
                                        ctypes.byref(length)) == 0
    assert buffer.value.decode() == full[:7] and length.value == len(full)
    with pytest.raises(_lib.InvalidArgument):
        _lib.check(lib.cie_hip_iteration_report(3, 2, ctypes.byref(status), ctypes.byref(settings), None, 0, None),
                   "cie_hip_iteration_report")

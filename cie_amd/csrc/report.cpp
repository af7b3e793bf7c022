// The iterative solvers' per-iteration report (host code).
//
// Follows IterativeSolver<TS>::makeIterationReport
// (linalg/packages/solvers/src/IterativeSolver.cpp:17-58): at verbosity >= 2 a
// table header when the iteration count is 0, one row "| iteration | absolute
// residual | relative residual |" on termination, on convergence, or on every
// iteration at verbosity >= 3, and a closing rule once converged. Each cell is
// coloured with a 24-bit ANSI foreground escape: TUM green where the value is
// below its tolerance (plain `<`, as the report tests it), TUM orange where
// not (cieutils/packages/types/src/Color.cpp:191-203, :255-258), then reset.
// The cells are right-aligned: the count in 9 columns, the residuals in 17 in
// scientific notation with 5 digits after the point and an upper-case E
// (std::format "{:>9}" and "{:>17.5E}"; printf's "%9llu" and "%17.5E" print
// the same characters).
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/cie_hip.h"

namespace ciehip {
int setError(int status, const std::string& message);
}

namespace {

constexpr const char* kGreen = "\x1b[38;2;162;173;0m";   // RGBAColor::TUMGreen (162, 173, 0)
constexpr const char* kOrange = "\x1b[38;2;227;114;34m"; // RGBAColor::TUMOrange (227, 114, 34)
constexpr const char* kReset = "\033[0m";
constexpr const char* kRule = "+ --------- + ----------------- + ----------------- +";

std::string cell(const char* format, const char* colour, double value) {
    char text[64];
    std::snprintf(text, sizeof text, format, value);
    return std::string(colour) + text + kReset;
}

}  // namespace

extern "C" int cie_hip_iteration_report(int verbosity, int reportType, const cie_hip_statistics* status,
                                        const cie_hip_statistics* settings, char* out, size_t capacity,
                                        size_t* length) {
    if (!status || !settings)
        return ciehip::setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_iteration_report: null statistics");
    if (reportType != CIE_HIP_REPORT_TERMINATION && reportType != CIE_HIP_REPORT_ITERATION)
        return ciehip::setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_iteration_report: unknown report type");
    std::string message;
    const bool absoluteMet = status->absoluteResidual < settings->absoluteResidual;
    const bool relativeMet = status->relativeResidual < settings->relativeResidual;
    const bool converged = absoluteMet || relativeMet;
    if (verbosity >= 2) {
        const bool first = status->iterationCount == 0;
        if (first) {
            message += kRule;
            message += "\n| iteration | absolute residual | relative residual |\n";
            message += kRule;
        }
        if (reportType == CIE_HIP_REPORT_TERMINATION || converged ||
            (verbosity >= 3 && reportType == CIE_HIP_REPORT_ITERATION)) {
            if (first) message += "\n";
            char count[32];
            std::snprintf(count, sizeof count, "%9" PRIu64, status->iterationCount);
            message += "| ";
            message += converged ? kGreen : kOrange;
            message += count;
            message += kReset;
            message += " | " + cell("%17.5E", absoluteMet ? kGreen : kOrange, status->absoluteResidual);
            message += " | " + cell("%17.5E", relativeMet ? kGreen : kOrange, status->relativeResidual);
            message += " |";
        }
        if (converged) {
            message += "\n";
            message += kRule;
        }
    }
    if (length) *length = message.size();
    if (out && capacity) {
        const size_t copied = message.size() < capacity ? message.size() : capacity - 1;
        std::memcpy(out, message.data(), copied);
        out[copied] = '\0';
    }
    return CIE_HIP_OK;
}

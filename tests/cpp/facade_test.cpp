// Drives the C++ face (include/cie_hip/linalg.hpp) the way the reference's
// poisson2D driver drives its solver layer (executables/fem_examples/src/
// poisson2D/solver.cpp:82-108): space + CSR operator + Jacobi + CG, then
// product(0, b, 1, x) and getStats().
//
// usage: facade_test <input.bin> <output-prefix>
//   input: int64 rows, int64 nnz, double tolerance, int32 extents[rows+1],
//          int32 columns[nnz], double entries[nnz], double rhs[rows]
//   writes <prefix>_{generic,fused,mixed}.bin (rows doubles) and
//   <prefix>_generic32.bin (rows floats: ConjugateGradients<HIPSpace<float>>,
//   the reference's FP32 instantiation, ConjugateGradients.cpp:10,16, run to
//   1e-6 as the driver's settings are) and prints one line per solver: name
//   iterations absolute relative; then "early <name> <scales> <0|1>", 1 when
//   an early exit (ConjugateGradients_impl.hpp:100-111) left x bit for bit;
//   and <prefix>_reports_{generic,fused}.txt, the solvers' per-iteration
//   reports at verbosity 3, one per line.
#include "cie_hip/linalg.hpp"

#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <vector>

using namespace cie::linalg;

template <class T>
static void readInto(std::ifstream& in, std::vector<T>& v, std::size_t n) {
    v.resize(n);
    in.read(reinterpret_cast<char*>(v.data()), std::streamsize(sizeof(T) * n));
}

static void write(const std::string& path, const std::vector<double>& v) {
    std::ofstream out(path, std::ios::binary);
    out.write(reinterpret_cast<const char*>(v.data()), std::streamsize(sizeof(double) * v.size()));
}

int main(int argc, char** argv) {
    if (argc != 3) {
        std::cerr << "usage: facade_test <input.bin> <output-prefix>\n";
        return 2;
    }
    std::ifstream in(argv[1], std::ios::binary);
    std::int64_t rows = 0, nnz = 0;
    double tolerance = 0.0;
    in.read(reinterpret_cast<char*>(&rows), 8);
    in.read(reinterpret_cast<char*>(&nnz), 8);
    in.read(reinterpret_cast<char*>(&tolerance), 8);
    std::vector<int> extents, columns;
    std::vector<double> entries, rhs;
    readInto(in, extents, std::size_t(rows + 1));
    readInto(in, columns, std::size_t(nnz));
    readInto(in, entries, std::size_t(nnz));
    readInto(in, rhs, std::size_t(rows));
    if (!in) {
        std::cerr << "short input\n";
        return 2;
    }
    const std::string prefix = argv[2];
    try {
        auto context = std::make_shared<const HIPContext>(0);
        auto space = std::make_shared<HIPSpace<double>>(context);
        const CSRView<const double, const int> view(std::size_t(rows), {extents.data(), extents.size()},
                                                    {columns.data(), columns.size()}, {entries.data(), entries.size()});
        auto matrix = std::make_shared<const HIPCSRMatrix>(context, view);
        auto lhs = std::make_shared<HIPCSROperator<int, double>>(matrix, space);
        auto preconditioner = std::make_shared<HIPDiagonalOperator<double>>(
            makeDiagonalOperator<double>(*matrix, std::shared_ptr<const HIPSpace<double>>(space)));
        using Solver = ConjugateGradients<HIPSpace<double>>;
        const Solver::Statistics settings{20000, 0.0, tolerance};

        auto b = space->makeVector(std::size_t(rows));
        space->assign(space->view(b), std::span<const double>(rhs));
        std::vector<double> host(static_cast<std::size_t>(rows));

        {
            Solver solver(lhs, space, preconditioner, settings, 0);
            auto x = space->makeVector(std::size_t(rows));
            space->fill(space->view(x), 0.0);
            solver.product(0, space->view(std::as_const(b)), 1, space->view(x));
            const auto stats = solver.getStats().value();
            space->assign(std::span<double>(host), space->view(std::as_const(x)));
            write(prefix + "_generic.bin", host);
            std::printf("generic %zu %.17g %.17g\n", stats.iterationCount, stats.absoluteResidual, stats.relativeResidual);
        }
        for (const bool mixed : {false, true}) {
            FusedConjugateGradients solver(matrix, space, {20000, 0.0, tolerance}, mixed);
            auto x = space->makeVector(std::size_t(rows));
            space->fill(space->view(x), 0.0);
            solver.product(0, space->view(std::as_const(b)), 1, space->view(x));
            const auto stats = solver.getStats().value();
            space->assign(std::span<double>(host), space->view(std::as_const(x)));
            write(prefix + (mixed ? "_mixed.bin" : "_fused.bin"), host);
            std::printf("%s %zu %.17g %.17g\n", mixed ? "mixed" : "fused", stats.iterationCount,
                        stats.absoluteResidual, stats.relativeResidual);
        }
        {
            // The reference's FP32 instantiation over the same CSR (FP32 plane).
            auto space32 = std::make_shared<HIPSpace<float>>(context);
            auto lhs32 = std::make_shared<HIPCSROperator<int, float>>(matrix, space32);
            auto preconditioner32 = std::make_shared<HIPDiagonalOperator<float>>(
                makeDiagonalOperator<float>(*matrix, std::shared_ptr<const HIPSpace<float>>(space32)));
            using Solver32 = ConjugateGradients<HIPSpace<float>>;
            Solver32 solver(lhs32, space32, preconditioner32, Solver32::Statistics{20000, 0.0f, 1e-6f}, 0);
            std::vector<float> rhs32(rhs.begin(), rhs.end()), host32(static_cast<std::size_t>(rows));
            auto b32 = space32->makeVector(std::size_t(rows));
            space32->assign(space32->view(b32), std::span<const float>(rhs32));
            auto x = space32->makeVector(std::size_t(rows));
            space32->fill(space32->view(x), 0.0f);
            solver.product(0, space32->view(std::as_const(b32)), 1, space32->view(x));
            const auto stats = solver.getStats().value();
            space32->assign(std::span<float>(host32), space32->view(std::as_const(x)));
            std::ofstream out(prefix + "_generic32.bin", std::ios::binary);
            out.write(reinterpret_cast<const char*>(host32.data()), std::streamsize(sizeof(float) * host32.size()));
            std::printf("generic32 %zu %.9g %.9g\n", stats.iterationCount, double(stats.absoluteResidual),
                        double(stats.relativeResidual));
        }
        {
            // Early exit: an absolute tolerance above any residual ends every solve at the start check,
            // which returns before the output step (:190-198), so x keeps its bits whatever the scales.
            const double scales[2][2] = {{1.0, 1.0}, {0.5, -2.0}};
            for (const char* name : {"generic", "fused", "mixed"}) {
                for (const auto& scale : scales) {
                    auto x = space->makeVector(std::size_t(rows));
                    std::vector<double> before(static_cast<std::size_t>(rows)), after(before.size());
                    for (std::size_t i = 0; i < before.size(); ++i) before[i] = double(i % 17) - 8.0;
                    space->assign(space->view(x), std::span<const double>(before));
                    const Solver::Statistics early{100, 1e300, 0.0};
                    if (std::string(name) == "generic") {
                        Solver solver(lhs, space, preconditioner, early, 0);
                        solver.product(scale[0], space->view(std::as_const(b)), scale[1], space->view(x));
                    } else {
                        FusedConjugateGradients solver(matrix, space, {100, 1e300, 0.0}, std::string(name) == "mixed");
                        solver.product(scale[0], space->view(std::as_const(b)), scale[1], space->view(x));
                    }
                    space->assign(std::span<double>(after), space->view(std::as_const(x)));
                    std::printf("early %s %g,%g %d\n", name, scale[0], scale[1], before == after ? 1 : 0);
                }
            }
        }
        {
            // Per-iteration reports at verbosity 3 (ConjugateGradients_impl.hpp:175-182), one file per solver.
            for (const std::string name : {"generic", "fused"}) {
                std::ofstream log(prefix + "_reports_" + name + ".txt", std::ios::binary);
                setReportStream(&log);
                auto x = space->makeVector(std::size_t(rows));
                space->fill(space->view(x), 0.0);
                if (name == "generic") {
                    Solver solver(lhs, space, preconditioner, settings, 3);
                    solver.product(0, space->view(std::as_const(b)), 1, space->view(x));
                } else {
                    FusedConjugateGradients solver(matrix, space, {20000, 0.0, tolerance}, false,
                                                   CIE_HIP_DEFAULT_REFINEMENT_FACTOR, 3);
                    solver.product(0, space->view(std::as_const(b)), 1, space->view(x));
                }
                setReportStream(&std::clog);
            }
        }
        // A missing diagonal raises like makeDiagonalOperator's CIE_THROW (DiagonalOperator.cpp:79-82).
        std::vector<int> badExtents{0, 1, 2}, badColumns{0, 0};
        std::vector<double> badEntries{1.0, 1.0};
        const CSRView<const double, const int> bad(2, {badExtents.data(), 3}, {badColumns.data(), 2},
                                                   {badEntries.data(), 2});
        HIPCSRMatrix badMatrix(context, bad);
        try {
            (void)makeDiagonalOperator<double>(badMatrix, std::shared_ptr<const HIPSpace<double>>(space));
            std::printf("missing-diagonal not-raised\n");
        } catch (const HIPError& error) {
            std::printf("missing-diagonal %d\n", error.status());
        }
    } catch (const HIPError& error) {
        std::cerr << "HIPError " << error.status() << ": " << error.what() << "\n";
        return 1;
    }
    return 0;
}

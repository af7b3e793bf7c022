// Drives the C++ face (include/cie_hip/linalg.hpp) the way the reference's
// poisson2D driver drives its solver layer (executables/fem_examples/src/
// poisson2D/solver.cpp:82-108): space + CSR operator + Jacobi + CG, then
// product(0, b, 1, x) and getStats().
//
// usage: facade_test <input.bin> <output-prefix>
//   input: int64 rows, int64 nnz, double tolerance, int32 extents[rows+1],
//          int32 columns[nnz], double entries[nnz], double rhs[rows]
//   writes <prefix>_{generic,fused,mixed}.bin (rows doubles) and prints one
//   line per solver: name iterations absolute relative
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

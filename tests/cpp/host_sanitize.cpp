// Host code under the sanitizers: the threaded CPU oracle (oracle/oracle.c,
// pthread pool), the structured-system generator (cie_amd/csrc/systems.cpp,
// worker threads) and the MatrixMarket writer and reader
// (cie_amd/csrc/matrix_market.cpp, threaded formatter), built from source with
// -fsanitize=address,undefined or -fsanitize=thread by
// tests/test_host_sanitizers.py, as the reference's CI runs its own tests
// (linalg/.github/workflows/github-actions.yml:21-55). Every call is checked
// against a property it must hold, so a run is a test and not only a trace:
// the process exits nonzero on the first failed check, and the sanitizers
// abort on the first report.
//
// usage: host_sanitize <scratch-directory>
#include "../../include/cie_hip.h"
#include "../../include/cie_hip_systems.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

// matrix_market.cpp reports errors through the library's error state, which
// lives in the HIP translation units; the host-only build keeps its own.
namespace ciehip {
static thread_local std::string lastError;
int setError(int status, const std::string& message) {
    lastError = message;
    return status;
}
}  // namespace ciehip
extern "C" const char* cie_hip_last_error(void) { return ciehip::lastError.c_str(); }

extern "C" {
typedef struct {
    uint64_t iterationCount;
    double absoluteResidual;
    double relativeResidual;
} oracle_statistics_t;
int oracle_csr_product(int64_t, const int32_t*, const int32_t*, const double*, double, const double*, double, double*,
                       int);
double oracle_inner_product(int64_t, const double*, const double*, int);
int oracle_vector_op(int, double*, const double*, double, int64_t, int);
int oracle_inverse_diagonal(int64_t, const int32_t*, const int32_t*, const double*, double*, int64_t*);
int oracle_pcg(int64_t, const int32_t*, const int32_t*, const double*, double, const double*, double, double*,
               uint64_t, double, double, int, int, oracle_statistics_t*, double*, int64_t, double*);
int oracle_element_gradient(int64_t, int32_t, const int32_t*, const double*, const double*, const double*, double*);
int oracle_assemble(int64_t, const int32_t*, const int32_t*, int64_t, int32_t, const int32_t*, const double*, int64_t,
                    const double*, double*);
int oracle_assemble_vector(int64_t, int64_t, int32_t, const int32_t*, const double*, int64_t, const double*, double*);
int oracle_make_reordering(int64_t, const int32_t*, const int32_t*, int, int32_t*);
int oracle_reorder_csr(int64_t, const int32_t*, int, const int32_t*, const int32_t*, const double*, int32_t*, int32_t*,
                       double*);
int oracle_reorder_vector(int64_t, const int32_t*, int, const double*, double*);
int oracle_masked_csr_product(int64_t, const int32_t*, const int32_t*, const double*, const uint16_t*, uint16_t,
                              double, const double*, double, double*, int);
int oracle_jacobi(int64_t, const int32_t*, const int32_t*, const double*, const uint16_t*, uint16_t, uint64_t, double,
                  double, const double*, double, double*, int);
int oracle_approximate_substitution(int64_t, const int32_t*, const int32_t*, const double*, uint64_t, double,
                                    const double*, double, double*);
int oracle_pmg(int64_t, const int32_t*, const int32_t*, const double*, const uint16_t*, int, const double*, double*,
               double, int64_t, double*, uint64_t*, int64_t, int64_t*, uint64_t*, int);
}

namespace {

int failures = 0;

void expect(bool condition, const char* what) {
    if (!condition) {
        std::fprintf(stderr, "FAILED: %s\n", what);
        ++failures;
    }
}

struct Csr {
    int64_t rows = 0, nnz = 0;
    std::vector<int32_t> extents, columns;
    std::vector<double> entries;
};

Csr lagrange(const cie_sys_lagrange_spec& spec, const double* conductivity, int threads) {
    Csr m;
    int64_t elements = 0;
    int32_t npe = 0;
    expect(cie_sys_lagrange_size(&spec, &m.rows, &m.nnz, &elements, &npe) == 0, "lagrange size");
    m.extents.resize(static_cast<size_t>(m.rows + 1));
    m.columns.resize(static_cast<size_t>(m.nnz));
    m.entries.resize(static_cast<size_t>(m.nnz));
    expect(cie_sys_lagrange_fill(&spec, conductivity, m.extents.data(), m.columns.data(), m.entries.data(),
                                 threads) == 0,
           "lagrange fill");
    return m;
}

double norm(const std::vector<double>& v) { return std::sqrt(oracle_inner_product(int64_t(v.size()), v.data(), v.data(), 1)); }

void generatorAndSolver() {
    // Random streams are reproducible.
    std::vector<double> u1(1000), u2(1000), g(1000);
    expect(cie_sys_uniform(1, 1000, u1.data()) == 0 && cie_sys_uniform(1, 1000, u2.data()) == 0, "uniform");
    expect(u1 == u2, "uniform stream reproducible");
    expect(cie_sys_normal(2, 0.5, 1000, g.data()) == 0, "normal");

    // 3D Q2 with lognormal conductivity on 4 generator threads equals the 1-thread build.
    // (19^3 = 6,859 rows: above the generator's 4,096-row threshold for threads)
    cie_sys_lagrange_spec spec{3, 2, 10, 1, CIE_SYS_BASIS_LAGRANGE};
    std::vector<double> k(10 * 10 * 10);
    for (size_t e = 0; e < k.size(); ++e) k[e] = std::exp(g[e]);
    const Csr a = lagrange(spec, k.data(), 4);
    const Csr a1 = lagrange(spec, k.data(), 1);
    expect(a.extents == a1.extents && a.columns == a1.columns && a.entries == a1.entries, "threaded generator");

    // Threaded products, dots and vector ops equal the sequential ones (rows are independent).
    std::vector<double> x(static_cast<size_t>(a.rows)), y1(static_cast<size_t>(a.rows), 0.0), y4(static_cast<size_t>(a.rows), 0.0);
    expect(cie_sys_uniform(3, a.rows, x.data()) == 0, "rhs");
    oracle_csr_product(a.rows, a.extents.data(), a.columns.data(), a.entries.data(), 0.0, x.data(), 1.0, y1.data(), 1);
    oracle_csr_product(a.rows, a.extents.data(), a.columns.data(), a.entries.data(), 0.0, x.data(), 1.0, y4.data(), 4);
    expect(y1 == y4, "threaded SpMV");
    const double dot1 = oracle_inner_product(a.rows, x.data(), y1.data(), 1);
    const double dot4 = oracle_inner_product(a.rows, x.data(), y1.data(), 4);
    expect(dot1 > 0 && std::fabs(dot1 - dot4) <= 1e-12 * dot1, "threaded dot (x^T A x > 0)");
    std::vector<double> t = x;
    oracle_vector_op(2, t.data(), y1.data(), 0.5, a.rows, 4);  // add
    expect(t[7] == x[7] + 0.5 * y1[7], "threaded add");

    // Jacobi-PCG on 4 pool threads: converges, and its solution solves the system.
    std::vector<double> sol(static_cast<size_t>(a.rows), 0.0), hist(200);
    oracle_statistics_t stats{};
    double seconds = 0.0;
    const int status = oracle_pcg(a.rows, a.extents.data(), a.columns.data(), a.entries.data(), 0.0, x.data(), 1.0,
                                  sol.data(), 5000, 0.0, 1e-10, 1, 4, &stats, hist.data(), int64_t(hist.size()), &seconds);
    expect(status == 0 && stats.relativeResidual < 1e-10, "threaded PCG converges");
    std::vector<double> r = x;
    oracle_csr_product(a.rows, a.extents.data(), a.columns.data(), a.entries.data(), 1.0, sol.data(), -1.0, r.data(), 4);
    expect(norm(r) < 1e-9 * norm(x), "PCG solution solves A x = b");

    // Inverse diagonal and its missing-diagonal error.
    std::vector<double> inverse(static_cast<size_t>(a.rows));
    int64_t failed = -1;
    expect(oracle_inverse_diagonal(a.rows, a.extents.data(), a.columns.data(), a.entries.data(), inverse.data(),
                                   &failed) == 0,
           "inverse diagonal");

    // Assembly of the per-element matrices equals the generator's system.
    int64_t rows = 0, nnz = 0, elements = 0;
    int32_t npe = 0;
    cie_sys_lagrange_size(&spec, &rows, &nnz, &elements, &npe);
    std::vector<int32_t> dofs(static_cast<size_t>(elements * npe));
    std::vector<double> ke(static_cast<size_t>(npe * npe)), assembled(static_cast<size_t>(nnz));
    expect(cie_sys_lagrange_element_dofs(&spec, dofs.data()) == 0, "element dofs");
    expect(cie_sys_lagrange_element_matrix(&spec, ke.data()) == 0, "element matrix");
    expect(oracle_assemble(rows, a.extents.data(), a.columns.data(), elements, npe, dofs.data(), ke.data(), 0, k.data(),
                           assembled.data()) == 0,
           "assemble");
    double worst = 0.0;
    for (int64_t i = 0; i < nnz; ++i) worst = std::fmax(worst, std::fabs(assembled[static_cast<size_t>(i)] - a.entries[static_cast<size_t>(i)]));
    expect(worst < 1e-12, "assembly equals the generator");
    std::vector<double> rhs(static_cast<size_t>(rows));
    std::vector<double> ones(static_cast<size_t>(elements * npe), 1.0);
    expect(oracle_assemble_vector(rows, elements, npe, dofs.data(), ones.data(), npe, nullptr, rhs.data()) == 0,
           "assemble vector");
    std::vector<double> gradient(static_cast<size_t>(elements));
    expect(oracle_element_gradient(elements, npe, dofs.data(), ke.data(), sol.data(), sol.data(), gradient.data()) == 0,
           "element gradient");
    for (double v : gradient) expect(v <= 1e-12, "-u_e^T K_e u_e <= 0");

    // Reordering: RCM map is a permutation; permuting there and back restores the matrix.
    std::vector<int32_t> map(static_cast<size_t>(rows)), seen(static_cast<size_t>(rows), 0);
    expect(oracle_make_reordering(rows, a.extents.data(), a.columns.data(), 1, map.data()) == 0, "RCM");
    for (int32_t v : map) expect(v >= 0 && v < rows && !seen[static_cast<size_t>(v)]++, "RCM map is a permutation");
    std::vector<int32_t> e2(static_cast<size_t>(rows + 1)), c2(static_cast<size_t>(nnz)), e3(static_cast<size_t>(rows + 1)), c3(static_cast<size_t>(nnz));
    std::vector<double> v2(static_cast<size_t>(nnz)), v3(static_cast<size_t>(nnz));
    expect(oracle_reorder_csr(rows, map.data(), 0, a.extents.data(), a.columns.data(), a.entries.data(), e2.data(),
                              c2.data(), v2.data()) == 0 &&
               oracle_reorder_csr(rows, map.data(), 1, e2.data(), c2.data(), v2.data(), e3.data(), c3.data(),
                                  v3.data()) == 0,
           "reorder");
    expect(e3 == a.extents && c3 == a.columns && v3 == a.entries, "reorder round trip");
    std::vector<double> xr(static_cast<size_t>(rows)), xb(static_cast<size_t>(rows));
    oracle_reorder_vector(rows, map.data(), 0, x.data(), xr.data());
    oracle_reorder_vector(rows, map.data(), 1, xr.data(), xb.data());
    expect(xb == x, "vector reorder round trip");
}

void multigrid() {
    // 2D integrated Legendre, order 2: masked operators, smoothers and the p-multigrid cycle on pool threads.
    cie_sys_lagrange_spec spec{2, 2, 12, 1, CIE_SYS_BASIS_LEGENDRE};
    const Csr a = lagrange(spec, nullptr, 2);
    std::vector<uint16_t> mask(static_cast<size_t>(a.rows));
    expect(cie_sys_lagrange_ansatz_mask(&spec, mask.data()) == 0, "ansatz mask");
    std::vector<double> b(static_cast<size_t>(a.rows)), y(static_cast<size_t>(a.rows), 0.0), z(static_cast<size_t>(a.rows), 0.0);
    cie_sys_uniform(4, a.rows, b.data());
    oracle_masked_csr_product(a.rows, a.extents.data(), a.columns.data(), a.entries.data(), mask.data(), 2, 0.0, b.data(),
                              1.0, y.data(), 3);
    oracle_masked_csr_product(a.rows, a.extents.data(), a.columns.data(), a.entries.data(), mask.data(), 2, 0.0, b.data(),
                              1.0, z.data(), 1);
    expect(y == z, "threaded masked SpMV");
    std::vector<double> s(static_cast<size_t>(a.rows), 0.0), s1(static_cast<size_t>(a.rows), 0.0);
    oracle_jacobi(a.rows, a.extents.data(), a.columns.data(), a.entries.data(), mask.data(), 3, 4, 2.0 / 3.0, 0.0,
                  b.data(), 1.0, s.data(), 3);
    oracle_jacobi(a.rows, a.extents.data(), a.columns.data(), a.entries.data(), mask.data(), 3, 4, 2.0 / 3.0, 0.0,
                  b.data(), 1.0, s1.data(), 1);
    expect(s == s1, "threaded Jacobi");
    std::vector<double> sub(static_cast<size_t>(a.rows), 0.0);
    expect(oracle_approximate_substitution(a.rows, a.extents.data(), a.columns.data(), a.entries.data(), 3, 0.0,
                                           b.data(), 1.0, sub.data()) == 0,
           "approximate substitution");
    std::vector<double> x(static_cast<size_t>(a.rows), 0.0), history(64);
    std::vector<uint64_t> coarse(64);
    int64_t cycles = 0;
    uint64_t coarseIterations = 0;
    expect(oracle_pmg(a.rows, a.extents.data(), a.columns.data(), a.entries.data(), mask.data(), 2, b.data(), x.data(),
                      1e-6, 60, history.data(), coarse.data(), 64, &cycles, &coarseIterations, 3) == 0,
           "p-multigrid");
    expect(cycles > 0 && coarseIterations > 0, "p-multigrid ran");
}

void matrixMarket(const std::string& directory) {
    // (608,400 entries: three 2^18-entry chunks, so the writer formats on three threads)
    cie_sys_lagrange_spec spec{2, 3, 130, 1, CIE_SYS_BASIS_LAGRANGE};
    const Csr a = lagrange(spec, nullptr, 2);
    expect(a.nnz > 2 * (int64_t(1) << 18), "large enough for threaded formatting");
    const std::string matrixPath = directory + "/lhs.mm", arrayPath = directory + "/rhs.mm";
    expect(cie_hip_mm_write_csr_f64(matrixPath.c_str(), a.rows, a.rows, a.nnz, a.extents.data(), a.columns.data(),
                                    a.entries.data(), 17) == 0,
           "write CSR");
    std::vector<double> b(static_cast<size_t>(a.rows));
    cie_sys_uniform(5, a.rows, b.data());
    expect(cie_hip_mm_write_array_f64(arrayPath.c_str(), b.data(), a.rows, 17) == 0, "write array");
    std::vector<float> b32(b.begin(), b.end());
    const std::string array32 = directory + "/rhs32.mm";
    expect(cie_hip_mm_write_array_f32(array32.c_str(), b32.data(), a.rows, 9) == 0, "write FP32 array");

    cie_hip_mm data = nullptr;
    expect(cie_hip_mm_read(matrixPath.c_str(), &data) == 0, "read CSR");
    int64_t rows = 0, columns = 0, nnz = 0;
    int dense = 1;
    cie_hip_mm_shape(data, &rows, &columns, &nnz, &dense);
    expect(rows == a.rows && columns == a.rows && nnz == a.nnz && !dense, "CSR shape");
    const int32_t *extents = nullptr, *cols = nullptr;
    const double* values = nullptr;
    cie_hip_mm_arrays(data, &extents, &cols, &values);
    expect(std::equal(a.extents.begin(), a.extents.end(), extents) && std::equal(a.columns.begin(), a.columns.end(), cols) &&
               std::equal(a.entries.begin(), a.entries.end(), values),
           "CSR round trip at 17 digits");
    cie_hip_mm_free(data);
    data = nullptr;
    expect(cie_hip_mm_read(arrayPath.c_str(), &data) == 0, "read array");
    cie_hip_mm_shape(data, &rows, &columns, &nnz, &dense);
    cie_hip_mm_arrays(data, &extents, &cols, &values);
    expect(dense && rows == a.rows && std::equal(b.begin(), b.end(), values), "array round trip");
    cie_hip_mm_free(data);
    const std::string missing = directory + "/missing.mm";
    expect(cie_hip_mm_read(missing.c_str(), &data) != 0, "a missing file is an error");
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: host_sanitize <scratch-directory>\n");
        return 2;
    }
    generatorAndSolver();
    multigrid();
    matrixMarket(argv[1]);
    if (failures) {
        std::fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("host_sanitize: all checks passed\n");
    return 0;
}

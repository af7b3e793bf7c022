// Structured finite-element systems for the solver workloads (host code).
//
// The reference assembles its systems with the fem library
// (fem/packages/graph/src/Assembler.cpp:55-175 builds the CSR pattern with
// sorted, unique column indices per row; graph/impl/Assembler_impl.hpp:156-194
// scatters element matrices; integrands/impl/LinearIsotropicStiffnessIntegrand_impl.hpp:74
// forms K_e = modulus * D^T D).  That assembler is not part of the solve path;
// this file produces the same kind of matrix directly for structured meshes:
//
//   * a 2D 5-point finite-difference Laplacian (stencil 4, -1) on an m x m
//     interior grid;
//   * the Laplace stiffness of tensor-product Lagrange elements of order p on
//     an n^d unit-square/cube mesh (d = 2, 3), nodes at the Chebyshev-Lobatto
//     points cos(i*pi/p) of each element (poisson2D/mesh.cpp:73-77), with a
//     per-element conductivity k_e multiplying the element matrix and the
//     Dirichlet boundary rows and columns eliminated;
//   * the same with the hierarchical integrated Legendre basis
//     (fem/packages/maths/src/LegendrePolynomial.cpp: N_0 = (1 - x)/2,
//     N_1 = (1 + x)/2, N_k = int_1^x P_{k-1} = (P_k - P_{k-2}) / (2k - 1) on
//     [-1, 1]), the basis poisson2D uses by default and its p-multigrid
//     needs. Along each dimension the element's vertex functions sit at the
//     element's end positions and N_k (k >= 2) at interior position k - 1, so
//     the node-grid numbering, the coupling pattern and the Dirichlet
//     elimination are those of the Lagrange case.
//
// Node numbering is lexicographic with x fastest, so every row's columns come
// out sorted and unique, as CSRView users expect (DiagonalOperator.cpp:144-148).
// Each nonzero is the sum over the elements that share both nodes, taken in
// increasing element index; the result does not depend on the thread count.
//
// The random inputs use SplitMix64 (S. Vigna, public domain):
//   u = (x >> 11) * 2^-53,  U(-1,1) = 2u - 1,
//   N(0, s^2) = s * sqrt(-2 ln(1 - u1)) * cos(2 pi u2)   (Box-Muller, two draws).

#include "../../include/cie_hip_systems.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_sysError;

int fail(const std::string& message) {
    g_sysError = message;
    return 1;
}

unsigned threadCount(int requested) {
    unsigned count = requested > 0 ? unsigned(requested) : std::thread::hardware_concurrency();
    if (requested <= 0) {
        if (const char* env = std::getenv("OMP_NUM_THREADS")) {
            const long value = std::strtol(env, nullptr, 10);
            if (0 < value && unsigned(value) < count) count = unsigned(value);
        }
    }
    return count ? count : 1u;
}

template <class F>
void parallelRanges(int64_t count, unsigned threads, F&& body) {
    if (threads <= 1 || count < 4096) {
        body(int64_t(0), count);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(threads);
    const int64_t chunk = (count + threads - 1) / threads;
    for (unsigned t = 0; t < threads; ++t) {
        const int64_t begin = std::min<int64_t>(count, int64_t(t) * chunk);
        const int64_t end = std::min<int64_t>(count, begin + chunk);
        if (begin < end) pool.emplace_back([&body, begin, end] { body(begin, end); });
    }
    for (auto& thread : pool) thread.join();
}

struct SplitMix64 {
    uint64_t state;
    uint64_t next() {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double unit() { return double(next() >> 11) * 0x1.0p-53; }
};

// ---------------------------------------------------------------------------
// 1D reference matrices on [0, 1].
// ---------------------------------------------------------------------------

// Gauss-Legendre points and weights on [-1, 1] by Newton iteration on P_q
// (the textbook method; see e.g. Press et al., Numerical Recipes, "gauleg").
void gaussLegendre(int q, std::vector<double>& points, std::vector<double>& weights) {
    points.assign(q, 0.0);
    weights.assign(q, 0.0);
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < (q + 1) / 2; ++i) {
        double z = std::cos(pi * (i + 0.75) / (q + 0.5));
        double derivative = 1.0;
        for (int iteration = 0; iteration < 100; ++iteration) {
            double p0 = 1.0, p1 = 0.0;
            for (int j = 1; j <= q; ++j) {
                const double p2 = p1;
                p1 = p0;
                p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p2) / j;
            }
            derivative = q * (z * p0 - p1) / (z * z - 1.0);
            const double step = p0 / derivative;
            z -= step;
            if (std::fabs(step) < 1e-16) break;
        }
        points[i] = -z;
        points[q - 1 - i] = z;
        weights[i] = weights[q - 1 - i] = 2.0 / ((1.0 - z * z) * derivative * derivative);
    }
}

// Element nodes of order p on [0, 1]: (1 - cos(i pi / p)) / 2, increasing.
std::vector<double> elementNodes(int p) {
    std::vector<double> nodes(p + 1);
    const double pi = 3.14159265358979323846;
    for (int i = 0; i <= p; ++i) nodes[i] = 0.5 * (1.0 - std::cos(i * pi / p));
    nodes[0] = 0.0;
    nodes[p] = 1.0;
    if (p % 2 == 0) nodes[p / 2] = 0.5;
    return nodes;
}

void lagrange(const std::vector<double>& nodes, double x, std::vector<double>& value, std::vector<double>& slope) {
    const int n = int(nodes.size());
    value.assign(n, 0.0);
    slope.assign(n, 0.0);
    for (int a = 0; a < n; ++a) {
        double v = 1.0;
        for (int j = 0; j < n; ++j)
            if (j != a) v *= (x - nodes[j]) / (nodes[a] - nodes[j]);
        double s = 0.0;
        for (int k = 0; k < n; ++k) {
            if (k == a) continue;
            double term = 1.0 / (nodes[a] - nodes[k]);
            for (int j = 0; j < n; ++j)
                if (j != a && j != k) term *= (x - nodes[j]) / (nodes[a] - nodes[j]);
            s += term;
        }
        value[a] = v;
        slope[a] = s;
    }
}

// Legendre polynomials P_0..P_n at x (three-term recurrence).
void legendre(int n, double x, std::vector<double>& values) {
    values.assign(size_t(n) + 1, 0.0);
    values[0] = 1.0;
    if (n >= 1) values[1] = x;
    for (int k = 2; k <= n; ++k) values[k] = ((2.0 * k - 1.0) * x * values[k - 1] - (k - 1.0) * values[k - 2]) / k;
}

// Integrated Legendre basis of order p in position order on t in [0, 1]
// (x = 2t - 1): position 0 holds N_0, position p holds N_1, position l
// (0 < l < p) holds N_{l+1}. Slopes are d/dt = 2 d/dx.
void integratedLegendre(int p, double t, std::vector<double>& value, std::vector<double>& slope) {
    const double x = 2.0 * t - 1.0;
    std::vector<double> P;
    legendre(p, x, P);
    value.assign(size_t(p) + 1, 0.0);
    slope.assign(size_t(p) + 1, 0.0);
    value[0] = 0.5 * (1.0 - x);
    slope[0] = -1.0;
    value[size_t(p)] = 0.5 * (1.0 + x);
    slope[size_t(p)] = 1.0;
    for (int l = 1; l < p; ++l) {
        const int k = l + 1;
        value[size_t(l)] = (P[size_t(k)] - P[size_t(k) - 2]) / (2.0 * k - 1.0);
        slope[size_t(l)] = 2.0 * P[size_t(k) - 1];
    }
}

// stiffness[a*(p+1)+b] = int_0^1 phi_a' phi_b',  mass[...] = int_0^1 phi_a phi_b.
void referenceMatrices(int p, int basis, std::vector<double>& stiffness, std::vector<double>& mass) {
    const int n = p + 1;
    const std::vector<double> nodes = elementNodes(p);
    std::vector<double> points, weights;
    gaussLegendre(p + 1, points, weights);
    stiffness.assign(n * n, 0.0);
    mass.assign(n * n, 0.0);
    std::vector<double> value, slope;
    for (size_t q = 0; q < points.size(); ++q) {
        const double x = 0.5 * (points[q] + 1.0);
        const double w = 0.5 * weights[q];
        if (basis == CIE_SYS_BASIS_LEGENDRE) integratedLegendre(p, x, value, slope);
        else lagrange(nodes, x, value, slope);
        for (int a = 0; a < n; ++a)
            for (int b = 0; b < n; ++b) {
                stiffness[a * n + b] += w * slope[a] * slope[b];
                mass[a * n + b] += w * value[a] * value[b];
            }
    }
}

struct Lagrange {
    int dim = 3, order = 2, elements = 1, eliminate = 1, basis = CIE_SYS_BASIS_LAGRANGE;
    int nodesPerDim = 0;       // global nodes per dimension, order*elements + 1
    int first = 0, last = 0;   // first/last retained node index per dimension
    int retained = 0;          // retained nodes per dimension
    int npe = 0;               // nodes per element
    std::vector<double> elementMatrix;  // npe x npe, k_e = 1

    bool init(const cie_sys_lagrange_spec* spec) {
        if (!spec) return false;
        dim = spec->dim;
        order = spec->order;
        elements = spec->elements;
        eliminate = spec->eliminate ? 1 : 0;
        basis = spec->basis;
        if ((dim != 2 && dim != 3) || order < 1 || order > 6 || elements < 1) return false;
        if (basis != CIE_SYS_BASIS_LAGRANGE && basis != CIE_SYS_BASIS_LEGENDRE) return false;
        nodesPerDim = order * elements + 1;
        first = eliminate ? 1 : 0;
        last = eliminate ? nodesPerDim - 2 : nodesPerDim - 1;
        retained = last - first + 1;
        if (retained < 1) return false;
        npe = 1;
        for (int d = 0; d < dim; ++d) npe *= order + 1;
        buildElementMatrix();
        return true;
    }

    void buildElementMatrix() {
        std::vector<double> k1, m1;
        referenceMatrices(order, basis, k1, m1);
        const int n = order + 1;
        const double h = 1.0 / elements;
        // Physical 1D blocks: K1 = k1 / h, M1 = h m1, so the d-dimensional
        // element matrix sum_d (K1 (x) M1 ...) carries the factor h^(d-2).
        const double scale = dim == 3 ? h : 1.0;
        elementMatrix.assign(size_t(npe) * npe, 0.0);
        for (int a = 0; a < npe; ++a) {
            int ai[3] = {a % n, (a / n) % n, dim == 3 ? a / (n * n) : 0};
            for (int b = 0; b < npe; ++b) {
                int bi[3] = {b % n, (b / n) % n, dim == 3 ? b / (n * n) : 0};
                double sum = 0.0;
                for (int d = 0; d < dim; ++d) {
                    double term = 1.0;
                    for (int e = 0; e < dim; ++e)
                        term *= (e == d ? k1 : m1)[ai[e] * n + bi[e]];
                    sum += term;
                }
                elementMatrix[size_t(a) * npe + b] = scale * sum;
            }
        }
    }

    int64_t rowCount() const {
        int64_t rows = 1;
        for (int d = 0; d < dim; ++d) rows *= retained;
        return rows;
    }

    int64_t elementCount() const {
        int64_t count = 1;
        for (int d = 0; d < dim; ++d) count *= elements;
        return count;
    }

    // Range of nodes coupled to node i along one dimension.
    void coupled(int i, int& lo, int& hi) const {
        if (i % order == 0) {
            lo = i - order;
            hi = i + order;
        } else {
            lo = order * (i / order);
            hi = lo + order;
        }
        lo = std::max(lo, first);
        hi = std::min(hi, last);
    }

    // Elements along one dimension holding both nodes i and j.
    void shared(int i, int j, int& lo, int& hi) const {
        const int upper = std::max(i, j), lower = std::min(i, j);
        // ceil((upper - order) / order) for upper >= 1, else 0
        lo = upper > 0 ? (upper - 1) / order : 0;
        hi = std::min(elements - 1, lower / order);
    }

    int64_t rowLength(int64_t row) const {
        int idx[3];
        unpack(row, idx);
        int64_t length = 1;
        for (int d = 0; d < dim; ++d) {
            int lo, hi;
            coupled(idx[d], lo, hi);
            length *= hi - lo + 1;
        }
        return length;
    }

    void unpack(int64_t row, int* idx) const {
        idx[0] = first + int(row % retained);
        idx[1] = first + int((row / retained) % retained);
        idx[2] = dim == 3 ? first + int(row / (int64_t(retained) * retained)) : 0;
    }

    int64_t pack(const int* idx) const {
        int64_t row = idx[0] - first;
        row += int64_t(retained) * (idx[1] - first);
        if (dim == 3) row += int64_t(retained) * retained * (idx[2] - first);
        return row;
    }

    void fillRow(int64_t row, const double* conductivity, int32_t* cols, double* vals) const {
        int idx[3];
        unpack(row, idx);
        int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
        for (int d = 0; d < dim; ++d) coupled(idx[d], lo[d], hi[d]);
        const int n = order + 1;
        int64_t k = 0;
        for (int jz = lo[2]; jz <= hi[2]; ++jz)
            for (int jy = lo[1]; jy <= hi[1]; ++jy)
                for (int jx = lo[0]; jx <= hi[0]; ++jx) {
                    const int jdx[3] = {jx, jy, jz};
                    int elo[3] = {0, 0, 0}, ehi[3] = {0, 0, 0};
                    for (int d = 0; d < dim; ++d) shared(idx[d], jdx[d], elo[d], ehi[d]);
                    double value = 0.0;
                    for (int ez = elo[2]; ez <= ehi[2]; ++ez)
                        for (int ey = elo[1]; ey <= ehi[1]; ++ey)
                            for (int ex = elo[0]; ex <= ehi[0]; ++ex) {
                                int a = idx[0] - order * ex + n * (idx[1] - order * ey);
                                int b = jx - order * ex + n * (jy - order * ey);
                                int64_t element = ex + int64_t(elements) * ey;
                                if (dim == 3) {
                                    a += n * n * (idx[2] - order * ez);
                                    b += n * n * (jz - order * ez);
                                    element += int64_t(elements) * elements * ez;
                                }
                                const double kappa = conductivity ? conductivity[element] : 1.0;
                                value += kappa * elementMatrix[size_t(a) * npe + b];
                            }
                    cols[k] = int32_t(pack(jdx));
                    vals[k] = value;
                    ++k;
                }
    }
};

}  // namespace

extern "C" {

const char* cie_sys_last_error(void) { return g_sysError.c_str(); }

int cie_sys_splitmix64(uint64_t seed, int64_t count, uint64_t* out) {
    if (count < 0 || (count && !out)) return fail("cie_sys_splitmix64: invalid arguments");
    SplitMix64 rng{seed};
    for (int64_t i = 0; i < count; ++i) out[i] = rng.next();
    return 0;
}

int cie_sys_uniform(uint64_t seed, int64_t count, double* out) {
    if (count < 0 || (count && !out)) return fail("cie_sys_uniform: invalid arguments");
    SplitMix64 rng{seed};
    for (int64_t i = 0; i < count; ++i) out[i] = 2.0 * rng.unit() - 1.0;
    return 0;
}

int cie_sys_normal(uint64_t seed, double sigma, int64_t count, double* out) {
    if (count < 0 || (count && !out)) return fail("cie_sys_normal: invalid arguments");
    SplitMix64 rng{seed};
    const double twoPi = 6.28318530717958647692;
    for (int64_t i = 0; i < count; ++i) {
        const double u1 = rng.unit();
        const double u2 = rng.unit();
        out[i] = sigma * std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(twoPi * u2);
    }
    return 0;
}

int cie_sys_poisson5_size(int m, int64_t* rows, int64_t* nnz) {
    if (m < 1 || !rows || !nnz) return fail("cie_sys_poisson5_size: invalid arguments");
    *rows = int64_t(m) * m;
    *nnz = 5 * int64_t(m) * m - 4 * int64_t(m);
    return 0;
}

int cie_sys_poisson5_fill(int m, int32_t* rowExtents, int32_t* columnIndices, double* entries) {
    if (m < 1 || !rowExtents || !columnIndices || !entries)
        return fail("cie_sys_poisson5_fill: invalid arguments");
    int64_t k = 0;
    rowExtents[0] = 0;
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < m; ++i) {
            const int64_t row = i + int64_t(m) * j;
            if (j > 0) { columnIndices[k] = int32_t(row - m); entries[k++] = -1.0; }
            if (i > 0) { columnIndices[k] = int32_t(row - 1); entries[k++] = -1.0; }
            columnIndices[k] = int32_t(row); entries[k++] = 4.0;
            if (i + 1 < m) { columnIndices[k] = int32_t(row + 1); entries[k++] = -1.0; }
            if (j + 1 < m) { columnIndices[k] = int32_t(row + m); entries[k++] = -1.0; }
            rowExtents[row + 1] = int32_t(k);
        }
    return 0;
}

int cie_sys_lagrange_size(const cie_sys_lagrange_spec* spec, int64_t* rows, int64_t* nnz,
                          int64_t* elementCount, int32_t* nodesPerElement) {
    Lagrange mesh;
    if (!mesh.init(spec)) return fail("cie_sys_lagrange_size: invalid spec");
    const int64_t rowCount = mesh.rowCount();
    int64_t total = 0;
    for (int64_t row = 0; row < rowCount; ++row) total += mesh.rowLength(row);
    if (rows) *rows = rowCount;
    if (nnz) *nnz = total;
    if (elementCount) *elementCount = mesh.elementCount();
    if (nodesPerElement) *nodesPerElement = mesh.npe;
    return 0;
}

int cie_sys_lagrange_fill(const cie_sys_lagrange_spec* spec, const double* conductivity,
                          int32_t* rowExtents, int32_t* columnIndices, double* entries, int threads) {
    Lagrange mesh;
    if (!mesh.init(spec)) return fail("cie_sys_lagrange_fill: invalid spec");
    if (!rowExtents || !columnIndices || !entries) return fail("cie_sys_lagrange_fill: null output");
    const int64_t rows = mesh.rowCount();
    std::vector<int64_t> offsets(rows + 1, 0);
    const unsigned workers = threadCount(threads);
    parallelRanges(rows, workers, [&](int64_t begin, int64_t end) {
        for (int64_t row = begin; row < end; ++row) offsets[row + 1] = mesh.rowLength(row);
    });
    for (int64_t row = 0; row < rows; ++row) offsets[row + 1] += offsets[row];
    if (offsets[rows] > INT32_MAX) return fail("cie_sys_lagrange_fill: more than 2^31-1 nonzeros");
    for (int64_t row = 0; row <= rows; ++row) rowExtents[row] = int32_t(offsets[row]);
    parallelRanges(rows, workers, [&](int64_t begin, int64_t end) {
        for (int64_t row = begin; row < end; ++row)
            mesh.fillRow(row, conductivity, columnIndices + offsets[row], entries + offsets[row]);
    });
    return 0;
}

int cie_sys_lagrange_element_matrix(const cie_sys_lagrange_spec* spec, double* elementMatrix) {
    Lagrange mesh;
    if (!mesh.init(spec)) return fail("cie_sys_lagrange_element_matrix: invalid spec");
    if (!elementMatrix) return fail("cie_sys_lagrange_element_matrix: null output");
    std::memcpy(elementMatrix, mesh.elementMatrix.data(), mesh.elementMatrix.size() * sizeof(double));
    return 0;
}

int cie_sys_lagrange_element_dofs(const cie_sys_lagrange_spec* spec, int32_t* dofs) {
    Lagrange mesh;
    if (!mesh.init(spec)) return fail("cie_sys_lagrange_element_dofs: invalid spec");
    if (!dofs) return fail("cie_sys_lagrange_element_dofs: null output");
    const int n = mesh.order + 1;
    const int64_t elements = mesh.elementCount();
    for (int64_t e = 0; e < elements; ++e) {
        const int ex = int(e % mesh.elements);
        const int ey = int((e / mesh.elements) % mesh.elements);
        const int ez = mesh.dim == 3 ? int(e / (int64_t(mesh.elements) * mesh.elements)) : 0;
        for (int a = 0; a < mesh.npe; ++a) {
            int idx[3] = {mesh.order * ex + a % n, mesh.order * ey + (a / n) % n,
                          mesh.dim == 3 ? mesh.order * ez + a / (n * n) : 0};
            bool kept = true;
            for (int d = 0; d < mesh.dim; ++d) kept = kept && mesh.first <= idx[d] && idx[d] <= mesh.last;
            dofs[e * mesh.npe + a] = kept ? int32_t(mesh.pack(idx)) : -1;
        }
    }
    return 0;
}

int cie_sys_lagrange_gradient_moments(const cie_sys_lagrange_spec* spec, double* moments) {
    Lagrange mesh;
    if (!mesh.init(spec)) return fail("cie_sys_lagrange_gradient_moments: invalid spec");
    if (!moments) return fail("cie_sys_lagrange_gradient_moments: null output");
    // 1D values and slopes on [-1, 1] at p + 1 Gauss-Legendre points (d/dxi = (d/dt) / 2 for t in [0, 1]).
    const int n = mesh.order + 1, dim = mesh.dim, npe = mesh.npe;
    std::vector<double> points, weights, value, slope;
    gaussLegendre(n, points, weights);
    const std::vector<double> nodes = elementNodes(mesh.order);
    std::vector<double> v1(size_t(n) * n), d1(size_t(n) * n);  // [point][local position]
    for (int q = 0; q < n; ++q) {
        const double t = 0.5 * (points[size_t(q)] + 1.0);
        if (mesh.basis == CIE_SYS_BASIS_LEGENDRE) integratedLegendre(mesh.order, t, value, slope);
        else lagrange(nodes, t, value, slope);
        for (int a = 0; a < n; ++a) {
            v1[size_t(q) * n + a] = value[size_t(a)];
            d1[size_t(q) * n + a] = 0.5 * slope[size_t(a)];
        }
    }
    int pointCount = 1;
    for (int d = 0; d < dim; ++d) pointCount *= n;
    std::vector<double> gradient(size_t(dim) * npe);
    std::fill(moments, moments + size_t(dim) * dim * npe * npe, 0.0);
    for (int q = 0; q < pointCount; ++q) {
        const int qi[3] = {q % n, (q / n) % n, dim == 3 ? q / (n * n) : 0};
        double weight = 1.0;
        for (int d = 0; d < dim; ++d) weight *= weights[size_t(qi[d])];
        for (int a = 0; a < npe; ++a) {
            const int ai[3] = {a % n, (a / n) % n, dim == 3 ? a / (n * n) : 0};
            for (int i = 0; i < dim; ++i) {
                double g = 1.0;
                for (int d = 0; d < dim; ++d)
                    g *= (d == i ? d1 : v1)[size_t(qi[d]) * n + ai[d]];
                gradient[size_t(i) * npe + a] = g;
            }
        }
        for (int i = 0; i < dim; ++i)
            for (int j = 0; j < dim; ++j) {
                double* block = moments + (size_t(i) * dim + j) * npe * npe;
                for (int a = 0; a < npe; ++a)
                    for (int b = 0; b < npe; ++b)
                        block[size_t(a) * npe + b] += weight * gradient[size_t(i) * npe + a] * gradient[size_t(j) * npe + b];
            }
    }
    return 0;
}

int cie_sys_lagrange_ansatz_mask(const cie_sys_lagrange_spec* spec, uint16_t* mask) {
    Lagrange mesh;
    if (!mesh.init(spec)) return fail("cie_sys_lagrange_ansatz_mask: invalid spec");
    if (mesh.basis != CIE_SYS_BASIS_LEGENDRE)
        return fail("cie_sys_lagrange_ansatz_mask: the mask is defined for the hierarchical (Legendre) basis");
    if (!mask) return fail("cie_sys_lagrange_ansatz_mask: null output");
    const int64_t rows = mesh.rowCount();
    for (int64_t row = 0; row < rows; ++row) {
        int idx[3];
        mesh.unpack(row, idx);
        int degree = 1;  // vertex functions N_0, N_1 count as degree 1
        for (int d = 0; d < mesh.dim; ++d) {
            const int local = idx[d] % mesh.order;
            if (local != 0) degree = std::max(degree, local + 1);  // N_{local+1}
        }
        mask[row] = uint16_t(degree);
    }
    return 0;
}

int cie_sys_csr_slots(int64_t rows, const int32_t* rowExtents, const int32_t* columnIndices,
                      int64_t pairCount, const int32_t* pairRows, const int32_t* pairColumns, int64_t* slots) {
    if (rows < 0 || !rowExtents || !columnIndices || pairCount < 0 || (pairCount && (!pairRows || !pairColumns || !slots)))
        return fail("cie_sys_csr_slots: invalid arguments");
    for (int64_t k = 0; k < pairCount; ++k) {
        const int32_t row = pairRows[k], col = pairColumns[k];
        if (row < 0 || col < 0 || row >= rows) { slots[k] = -1; continue; }
        const int32_t* begin = columnIndices + rowExtents[row];
        const int32_t* end = columnIndices + rowExtents[row + 1];
        const int32_t* it = std::lower_bound(begin, end, col);
        slots[k] = (it != end && *it == col) ? int64_t(it - columnIndices) : -1;
    }
    return 0;
}

}  // extern "C"

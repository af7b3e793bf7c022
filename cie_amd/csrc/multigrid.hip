// Masked operators, Jacobi smoothers and the p-multigrid cycle of the
// reference's other iterative path (linalg/packages/solvers/src:
// MaskedCSROperator, MaskedIdentityOperator, JacobiOperator,
// MaskedJacobiOperator, ApproximateSubstitution; the cycle of
// executables/fem_examples/src/poisson2D/solver.cpp:111-232).
//
// A masked CSR operator (rows and columns with mask >= threshold left out of
// the products) is built once as a matrix of its own: a device pass compacts
// the kept entries, so every masked product afterwards is an ordinary SpMV on
// the block-tile kernel, and a masked row is an empty row (sum 0), which is
// the reference's contribution for it. The coarsest multigrid grid is the same
// compaction with the kept rows and columns renumbered, solved by the fused
// Jacobi-PCG. A Jacobi sweep is one SpMV whose epilogue forms
// x + (w (b - A x)) / d row by row, with the roundings of
// JacobiOperator.cpp:52-57 (the residual, then DiagonalOperator's scale and
// add, DiagonalOperator.cpp:37-53).
#include "dispatch.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

using namespace ciehip;

struct cie_hip_mask_s {
    Holds holds;
    cie_hip_context context = nullptr;
    int64_t size = 0;
    uint16_t* values = nullptr;     // device
    std::vector<uint16_t> host;
    uint16_t largest = 0;
};

namespace {

constexpr int kRowGroup = 8;  // lanes per row of the substitution kernels

// ---- compaction -------------------------------------------------------------

// Kept entries of row i of the result: row source[i] (or i) of the input, if
// that row is kept, with the columns whose renumber[] is not negative.
__global__ __launch_bounds__(kReduceBlock) void compactCountKernel(
    int rows, const int32_t* __restrict__ source, const int32_t* __restrict__ renumber,
    const int32_t* __restrict__ extents, const int32_t* __restrict__ columns, int32_t* __restrict__ counts) {
    for (int i = int(blockIdx.x) * kReduceBlock + int(threadIdx.x); i < rows; i += int(gridDim.x) * kReduceBlock) {
        const int old = source ? source[i] : i;
        int count = 0;
        if (renumber[old] >= 0)
            for (int k = extents[old]; k < extents[old + 1]; ++k) count += renumber[columns[k]] >= 0;
        counts[i] = count;
    }
}

__global__ __launch_bounds__(kReduceBlock) void compactCopyKernel(
    int rows, const int32_t* __restrict__ source, const int32_t* __restrict__ renumber,
    const int32_t* __restrict__ extents, const int32_t* __restrict__ columns, const double* __restrict__ entries64,
    const float* __restrict__ entries32, const int32_t* __restrict__ newExtents, int32_t* __restrict__ newColumns,
    double* __restrict__ newEntries64, float* __restrict__ newEntries32) {
    for (int i = int(blockIdx.x) * kReduceBlock + int(threadIdx.x); i < rows; i += int(gridDim.x) * kReduceBlock) {
        const int old = source ? source[i] : i;
        if (renumber[old] < 0) continue;
        int out = newExtents[i];
        for (int k = extents[old]; k < extents[old + 1]; ++k) {
            const int column = renumber[columns[k]];
            if (column < 0) continue;
            newColumns[out] = column;
            newEntries64[out] = entries64[k];
            if (newEntries32) newEntries32[out] = entries32[k];
            ++out;
        }
    }
}

// A new matrix of the kept rows and entries of m. source (new row -> old row)
// may be empty (all rows, same numbering); renumber maps old columns to new
// ones, -1 for dropped columns and rows.
int compactMatrix(cie_hip_csr m, const std::vector<int32_t>& source, const std::vector<int32_t>& renumber,
                  int64_t newColumnCount, cie_hip_csr* out) {
    cie_hip_context ctx = m->context;
    CIE_HIP_CHECK(hipSetDevice(ctx->device));
    const int64_t rows = source.empty() ? m->rows : int64_t(source.size());
    int32_t *dSource = nullptr, *dRenumber = nullptr, *dCounts = nullptr;
    int32_t *dExtents = nullptr, *dColumns = nullptr;
    double* dEntries64 = nullptr;
    float* dEntries32 = nullptr;
    auto release = [&](bool all) {
        if (dSource) (void)hipFree(dSource);
        if (dRenumber) (void)hipFree(dRenumber);
        if (dCounts) (void)hipFree(dCounts);
        if (all) {
            if (dExtents) (void)hipFree(dExtents);
            if (dColumns) (void)hipFree(dColumns);
            if (dEntries64) (void)hipFree(dEntries64);
            if (dEntries32) (void)hipFree(dEntries32);
        }
    };
    bool ok = (source.empty() || hipMalloc(&dSource, sizeof(int32_t) * source.size()) == hipSuccess) &&
              hipMalloc(&dRenumber, sizeof(int32_t) * std::max<size_t>(1, renumber.size())) == hipSuccess &&
              hipMalloc(&dCounts, sizeof(int32_t) * size_t(std::max<int64_t>(1, rows))) == hipSuccess &&
              hipMalloc(&dExtents, sizeof(int32_t) * size_t(rows + 1)) == hipSuccess;
    ok = ok && (source.empty() || hipMemcpyAsync(dSource, source.data(), sizeof(int32_t) * source.size(),
                                                 hipMemcpyHostToDevice, ctx->stream) == hipSuccess) &&
         (renumber.empty() || hipMemcpyAsync(dRenumber, renumber.data(), sizeof(int32_t) * renumber.size(),
                                             hipMemcpyHostToDevice, ctx->stream) == hipSuccess);
    std::vector<int32_t> counts(static_cast<size_t>(rows)), newExtents(static_cast<size_t>(rows) + 1, 0);
    if (ok && rows) {
        hipLaunchKernelGGL(compactCountKernel, dim3(vectorGrid(rows)), dim3(kReduceBlock), 0, ctx->stream, int(rows),
                           dSource, dRenumber, m->rowExtents, m->columnIndices, dCounts);
        ok = hipGetLastError() == hipSuccess &&
             hipMemcpyAsync(counts.data(), dCounts, sizeof(int32_t) * size_t(rows), hipMemcpyDeviceToHost,
                            ctx->stream) == hipSuccess &&
             hipStreamSynchronize(ctx->stream) == hipSuccess;
    }
    if (!ok) {
        release(true);
        return setError(CIE_HIP_ERROR_HIP, "masked matrix: counting the kept entries failed");
    }
    for (int64_t i = 0; i < rows; ++i) newExtents[size_t(i) + 1] = newExtents[size_t(i)] + counts[size_t(i)];
    const int64_t nnz = newExtents[size_t(rows)];
    ok = hipMalloc(&dColumns, sizeof(int32_t) * size_t(std::max<int64_t>(1, nnz))) == hipSuccess &&
         hipMalloc(&dEntries64, sizeof(double) * size_t(std::max<int64_t>(1, nnz))) == hipSuccess &&
         (!m->entries32 || hipMalloc(&dEntries32, sizeof(float) * size_t(std::max<int64_t>(1, nnz))) == hipSuccess) &&
         hipMemcpyAsync(dExtents, newExtents.data(), sizeof(int32_t) * size_t(rows + 1), hipMemcpyHostToDevice,
                        ctx->stream) == hipSuccess;
    if (ok && rows) {
        hipLaunchKernelGGL(compactCopyKernel, dim3(vectorGrid(rows)), dim3(kReduceBlock), 0, ctx->stream, int(rows),
                           dSource, dRenumber, m->rowExtents, m->columnIndices, m->entries64, m->entries32, dExtents,
                           dColumns, dEntries64, dEntries32);
        ok = hipGetLastError() == hipSuccess;
    }
    ok = ok && hipStreamSynchronize(ctx->stream) == hipSuccess;
    release(!ok);
    if (!ok) return setError(CIE_HIP_ERROR_HIP, "masked matrix: copying the kept entries failed");
    return adoptMatrix(ctx, rows, newColumnCount, dExtents, dColumns, dEntries64, dEntries32, std::move(newExtents),
                       out);
}

// ---- elementwise ------------------------------------------------------------

// out = op(out, right) for one (inScale, outScale) class (the documented
// contract; exact wherever the reference's own case table is exact).
__device__ __forceinline__ double applyScales(double inScale, double outScale, double out, double right) {
    if (inScale == 0.0) return outScale == 1.0 ? right : outScale == -1.0 ? -right : outScale * right;
    if (inScale == 1.0) return outScale == 1.0 ? out + right : outScale == -1.0 ? out - right : out + outScale * right;
    if (inScale == -1.0) return outScale == 1.0 ? right - out : outScale == -1.0 ? -right - out : outScale * right - out;
    return outScale == 1.0 ? inScale * out + right : outScale == -1.0 ? inScale * out - right
                                                                      : inScale * out + outScale * right;
}

template <class T>
__global__ __launch_bounds__(kReduceBlock) void maskedIdentityKernel(int64_t n, const uint16_t* __restrict__ mask,
                                                                     uint16_t threshold, T inScale,
                                                                     const T* __restrict__ in, T outScale,
                                                                     T* __restrict__ out) {
    for (int64_t i = int64_t(blockIdx.x) * kReduceBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kReduceBlock)
        out[i] = T(applyScales(double(inScale), double(outScale), double(out[i]),
                               mask[i] < threshold ? double(in[i]) : 0.0));
}

// First Jacobi sweep from x = 0: x = 0 + (w (b - 0)) / d, with b = the masked
// right-hand side (mask may be null: nothing masked).
__global__ __launch_bounds__(kReduceBlock) void jacobiStartKernel(int64_t n, const uint16_t* __restrict__ mask,
                                                                  uint16_t threshold, const double* __restrict__ b,
                                                                  const double* __restrict__ inverse, double relaxation,
                                                                  double* __restrict__ x) {
    for (int64_t i = int64_t(blockIdx.x) * kReduceBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kReduceBlock) {
        const double rhs = (!mask || mask[i] < threshold) ? b[i] : 0.0;
        x[i] = 0.0 + relaxation * (rhs - 0.0) * inverse[i];
    }
}

// out = outScale x + inScale memory (JacobiOperator.cpp:59-60: scale, then add).
__global__ __launch_bounds__(kReduceBlock) void jacobiFinishKernel(int64_t n, const double* __restrict__ x,
                                                                   const double* __restrict__ memory, double inScale,
                                                                   double outScale, double* __restrict__ out) {
    for (int64_t i = int64_t(blockIdx.x) * kReduceBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kReduceBlock) {
        const double scaled = x[i] * outScale;
        out[i] = inScale == 1.0 ? scaled + memory[i] : inScale == -1.0 ? scaled - memory[i] : scaled + inScale * memory[i];
    }
}

__global__ __launch_bounds__(kReduceBlock) void gatherKernel(int64_t n, const int32_t* __restrict__ index,
                                                             const double* __restrict__ in, double* __restrict__ out) {
    for (int64_t i = int64_t(blockIdx.x) * kReduceBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kReduceBlock)
        out[i] = in[index[i]];
}

// ---- SpMV-based sweeps ------------------------------------------------------

// x' = x + (w (b - A x)) / d for every row; b masked as in jacobiStartKernel.
template <class Walker>
__global__ __launch_bounds__(Walker::kBlock) void jacobiSweepKernel(RowSource source, const double* __restrict__ entries,
                                                                   const double* __restrict__ x,
                                                                   const uint16_t* __restrict__ mask, uint16_t threshold,
                                                                   const double* __restrict__ b,
                                                                   const double* __restrict__ inverse, double relaxation,
                                                                   double* __restrict__ next) {
    Walker::run(source, entries, x, [&](int row, double sum) {
        const double rhs = (!mask || mask[row] < threshold) ? b[row] : 0.0;
        next[row] = x[row] + relaxation * (rhs - sum) * inverse[row];
    });
}

// residual -= A update and solution += update, row by row.
template <class Walker>
__global__ __launch_bounds__(Walker::kBlock) void applyUpdateKernel(RowSource source, const double* __restrict__ entries,
                                                                   const double* __restrict__ update,
                                                                   double* __restrict__ residual,
                                                                   double* __restrict__ solution) {
    Walker::run(source, entries, update, [&](int row, double sum) {
        residual[row] = residual[row] - sum;
        solution[row] = solution[row] + update[row];
    });
}

// The same on a compacted grid: grid row i is row keep[i].
template <class Walker>
__global__ __launch_bounds__(Walker::kBlock) void applyUpdateScatterKernel(RowSource source,
                                                                          const double* __restrict__ entries,
                                                                          const double* __restrict__ update,
                                                                          const int32_t* __restrict__ keep,
                                                                          double* __restrict__ residual,
                                                                          double* __restrict__ solution) {
    Walker::run(source, entries, update, [&](int row, double sum) {
        const int32_t i = keep[row];
        residual[i] = residual[i] - sum;
        solution[i] = solution[i] + update[row];
    });
}

template <template <class> class Kernel, class... Args>
int launchSweep(cie_hip_csr m, Args... args) {
    int status = CIE_HIP_OK;
    const bool ok = withWalker<double, double>(m->launch64, [&](auto walker) {
        using Walker = decltype(walker);
        RowSource source;
        int grid = 0;
        status = prepareLaunch(m, m->launch64, Kernel<Walker>::kernel, source, grid);
        if (status != CIE_HIP_OK) return;
        hipLaunchKernelGGL(Kernel<Walker>::kernel, dim3(grid), dim3(Walker::kBlock), 0, m->context->stream, source,
                           m->entries64, args...);
    });
    if (!ok) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "unsupported SpMV launch shape");
    if (status != CIE_HIP_OK) return status;
    CIE_HIP_CHECK(hipGetLastError());
    return CIE_HIP_OK;
}

template <class Walker>
struct JacobiSweep {
    static constexpr auto kernel = jacobiSweepKernel<Walker>;
};
template <class Walker>
struct ApplyUpdate {
    static constexpr auto kernel = applyUpdateKernel<Walker>;
};
template <class Walker>
struct ApplyUpdateScatter {
    static constexpr auto kernel = applyUpdateScatterKernel<Walker>;
};

// MaskedJacobiOperator (masked) and ApproximateSubstitution sweeps from
// `previous` into out: a group of kRowGroup lanes per row sums the kept
// off-diagonal products; the group's first lane applies the formula.
__global__ __launch_bounds__(kReduceBlock) void substitutionKernel(
    int rows, const int32_t* __restrict__ extents, const int32_t* __restrict__ columns,
    const double* __restrict__ entries, const uint16_t* __restrict__ mask, uint16_t threshold, int masked,
    double relaxation, double inScale, const double* __restrict__ in, double outScale,
    const double* __restrict__ previous, double* __restrict__ out, int* __restrict__ missing) {
    const int lane = int(threadIdx.x) % kRowGroup;
    const int row = (int(blockIdx.x) * kReduceBlock + int(threadIdx.x)) / kRowGroup;
    if (row >= rows) return;  // whole groups leave together
    const bool kept = !masked || mask[row] < threshold;
    double contribution = 0.0;
    int diagonal = -1;
    if (kept)
        for (int k = extents[row] + lane; k < extents[row + 1]; k += kRowGroup) {
            const int column = columns[k];
            if (column == row) diagonal = k;
            else if (!masked || mask[column] < threshold) contribution += entries[k] * previous[column];
        }
#pragma unroll
    for (int offset = kRowGroup / 2; offset > 0; offset /= 2) {
        contribution += __shfl_xor(contribution, offset, kRowGroup);
        diagonal = max(diagonal, __shfl_xor(diagonal, offset, kRowGroup));
    }
    if (lane != 0) return;
    double value = 0.0;
    if (kept) {
        if (diagonal < 0) {
            atomicOr(missing, 1);
            return;
        }
        const double d = entries[diagonal];
        value = masked ? relaxation * (in[row] - contribution) / d + (1.0 - relaxation) * previous[row]
                       : previous[row] + (in[row] - contribution) / d;
    }
    out[row] = applyScales(inScale, outScale, out[row], value);
}

int checkVectors(cie_hip_csr m, int64_t size, const char* name) {
    if (size != m->rows || m->rows != m->columns)
        return setError(CIE_HIP_ERROR_SHAPE, std::string(name) + ": vector size does not match the square matrix");
    return CIE_HIP_OK;
}

int substitution(cie_hip_context ctx, cie_hip_csr m, const cie_hip_mask_s* mask, uint16_t threshold, uint64_t iterations,
                 double relaxation, double inScale, const double* in, double outScale, double* out, int64_t size,
                 const char* name) {
    if (!ctx || !m || (size && (!in || !out)))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, std::string(name) + ": invalid arguments");
    CIE_HIP_REQUIRE_CONTEXT(ctx, m->context, name);
    if (mask) CIE_HIP_REQUIRE_CONTEXT(ctx, mask->context, name);
    if (const int s = checkVectors(m, size, name)) return s;
    if (mask && mask->size != size) return setError(CIE_HIP_ERROR_SHAPE, std::string(name) + ": mask size differs");
    if (inScale != 0.0 && iterations != 1)  // MaskedJacobiOperator.cpp:64-65, ApproximateSubstitution.cpp:40-41
        return setError(CIE_HIP_ERROR_NOT_IMPLEMENTED, std::string(name) + ": inScale != 0 needs exactly one sweep");
    if (!size || !iterations) return CIE_HIP_OK;
    CIE_HIP_CHECK(hipSetDevice(ctx->device));
    void* workspace = nullptr;
    if (const int s = matrixWorkspace(m, sizeof(double) * size_t(size), &workspace)) return s;
    double* previous = static_cast<double*>(workspace);
    int* missing = reinterpret_cast<int*>(ctx->scratch.scalars + 28);
    int* pinned = reinterpret_cast<int*>(ctx->scratch.pinned + 28);
    CIE_HIP_CHECK(hipMemsetAsync(missing, 0, sizeof(int), ctx->stream));
    const unsigned grid = unsigned((size * kRowGroup + kReduceBlock - 1) / kReduceBlock);
    for (uint64_t k = 0; k < iterations; ++k) {
        CIE_HIP_CHECK(hipMemcpyAsync(previous, out, sizeof(double) * size_t(size), hipMemcpyDeviceToDevice, ctx->stream));
        hipLaunchKernelGGL(substitutionKernel, dim3(grid), dim3(kReduceBlock), 0, ctx->stream, int(size), m->rowExtents,
                           m->columnIndices, m->entries64, mask ? mask->values : nullptr, threshold, mask ? 1 : 0,
                           relaxation, inScale, in, outScale, previous, out, missing);
        CIE_HIP_CHECK(hipGetLastError());
    }
    CIE_HIP_CHECK(hipMemcpyAsync(pinned, missing, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (*pinned) return setError(CIE_HIP_ERROR_MISSING_DIAGONAL, std::string(name) + ": a kept row has no diagonal entry");
    return CIE_HIP_OK;
}

}  // namespace

// ---- p-multigrid ------------------------------------------------------------

struct cie_hip_pmg_s {
    cie_hip_csr matrix = nullptr;
    cie_hip_mask_s* mask = nullptr;   // held by the plan
    int order = 0;
    std::vector<cie_hip_csr> grids;   // grid g: the matrix masked to mask < g + 2 (owned unless it is `matrix`);
                                      // for g >= 1 its kept rows and columns renumbered (keep[g] maps them back)
    std::vector<int32_t*> keep;       // g >= 1, compacted grids: grid row -> row (null for grid 0 and the matrix)
    std::vector<int64_t> keepRows;    // rows of each compacted grid
    std::vector<double*> keepInverse; // the inverse diagonal on each compacted grid's rows (gathered per solve)
    cie_hip_csr coarse = nullptr;      // mask < 2, rows and columns renumbered
    int64_t coarseRows = 0;
    int32_t* coarseIndex = nullptr;    // coarse row -> row
    double* work = nullptr;            // residual, update, x, x', coarse b, coarse x
};

namespace {

void releaseMask(cie_hip_mask mask) {
    cie_hip_context context = mask->context;
    (void)hipSetDevice(context->device);
    if (mask->values) (void)hipFree(mask->values);
    delete mask;
    dropContext(context);
}

void releasePmg(cie_hip_pmg plan) {
    if (!plan) return;
    for (cie_hip_csr g : plan->grids)
        if (g && g != plan->matrix && g != plan->coarse) cie_hip_csr_destroy(g);
    if (plan->coarse) cie_hip_csr_destroy(plan->coarse);
    if (plan->coarseIndex) (void)hipFree(plan->coarseIndex);
    for (int32_t* k : plan->keep)
        if (k) (void)hipFree(k);
    for (double* v : plan->keepInverse)
        if (v) (void)hipFree(v);
    if (plan->work) (void)hipFree(plan->work);
    cie_hip_csr matrix = plan->matrix;
    cie_hip_mask mask = plan->mask;
    delete plan;
    if (matrix) dropMatrix(matrix);
    if (mask && --mask->holds.holders == 0 && mask->holds.destroyRequested) releaseMask(mask);
}

}  // namespace

extern "C" {

int cie_hip_mask_create(cie_hip_context context, const uint16_t* values, int64_t size, cie_hip_mask* mask) {
    if (!mask) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_mask_create: null output");
    *mask = nullptr;
    if (!context || size < 0 || (size && !values))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_mask_create: invalid arguments");
    CIE_HIP_CHECK(hipSetDevice(context->device));
    auto* m = new cie_hip_mask_s;
    m->context = context;
    holdContext(context);
    m->size = size;
    m->host.assign(values, values + size);
    m->largest = size ? *std::max_element(values, values + size) : 0;
    if (hipMalloc(&m->values, sizeof(uint16_t) * size_t(std::max<int64_t>(1, size))) != hipSuccess ||
        (size && hipMemcpy(m->values, values, sizeof(uint16_t) * size_t(size), hipMemcpyHostToDevice) != hipSuccess)) {
        if (m->values) (void)hipFree(m->values);
        delete m;
        dropContext(context);
        return setError(CIE_HIP_ERROR_OUT_OF_MEMORY, "cie_hip_mask_create: upload failed");
    }
    *mask = m;
    return CIE_HIP_OK;
}

int cie_hip_mask_destroy(cie_hip_mask mask) {
    if (!mask || mask->holds.destroyRequested) return CIE_HIP_OK;
    if (mask->holds.holders > 0) {
        mask->holds.destroyRequested = true;  // released with the last multigrid plan over it
        return CIE_HIP_OK;
    }
    releaseMask(mask);
    return CIE_HIP_OK;
}

int cie_hip_csr_masked(cie_hip_csr matrix, cie_hip_mask mask, uint16_t threshold, cie_hip_csr* masked) {
    if (!masked) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_masked: null output");
    *masked = nullptr;
    if (!matrix || !mask) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_masked: null argument");
    CIE_HIP_REQUIRE_CONTEXT(mask->context, matrix->context, "cie_hip_csr_masked");
    if (matrix->rows != matrix->columns || mask->size != matrix->rows)
        return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_csr_masked: the mask needs one value per row of a square matrix");
    std::vector<int32_t> renumber(size_t(matrix->rows));
    for (int64_t i = 0; i < matrix->rows; ++i) renumber[size_t(i)] = mask->host[size_t(i)] < threshold ? int32_t(i) : -1;
    return compactMatrix(matrix, {}, renumber, matrix->columns, masked);
}

int cie_hip_masked_identity_f64(cie_hip_context context, cie_hip_mask mask, uint16_t threshold, double inScale,
                                const double* in, double outScale, double* out, int64_t size) {
    if (!context || !mask || (size && (!in || !out)))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_masked_identity_f64: invalid arguments");
    CIE_HIP_REQUIRE_CONTEXT(context, mask->context, "cie_hip_masked_identity_f64");
    if (mask->size != size) return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_masked_identity_f64: mask size differs");
    if (!size) return CIE_HIP_OK;
    CIE_HIP_CHECK(hipSetDevice(context->device));
    hipLaunchKernelGGL((maskedIdentityKernel<double>), dim3(vectorGrid(size)), dim3(kReduceBlock), 0, context->stream,
                       size, mask->values, threshold, inScale, in, outScale, out);
    CIE_HIP_CHECK(hipGetLastError());
    CIE_HIP_CHECK(hipStreamSynchronize(context->stream));
    return CIE_HIP_OK;
}

int cie_hip_jacobi_f64(cie_hip_context context, cie_hip_csr lhs, const double* inverseDiagonal, uint64_t iterations,
                       double relaxation, double inScale, const double* in, double outScale, double* out,
                       int64_t size) {
    if (!context || !lhs || (size && (!inverseDiagonal || !in || !out)))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_jacobi_f64: invalid arguments");
    CIE_HIP_REQUIRE_CONTEXT(context, lhs->context, "cie_hip_jacobi_f64");
    if (const int s = checkVectors(lhs, size, "cie_hip_jacobi_f64")) return s;
    if (!size) return CIE_HIP_OK;
    CIE_HIP_CHECK(hipSetDevice(context->device));
    void* workspace = nullptr;
    if (const int s = matrixWorkspace(lhs, sizeof(double) * 3 * size_t(size), &workspace)) return s;
    double* x = static_cast<double*>(workspace);
    double* next = x + size;
    double* memory = next + size;
    CIE_HIP_CHECK(hipMemcpyAsync(memory, out, sizeof(double) * size_t(size), hipMemcpyDeviceToDevice, context->stream));
    if (iterations == 0) {
        CIE_HIP_CHECK(hipMemsetAsync(x, 0, sizeof(double) * size_t(size), context->stream));
    } else {
        hipLaunchKernelGGL(jacobiStartKernel, dim3(vectorGrid(size)), dim3(kReduceBlock), 0, context->stream, size,
                           nullptr, uint16_t(0), in, inverseDiagonal, relaxation, x);
        CIE_HIP_CHECK(hipGetLastError());
        for (uint64_t k = 1; k < iterations; ++k) {
            if (const int s = launchSweep<JacobiSweep>(lhs, static_cast<const double*>(x),
                                                       static_cast<const uint16_t*>(nullptr), uint16_t(0), in,
                                                       inverseDiagonal, relaxation, next))
                return s;
            std::swap(x, next);
        }
    }
    hipLaunchKernelGGL(jacobiFinishKernel, dim3(vectorGrid(size)), dim3(kReduceBlock), 0, context->stream, size, x,
                       memory, inScale, outScale, out);
    CIE_HIP_CHECK(hipGetLastError());
    CIE_HIP_CHECK(hipStreamSynchronize(context->stream));
    return CIE_HIP_OK;
}

int cie_hip_masked_jacobi_f64(cie_hip_context context, cie_hip_csr matrix, cie_hip_mask mask, uint16_t threshold,
                              uint64_t iterations, double relaxation, double inScale, const double* in,
                              double outScale, double* out, int64_t size) {
    if (!mask) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_masked_jacobi_f64: null mask");
    return substitution(context, matrix, mask, threshold, iterations, relaxation, inScale, in, outScale, out, size,
                        "cie_hip_masked_jacobi_f64");
}

int cie_hip_approximate_substitution_f64(cie_hip_context context, cie_hip_csr matrix, uint64_t iterations,
                                         double inScale, const double* in, double outScale, double* out,
                                         int64_t size) {
    return substitution(context, matrix, nullptr, 0, iterations, 1.0, inScale, in, outScale, out, size,
                        "cie_hip_approximate_substitution_f64");
}

int cie_hip_pmg_create(cie_hip_csr matrix, cie_hip_mask mask, int order, cie_hip_pmg* plan) {
    if (!plan) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_pmg_create: null output");
    *plan = nullptr;
    if (!matrix || !mask || order < 1 || order > 65533)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_pmg_create: invalid arguments");
    CIE_HIP_REQUIRE_CONTEXT(mask->context, matrix->context, "cie_hip_pmg_create");
    if (matrix->rows != matrix->columns || mask->size != matrix->rows)
        return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_pmg_create: the mask needs one value per row of a square matrix");
    if (const int s = ensureInverseDiagonal(matrix, 64)) return s;
    auto* p = new cie_hip_pmg_s;
    p->matrix = matrix;
    p->mask = mask;
    holdMatrix(matrix);
    ++mask->holds.holders;
    p->order = order;
    auto bail = [&](int status) {
        releasePmg(p);
        return status;
    };
    const int64_t n = matrix->rows;
    // Grid g keeps mask < g + 2 (solver.cpp:148-205); a grid that keeps every row is the matrix itself. The
    // masked rows of a grid stay zero through its sweeps and update, so grids g >= 1 are compacted to their
    // kept rows and columns (their sweeps then stream only those); grid 0 is solved on the coarse matrix.
    for (int g = 0; g < order; ++g) {
        const uint16_t threshold = uint16_t(g + 2);
        cie_hip_csr grid = matrix;
        int32_t* keep = nullptr;
        int64_t keepRows = 0;
        double* keepInverse = nullptr;
        if (mask->largest >= threshold && g == 0) {
            grid = nullptr;  // grid 0 is the coarse matrix, set below
        } else if (mask->largest >= threshold) {
            std::vector<int32_t> rows, renumber(size_t(n), -1);
            for (int64_t i = 0; i < n; ++i)
                if (mask->host[size_t(i)] < threshold) {
                    renumber[size_t(i)] = int32_t(rows.size());
                    rows.push_back(int32_t(i));
                }
            keepRows = int64_t(rows.size());
            if (const int s = compactMatrix(matrix, rows, renumber, keepRows, &grid)) return bail(s);
            if (hipMalloc(&keep, sizeof(int32_t) * size_t(std::max<int64_t>(1, keepRows))) != hipSuccess ||
                hipMalloc(&keepInverse, sizeof(double) * size_t(std::max<int64_t>(1, keepRows))) != hipSuccess ||
                (keepRows && hipMemcpy(keep, rows.data(), sizeof(int32_t) * rows.size(), hipMemcpyHostToDevice) !=
                                 hipSuccess)) {
                if (keep) (void)hipFree(keep);
                if (keepInverse) (void)hipFree(keepInverse);
                p->grids.push_back(grid);
                return bail(setError(CIE_HIP_ERROR_OUT_OF_MEMORY, "cie_hip_pmg_create: grid index upload failed"));
            }
        }
        p->grids.push_back(grid);
        p->keep.push_back(keep);
        p->keepRows.push_back(keepRows);
        p->keepInverse.push_back(keepInverse);
    }
    std::vector<int32_t> coarseIndex, renumber(size_t(n), -1);
    for (int64_t i = 0; i < n; ++i)
        if (mask->host[size_t(i)] < 2) {
            renumber[size_t(i)] = int32_t(coarseIndex.size());
            coarseIndex.push_back(int32_t(i));
        }
    p->coarseRows = int64_t(coarseIndex.size());
    if (p->coarseRows) {
        if (const int s = compactMatrix(matrix, coarseIndex, renumber, p->coarseRows, &p->coarse)) return bail(s);
        if (!p->grids[0]) p->grids[0] = p->coarse;  // grid 0 keeps exactly the coarse rows
        if (hipMalloc(&p->coarseIndex, sizeof(int32_t) * coarseIndex.size()) != hipSuccess ||
            hipMemcpy(p->coarseIndex, coarseIndex.data(), sizeof(int32_t) * coarseIndex.size(), hipMemcpyHostToDevice) !=
                hipSuccess)
            return bail(setError(CIE_HIP_ERROR_OUT_OF_MEMORY, "cie_hip_pmg_create: coarse index upload failed"));
    }
    if (hipMalloc(&p->work, sizeof(double) * size_t(4 * n + 2 * std::max<int64_t>(1, p->coarseRows))) != hipSuccess)
        return bail(setError(CIE_HIP_ERROR_OUT_OF_MEMORY, "cie_hip_pmg_create: workspace allocation failed"));
    *plan = p;
    return CIE_HIP_OK;
}

int cie_hip_pmg_destroy(cie_hip_pmg plan) {
    releasePmg(plan);
    return CIE_HIP_OK;
}

int cie_hip_pmg_grid(cie_hip_pmg plan, int grid, cie_hip_csr* matrix) {
    if (!plan || !matrix || grid < -1 || grid >= int(plan->grids.size()))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_pmg_grid: invalid arguments");
    *matrix = grid < 0 ? plan->coarse : plan->grids[size_t(grid)];
    return CIE_HIP_OK;
}

int cie_hip_pmg_solve(cie_hip_pmg plan, const double* b, double* x, int64_t size, double relativeTolerance,
                      uint64_t maxCycles, cie_hip_pmg_info* info) {
    if (info) *info = cie_hip_pmg_info{};
    if (!plan || (size && (!b || !x)))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_pmg_solve: invalid arguments");
    cie_hip_csr m = plan->matrix;
    if (size != m->rows) return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_pmg_solve: vector size differs from the matrix");
    cie_hip_context ctx = m->context;
    CIE_HIP_CHECK(hipSetDevice(ctx->device));
    if (const int s = ensureInverseDiagonal(m, 64)) return s;
    const int64_t n = size, nc = plan->coarseRows;
    double* residual = plan->work;
    double* update = residual + n;
    double* xa = update + n;
    double* xb = xa + n;
    double* coarseB = xb + n;
    double* coarseX = coarseB + std::max<int64_t>(1, nc);
    const double* inverse = m->inverseDiagonal64;
    // The matrix's current inverse diagonal onto each compacted grid's rows (what the sweeps read per row).
    for (size_t g = 0; g < plan->keep.size(); ++g)
        if (plan->keep[g] && plan->keepRows[g])
            hipLaunchKernelGGL(gatherKernel, dim3(vectorGrid(plan->keepRows[g])), dim3(kReduceBlock), 0, ctx->stream,
                               plan->keepRows[g], static_cast<const int32_t*>(plan->keep[g]), inverse,
                               plan->keepInverse[g]);
    constexpr double kRelaxation = 2.0 / 3.0;  // solver.cpp:196
    constexpr int kSweeps = 6;                 // solver.cpp:195
    ScopedEvents<2> events;  // start, stop
    CIE_HIP_CHECK(events.create());
    CIE_HIP_CHECK(hipEventRecord(events.event[0], ctx->stream));

    CIE_HIP_CHECK(hipMemcpyAsync(residual, b, sizeof(double) * size_t(n), hipMemcpyDeviceToDevice, ctx->stream));
    double initial = 0.0;
    if (const int s = cie_hip_inner_product_f64(ctx, residual, residual, n, &initial)) return s;
    initial = std::sqrt(initial);
    double norm = initial;
    uint64_t cycles = 0, coarseIterations = 0;
    cie_hip_statistics coarseSettings{5000, 0.0, 0.5};  // solver.cpp:156-159
    int status = CIE_HIP_OK;

    auto visit = [&](int g) -> int {
        cie_hip_csr grid = plan->grids[size_t(g)];
        const uint16_t threshold = uint16_t(g + 2);
        if (g == 0) {
            if (!nc) return CIE_HIP_OK;  // no row has mask < 2: grid 0 is empty and its update zero
            hipLaunchKernelGGL(gatherKernel, dim3(vectorGrid(nc)), dim3(kReduceBlock), 0, ctx->stream, nc,
                               plan->coarseIndex, static_cast<const double*>(residual), coarseB);
            CIE_HIP_CHECK(hipMemsetAsync(coarseX, 0, sizeof(double) * size_t(nc), ctx->stream));
            cie_hip_solve_info coarseInfo{};
            const int s = cie_hip_pcg_f64(ctx, plan->coarse, 0.0, coarseB, 1.0, coarseX, nc, &coarseSettings, 0,
                                          &coarseInfo);
            if (s != CIE_HIP_OK) return s;
            coarseIterations += coarseInfo.stats.iterationCount;
            // solution += update; residual -= A_grid update on the coarse rows (solver.cpp:217-218).
            return launchSweep<ApplyUpdateScatter>(plan->coarse, static_cast<const double*>(coarseX),
                                                   static_cast<const int32_t*>(plan->coarseIndex), residual, x);
        } else if (const int32_t* keep = plan->keep[size_t(g)]) {
            // NestedProductOperator(Jacobi, mask) of the restricted residual (solver.cpp:177-205) on the
            // compacted grid: its right-hand side is the residual on the kept rows (in `update`).
            const int64_t ng = plan->keepRows[size_t(g)];
            const double* gridInverse = plan->keepInverse[size_t(g)];
            double* gridB = update;
            hipLaunchKernelGGL(gatherKernel, dim3(vectorGrid(ng)), dim3(kReduceBlock), 0, ctx->stream, ng, keep,
                               static_cast<const double*>(residual), gridB);
            hipLaunchKernelGGL(jacobiStartKernel, dim3(vectorGrid(ng)), dim3(kReduceBlock), 0, ctx->stream, ng,
                               static_cast<const uint16_t*>(nullptr), threshold, static_cast<const double*>(gridB),
                               gridInverse, kRelaxation, xa);
            CIE_HIP_CHECK(hipGetLastError());
            double *current = xa, *next = xb;
            for (int k = 1; k < kSweeps; ++k) {
                if (const int s = launchSweep<JacobiSweep>(grid, static_cast<const double*>(current),
                                                           static_cast<const uint16_t*>(nullptr), threshold,
                                                           static_cast<const double*>(gridB), gridInverse,
                                                           kRelaxation, next))
                    return s;
                std::swap(current, next);
            }
            // solution += update; residual -= A_grid update on the kept rows (solver.cpp:217-218).
            return launchSweep<ApplyUpdateScatter>(grid, static_cast<const double*>(current), keep, residual, x);
        }
        // A grid that keeps every row (the matrix itself): the same Jacobi sweeps with nothing masked.
        hipLaunchKernelGGL(jacobiStartKernel, dim3(vectorGrid(n)), dim3(kReduceBlock), 0, ctx->stream, n,
                           static_cast<const uint16_t*>(nullptr), threshold, static_cast<const double*>(residual),
                           inverse, kRelaxation, xa);
        CIE_HIP_CHECK(hipGetLastError());
        double *current = xa, *next = xb;
        for (int k = 1; k < kSweeps; ++k) {
            if (const int s = launchSweep<JacobiSweep>(grid, static_cast<const double*>(current),
                                                       static_cast<const uint16_t*>(nullptr), threshold,
                                                       static_cast<const double*>(residual), inverse, kRelaxation,
                                                       next))
                return s;
            std::swap(current, next);
        }
        // solution += update; residual -= A_grid update (solver.cpp:217-218).
        return launchSweep<ApplyUpdate>(grid, static_cast<const double*>(current), residual, x);
    };

    while (relativeTolerance * initial < norm && cycles < maxCycles) {  // solver.cpp:212
        for (int g = plan->order - 1; g >= 0 && status == CIE_HIP_OK; --g) status = visit(g);
        for (int g = 1; g < plan->order && status == CIE_HIP_OK; ++g) status = visit(g);
        if (status != CIE_HIP_OK) break;
        double squared = 0.0;
        status = cie_hip_inner_product_f64(ctx, residual, residual, n, &squared);
        if (status != CIE_HIP_OK) break;
        norm = std::sqrt(squared);
        ++cycles;
    }
    CIE_HIP_CHECK(hipEventRecord(events.event[1], ctx->stream));
    CIE_HIP_CHECK(hipEventSynchronize(events.event[1]));
    float ms = 0.0f;
    CIE_HIP_CHECK(hipEventElapsedTime(&ms, events.event[0], events.event[1]));
    if (info) {
        info->cycles = cycles;
        info->coarseIterations = coarseIterations;
        info->initialResidual = initial;
        info->residual = norm;
        info->converged = !(relativeTolerance * initial < norm) ? 1 : 0;
        info->seconds = double(ms) * 1e-3;
        info->status = status;
    }
    return status;
}

}  // extern "C"

// Parameterised stiffness assembly on the device:
//     K = sum_e scale_e P_e^T K_e0 P_e
// for a given CSR pattern, element -> row map and element matrix K_e0. This is
// the scatter of fem/packages/graph/impl/Assembler_impl.hpp:156-194
// (Assembler::addContribution) for an element integrand linear in the cell
// modulus (fem/packages/integrands/impl/LinearIsotropicStiffnessIntegrand_impl.hpp:74:
// K_e = modulus * D^T D), restated as a gather so that it needs no atomics
// and is deterministic: each row is owned by a group of lanes of one
// wavefront, which visits the elements of the row node in increasing element
// order and adds scale_e * K_e0[a][b] into the slot of every column node b
// (one lane per column node). Every slot thus
// receives its contributions in element order, starting from zero -- the
// order of a sequential element loop -- so the values are bitwise those of
// the structured generator and of a CPU scatter in element order.
//
// The plan (built once per pattern): the elements of each row node in
// increasing element order with the node's local index, and the element
// -> row map on the device; every (row, column) pair of an element is checked
// to be in the CSR pattern when the plan is made.
#include "dispatch.hpp"
#include "kernels.hpp"

#include <type_traits>
#include <vector>

using namespace ciehip;

namespace {

constexpr int kAssemblyBlock = 256;

// Position of `column` in columns[begin, end) (sorted), or -1.
__device__ __forceinline__ int findSlot(const int32_t* __restrict__ columns, int begin, int end, int column) {
    int lo = begin, hi = end;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (columns[mid] < column) lo = mid + 1;
        else hi = mid;
    }
    return lo < end && columns[lo] == column ? lo : -1;
}

// Flags (row, column) pairs of kept element nodes that are not in the pattern.
__global__ __launch_bounds__(kAssemblyBlock) void checkPatternKernel(
    int rows, const int32_t* __restrict__ rowExtents, const int32_t* __restrict__ columns,
    const int64_t* __restrict__ nodeBegin, const int2* __restrict__ nodeElements, const int32_t* __restrict__ dofs,
    int nodes, int* __restrict__ missing) {
    const int row = int(blockIdx.x) * kAssemblyBlock + int(threadIdx.x);
    if (row >= rows) return;
    const int begin = rowExtents[row], end = rowExtents[row + 1];
    for (int64_t c = nodeBegin[row]; c < nodeBegin[row + 1]; ++c) {
        const int32_t* d = dofs + int64_t(nodeElements[c].x) * nodes;
        for (int b = 0; b < nodes; ++b)
            if (d[b] >= 0 && findSlot(columns, begin, end, d[b]) < 0) atomicOr(missing, 1);
    }
}

// A group of G lanes (G divides the wavefront) per row. The group zeroes the
// row, then visits the row node's elements in increasing element order; for
// one element its lanes take the element's column nodes b = lane, lane + G, ...
// and add scale_e * K_e0[a][b] into the slot of d_e[b]. The nodes of one
// element are distinct rows, so the lanes of one element write distinct slots;
// a slot hit by two elements is hit in element order, the later add reading
// the earlier one's store. Both live in one wavefront, whose vector memory
// operations are performed in program order, and a wavefront-scope fence after
// every element keeps the compiler from moving the loads and stores across it.
template <int G>
__global__ __launch_bounds__(kAssemblyBlock) void assembleRowsKernel(
    int rows, const int32_t* __restrict__ rowExtents, const int32_t* __restrict__ columns,
    const int64_t* __restrict__ nodeBegin, const int2* __restrict__ nodeElements, const int32_t* __restrict__ dofs,
    int nodes, const double* __restrict__ elementMatrices, int64_t elementStride, const double* __restrict__ scale,
    double* entries64, float* __restrict__ entries32) {
    static_assert(G >= 1 && G <= 64 && (64 % G) == 0, "a row group lies inside one wavefront");
    const int lane = int(threadIdx.x) % G;
    const int row = int(blockIdx.x) * (kAssemblyBlock / G) + int(threadIdx.x) / G;
    if (row >= rows) return;
    const int begin = rowExtents[row], end = rowExtents[row + 1];
    for (int k = begin + lane; k < end; k += G) entries64[k] = 0.0;
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    for (int64_t c = nodeBegin[row]; c < nodeBegin[row + 1]; ++c) {
        const int2 entry = nodeElements[c];  // {element, local index of the row node}
        const double factor = scale ? scale[entry.x] : 1.0;
        const int32_t* d = dofs + int64_t(entry.x) * nodes;
        const double* matrixRow = elementMatrices + int64_t(entry.x) * elementStride + int64_t(entry.y) * nodes;
        for (int b = lane; b < nodes; b += G) {
            const int column = d[b];
            if (column < 0) continue;
            const int slot = findSlot(columns, begin, end, column);  // present: checked with the plan
            entries64[slot] += factor * matrixRow[b];
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    }
    if (entries32)
        for (int k = begin + lane; k < end; k += G) entries32[k] = float(entries64[k]);
}

template <int G>
void launchAssembly(cie_hip_csr m, const int64_t* nodeBegin, const int2* nodeElements, const int32_t* dofs, int nodes,
                    const double* elementMatrices, int64_t elementStride, const double* scale, hipStream_t stream) {
    constexpr int kRowsPerBlock = kAssemblyBlock / G;
    const unsigned grid = unsigned((m->rows + kRowsPerBlock - 1) / kRowsPerBlock);
    hipLaunchKernelGGL((assembleRowsKernel<G>), dim3(grid), dim3(kAssemblyBlock), 0, stream, int(m->rows),
                       m->rowExtents, m->columnIndices, nodeBegin, nodeElements, dofs, nodes, elementMatrices,
                       elementStride, scale, m->entries64, m->entries32);
}

// One thread per row: out[row] = sum over the row node's elements, in
// increasing element order, of scale_e * v_e[a] (a the row node's local index).
__global__ __launch_bounds__(kAssemblyBlock) void assembleVectorKernel(
    int rows, const int64_t* __restrict__ nodeBegin, const int2* __restrict__ nodeElements,
    const double* __restrict__ elementVectors, int64_t elementStride, const double* __restrict__ scale,
    double* __restrict__ out) {
    const int row = int(blockIdx.x) * kAssemblyBlock + int(threadIdx.x);
    if (row >= rows) return;
    double sum = 0.0;
    for (int64_t c = nodeBegin[row]; c < nodeBegin[row + 1]; ++c) {
        const int2 entry = nodeElements[c];
        const double value = elementVectors[int64_t(entry.x) * elementStride + entry.y];
        sum += scale ? scale[entry.x] * value : value;
    }
    out[row] = sum;
}

}  // namespace

struct cie_hip_assembly_s {
    cie_hip_csr matrix = nullptr;
    int64_t elements = 0;
    int32_t nodes = 0;
    int64_t* nodeBegin = nullptr;   // rows + 1
    int2* nodeElements = nullptr;   // {element, local index}, per row in element order
    int32_t* dofs = nullptr;        // elements x nodes
};

namespace {

void releasePlan(cie_hip_assembly plan) {
    if (!plan) return;
    if (plan->nodeBegin) (void)hipFree(plan->nodeBegin);
    if (plan->nodeElements) (void)hipFree(plan->nodeElements);
    if (plan->dofs) (void)hipFree(plan->dofs);
    delete plan;
}

}  // namespace

extern "C" {

int cie_hip_assembly_create(cie_hip_csr matrix, int64_t elements, int32_t nodesPerElement, const int32_t* dofs,
                            cie_hip_assembly* plan) {
    if (!plan) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_assembly_create: null output");
    *plan = nullptr;
    if (!matrix || elements < 0 || nodesPerElement < 1 || (elements && !dofs))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_assembly_create: invalid arguments");
    if (matrix->rows != matrix->columns)
        return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_assembly_create: the stiffness matrix must be square");
    const int64_t rows = matrix->rows;
    // Elements of each row node in increasing element order (a stable counting sort).
    std::vector<int64_t> begin(size_t(rows) + 1, 0);
    for (int64_t e = 0; e < elements; ++e)
        for (int a = 0; a < nodesPerElement; ++a) {
            const int32_t d = dofs[e * nodesPerElement + a];
            if (d >= rows) return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_assembly_create: element row index out of range");
            if (d >= 0) begin[size_t(d) + 1] += 1;
        }
    if (elements > INT32_MAX) return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_assembly_create: too many elements");
    for (int64_t r = 0; r < rows; ++r) begin[size_t(r) + 1] += begin[size_t(r)];
    std::vector<int2> nodeElements(static_cast<size_t>(begin[static_cast<size_t>(rows)]));
    std::vector<int64_t> fill(begin.begin(), begin.end() - 1);
    for (int64_t e = 0; e < elements; ++e)
        for (int a = 0; a < nodesPerElement; ++a) {
            const int32_t d = dofs[e * nodesPerElement + a];
            if (d >= 0) nodeElements[size_t(fill[size_t(d)]++)] = make_int2(int(e), a);
        }
    cie_hip_context ctx = matrix->context;
    CIE_HIP_CHECK(hipSetDevice(ctx->device));
    auto* p = new cie_hip_assembly_s;
    p->matrix = matrix;
    p->elements = elements;
    p->nodes = nodesPerElement;
    auto bail = [&](int status, const std::string& message) {
        releasePlan(p);
        return setError(status, message);
    };
    const size_t dofBytes = sizeof(int32_t) * size_t(std::max<int64_t>(1, elements * nodesPerElement));
    if (hipMalloc(&p->nodeBegin, sizeof(int64_t) * begin.size()) != hipSuccess ||
        hipMalloc(&p->nodeElements, sizeof(int2) * std::max<size_t>(1, nodeElements.size())) != hipSuccess ||
        hipMalloc(&p->dofs, dofBytes) != hipSuccess)
        return bail(CIE_HIP_ERROR_OUT_OF_MEMORY, "cie_hip_assembly_create: device allocation failed");
    int* missing = reinterpret_cast<int*>(ctx->scratch.scalars + 16);
    int* pinned = reinterpret_cast<int*>(ctx->scratch.pinned + 16);
    bool ok = hipMemcpy(p->nodeBegin, begin.data(), sizeof(int64_t) * begin.size(), hipMemcpyHostToDevice) == hipSuccess &&
              (nodeElements.empty() || hipMemcpy(p->nodeElements, nodeElements.data(), sizeof(int2) * nodeElements.size(),
                                                 hipMemcpyHostToDevice) == hipSuccess) &&
              (!elements || hipMemcpy(p->dofs, dofs, sizeof(int32_t) * size_t(elements * nodesPerElement),
                                      hipMemcpyHostToDevice) == hipSuccess) &&
              hipMemsetAsync(missing, 0, sizeof(int), ctx->stream) == hipSuccess;
    if (ok && rows) {
        hipLaunchKernelGGL(checkPatternKernel, dim3(unsigned((rows + kAssemblyBlock - 1) / kAssemblyBlock)),
                           dim3(kAssemblyBlock), 0, ctx->stream, int(rows), matrix->rowExtents, matrix->columnIndices,
                           p->nodeBegin, p->nodeElements, p->dofs, nodesPerElement, missing);
        ok = hipGetLastError() == hipSuccess;
    }
    ok = ok && hipMemcpyAsync(pinned, missing, sizeof(int), hipMemcpyDeviceToHost, ctx->stream) == hipSuccess &&
         hipStreamSynchronize(ctx->stream) == hipSuccess;
    if (!ok) return bail(CIE_HIP_ERROR_HIP, "cie_hip_assembly_create: upload or pattern check failed");
    if (*pinned) return bail(CIE_HIP_ERROR_SHAPE, "cie_hip_assembly_create: an element couples rows the CSR pattern lacks");
    *plan = p;
    return CIE_HIP_OK;
}

int cie_hip_assembly_destroy(cie_hip_assembly plan) {
    releasePlan(plan);
    return CIE_HIP_OK;
}

int cie_hip_assemble(cie_hip_assembly plan, const double* elementMatrices, int perElement, const double* scale) {
    if (!plan || !elementMatrices) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_assemble: null argument");
    cie_hip_csr m = plan->matrix;
    if (!m->ownsEntries64)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_assemble: the matrix must own its FP64 entries");
    cie_hip_context ctx = m->context;
    CIE_HIP_CHECK(hipSetDevice(ctx->device));
    if (m->rows) {
        // The smallest group that covers one element's nodes, at most a wavefront.
        const int n = plan->nodes;
        const int64_t stride = perElement ? int64_t(n) * n : 0;
        auto launch = [&](auto group) {
            launchAssembly<decltype(group)::value>(m, plan->nodeBegin, plan->nodeElements, plan->dofs, n,
                                                   elementMatrices, stride, scale, ctx->stream);
        };
        if (n <= 4) launch(std::integral_constant<int, 4>{});
        else if (n <= 8) launch(std::integral_constant<int, 8>{});
        else if (n <= 16) launch(std::integral_constant<int, 16>{});
        else if (n <= 32) launch(std::integral_constant<int, 32>{});
        else launch(std::integral_constant<int, 64>{});
        CIE_HIP_CHECK(hipGetLastError());
    }
    CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    // The Jacobi diagonal belongs to the old values.
    if (m->inverseDiagonal64) { CIE_HIP_CHECK(hipFree(m->inverseDiagonal64)); m->inverseDiagonal64 = nullptr; }
    if (m->inverseDiagonal32) { CIE_HIP_CHECK(hipFree(m->inverseDiagonal32)); m->inverseDiagonal32 = nullptr; }
    return CIE_HIP_OK;
}

int cie_hip_assemble_vector(cie_hip_assembly plan, const double* elementVectors, int perElement, const double* scale,
                            double* out) {
    if (!plan || !elementVectors || !out)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_assemble_vector: null argument");
    cie_hip_csr m = plan->matrix;
    cie_hip_context ctx = m->context;
    CIE_HIP_CHECK(hipSetDevice(ctx->device));
    if (m->rows) {
        hipLaunchKernelGGL(assembleVectorKernel, dim3(unsigned((m->rows + kAssemblyBlock - 1) / kAssemblyBlock)),
                           dim3(kAssemblyBlock), 0, ctx->stream, int(m->rows), plan->nodeBegin, plan->nodeElements,
                           elementVectors, perElement ? int64_t(plan->nodes) : int64_t(0), scale, out);
        CIE_HIP_CHECK(hipGetLastError());
    }
    CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return CIE_HIP_OK;
}

}  // extern "C"

// Parameterised stiffness assembly on the device:
//     K = sum_e scale_e P_e^T K_e0 P_e
// for a given CSR pattern, element -> row map and element matrix K_e0. This is
// the scatter of fem/packages/graph/impl/Assembler_impl.hpp:156-194
// (Assembler::addContribution) for an element integrand linear in the cell
// modulus (fem/packages/integrands/impl/LinearIsotropicStiffnessIntegrand_impl.hpp:74:
// K_e = modulus * D^T D), restated as a gather so that it needs no atomics
// and is deterministic: every CSR slot (r, c) is owned by one thread, which
// sums the contributions of the elements holding both nodes in increasing
// element order, starting from zero -- the order of a sequential element
// loop -- so the values are bitwise those of the structured generator and of
// a CPU scatter in element order.
//
// The plan (built once per pattern): the elements of each row node in
// increasing element order with the node's local index; every (row, column)
// pair of an element is checked to be in the CSR pattern; and, for every CSR
// slot, its contributions {element, local pair} in element order, found once
// by walking the two nodes' element lists together (int32 slot offsets, int32
// elements, uint16 local pairs up to 256 nodes per element). Each assembly is
// then one streaming pass over tiles of consecutive slots (6 B per
// contribution and 4 B per slot read, 12 B per slot written).
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

// The contributions to CSR slot (r, c): the elements holding both nodes, in
// increasing element order, and for each the local indices a of r and b of c
// (a-major) -- the order in which a sequential scatter in element order
// (local row, then local column) reaches the slot. Found by walking the
// element lists of r and c together (both sorted by element, then by local
// index). With `elements` null the walk only counts; else it writes the
// element and the local pair a * nodes + b of each contribution.
template <class TLocal>
__device__ __forceinline__ int walkSlot(int row, int column, const int64_t* __restrict__ nodeBegin,
                                        const int2* __restrict__ nodeElements, int nodes,
                                        int32_t* __restrict__ elements, TLocal* __restrict__ locals) {
    int64_t i = nodeBegin[row], j = nodeBegin[column];
    const int64_t rowLast = nodeBegin[row + 1], columnLast = nodeBegin[column + 1];
    int count = 0;
    while (i < rowLast && j < columnLast) {
        const int2 r = nodeElements[i], c = nodeElements[j];
        if (r.x < c.x) {
            ++i;
        } else if (c.x < r.x) {
            ++j;
        } else {
            int64_t j2 = j;
            while (j2 < columnLast && nodeElements[j2].x == r.x) ++j2;
            for (; i < rowLast && nodeElements[i].x == r.x; ++i)
                for (int64_t jj = j; jj < j2; ++jj) {
                    if (elements) {
                        elements[count] = r.x;
                        locals[count] = TLocal(nodeElements[i].y * nodes + nodeElements[jj].y);
                    }
                    ++count;
                }
            j = j2;
        }
    }
    return count;
}

// counts[k] (elements null) or the contributions of every slot, one thread per row.
template <class TLocal>
__global__ __launch_bounds__(kAssemblyBlock) void slotPairsKernel(
    int rows, const int32_t* __restrict__ rowExtents, const int32_t* __restrict__ columns,
    const int64_t* __restrict__ nodeBegin, const int2* __restrict__ nodeElements, int nodes,
    int32_t* __restrict__ counts, const int32_t* __restrict__ slotBegin, int32_t* __restrict__ elements,
    TLocal* __restrict__ locals) {
    const int row = int(blockIdx.x) * kAssemblyBlock + int(threadIdx.x);
    if (row >= rows) return;
    for (int k = rowExtents[row]; k < rowExtents[row + 1]; ++k) {
        if (elements)
            walkSlot(row, columns[k], nodeBegin, nodeElements, nodes, elements + slotBegin[k], locals + slotBegin[k]);
        else
            counts[k] = walkSlot<TLocal>(row, columns[k], nodeBegin, nodeElements, nodes, nullptr, nullptr);
    }
}

constexpr int kSlotTile = 2048;   // contributions per tile
constexpr int kTileSlots = 1024;  // slots per tile (four per lane)

// One block per tile of consecutive slots: the lanes stream the tile's
// contributions (coalesced), stage each product scale_e * K_e[a][b] in LDS,
// and after one barrier each lane adds, for each of its slots (tid, tid + 256,
// ...), that slot's products from zero in their order; no two lanes share a
// slot. A slot holds about 1.4 contributions on a 3D Q2 mesh, so a tile of
// whole slots reaches its contribution capacity only with a few slots per lane.
template <class TLocal>
__global__ __launch_bounds__(kAssemblyBlock) void assembleSlotsKernel(
    const int2* __restrict__ tiles, const int32_t* __restrict__ slotBegin, const int32_t* __restrict__ elements,
    const TLocal* __restrict__ locals, const double* __restrict__ elementMatrices, int64_t elementStride,
    const double* __restrict__ scale, double* __restrict__ entries64, float* __restrict__ entries32) {
    constexpr int kPerLane = kSlotTile / kAssemblyBlock;  // contributions per lane at most
    constexpr int kSlotsPerLane = kTileSlots / kAssemblyBlock;
    __shared__ double stage[kSlotTile];
    const int2 tile = tiles[blockIdx.x], next = tiles[blockIdx.x + 1];  // {first slot, first contribution}
    const int tid = int(threadIdx.x);
    // The lane's slot ranges, in flight while the contributions stream.
    int begin[kSlotsPerLane], end[kSlotsPerLane];
#pragma unroll
    for (int j = 0; j < kSlotsPerLane; ++j) {
        const int slot = tile.x + tid + kAssemblyBlock * j;
        begin[j] = end[j] = 0;
        if (slot < next.x) {
            begin[j] = slotBegin[slot];
            end[j] = slotBegin[slot + 1];
        }
    }
    // All of the lane's contribution loads first, then their dependent scale and K_e loads (masked lanes load
    // nothing), so a lane's loads overlap instead of waiting on each other.
    int element[kPerLane], local[kPerLane];
#pragma unroll
    for (int u = 0; u < kPerLane; ++u) {
        const int p = tile.y + tid + kAssemblyBlock * u;
        element[u] = 0;
        local[u] = 0;
        if (p < next.y) {
            element[u] = elements[p];
            local[u] = int(locals[p]);
        }
    }
    double product[kPerLane];
#pragma unroll
    for (int u = 0; u < kPerLane; ++u) {
        product[u] = 0.0;
        if (tile.y + tid + kAssemblyBlock * u < next.y) {
            const double factor = scale ? scale[element[u]] : 1.0;
            product[u] = factor * elementMatrices[int64_t(element[u]) * elementStride + int64_t(local[u])];
        }
    }
#pragma unroll
    for (int u = 0; u < kPerLane; ++u)
        if (tile.y + tid + kAssemblyBlock * u < next.y) stage[tid + kAssemblyBlock * u] = product[u];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSlotsPerLane; ++j) {
        const int slot = tile.x + tid + kAssemblyBlock * j;
        if (slot < next.x) {
            double sum = 0.0;
            for (int p = begin[j]; p < end[j]; ++p) sum += stage[p - tile.y];
            entries64[slot] = sum;
            if (entries32) entries32[slot] = float(sum);
        }
    }
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

// ---- element stiffness of affine cells ---------------------------------------

constexpr int kStiffnessBlock = 256;
constexpr int kStiffnessEntries = 4096;  // entries of the block's elements per pass
constexpr int kStiffnessElements = 256;  // at most this many elements per block

// |det J| k_e and G = J^-1 J^-T of one affine cell x = x0 + J xi (J row-major,
// dim 2 or 3), by cofactors: inverse[i][j] = cofactor[j][i] / det.
__device__ __forceinline__ void cellGeometry(int dim, const double* J, double modulus, double* G, double* factor) {
    double inverse[9];
    double det;
    if (dim == 2) {
        det = J[0] * J[3] - J[1] * J[2];
        inverse[0] = J[3] / det;
        inverse[1] = -J[1] / det;
        inverse[2] = -J[2] / det;
        inverse[3] = J[0] / det;
    } else {
        const double c00 = J[4] * J[8] - J[5] * J[7], c01 = J[5] * J[6] - J[3] * J[8], c02 = J[3] * J[7] - J[4] * J[6];
        const double c10 = J[2] * J[7] - J[1] * J[8], c11 = J[0] * J[8] - J[2] * J[6], c12 = J[1] * J[6] - J[0] * J[7];
        const double c20 = J[1] * J[5] - J[2] * J[4], c21 = J[2] * J[3] - J[0] * J[5], c22 = J[0] * J[4] - J[1] * J[3];
        det = J[0] * c00 + J[1] * c01 + J[2] * c02;
        inverse[0] = c00 / det; inverse[1] = c10 / det; inverse[2] = c20 / det;
        inverse[3] = c01 / det; inverse[4] = c11 / det; inverse[5] = c21 / det;
        inverse[6] = c02 / det; inverse[7] = c12 / det; inverse[8] = c22 / det;
    }
    for (int i = 0; i < dim; ++i)
        for (int j = 0; j < dim; ++j) {
            double g = 0.0;
            for (int k = 0; k < dim; ++k) g += inverse[i * dim + k] * inverse[j * dim + k];
            G[i * dim + j] = g;
        }
    *factor = fabs(det) * modulus;
}

// out[e][a][b] = |det J_e| k_e sum_ij G_ij moments[i][j][a][b] for a run of
// consecutive elements per block: the block's elements get their geometry in
// LDS first, then every lane writes entries of those elements in order, so
// the stores are contiguous and the moments (the same for every element) are
// read in the same order by every element.
__global__ __launch_bounds__(kStiffnessBlock) void elementStiffnessKernel(
    int64_t elements, int dim, int entries, int perBlock, const double* __restrict__ moments,
    const double* __restrict__ jacobians, const double* __restrict__ modulus, double* __restrict__ out) {
    __shared__ double geometry[kStiffnessElements][10];  // G (up to 9) and the factor
    for (int64_t first = int64_t(blockIdx.x) * perBlock; first < elements; first += int64_t(gridDim.x) * perBlock) {
        const int count = int(min<int64_t>(perBlock, elements - first));
        __syncthreads();  // the previous run's geometry is no longer read
        if (int(threadIdx.x) < count) {
            const int64_t e = first + threadIdx.x;
            cellGeometry(dim, jacobians + e * dim * dim, modulus ? modulus[e] : 1.0, geometry[threadIdx.x],
                         &geometry[threadIdx.x][9]);
        }
        __syncthreads();
        const int total = count * entries;
        const int planes = dim * dim;
        for (int t = int(threadIdx.x); t < total; t += kStiffnessBlock) {
            const int local = t / entries, entry = t - local * entries;
            const double* G = geometry[local];
            double v = 0.0;
            for (int ij = 0; ij < planes; ++ij) v += G[ij] * moments[int64_t(ij) * entries + entry];
            out[(first + local) * entries + entry] = geometry[local][9] * v;
        }
    }
}

}  // namespace

struct cie_hip_assembly_s {
    cie_hip_csr matrix = nullptr;
    int64_t elements = 0;
    int32_t nodes = 0;
    int64_t* nodeBegin = nullptr;   // rows + 1
    int2* nodeElements = nullptr;   // {element, local index}, per row in element order
    int32_t* dofs = nullptr;        // elements x nodes
    int32_t* slotBegin = nullptr;   // nnz + 1: each slot's contributions, in element order
    int32_t* pairElements = nullptr;  // element of each contribution
    void* pairLocals = nullptr;     // a * nodes + b: uint16 if nodes^2 <= 65536, else int32
    bool wideLocals = false;
    int tileCount = 0;
    int2* tiles = nullptr;          // tileCount + 1: {first slot, first contribution}
};

namespace {

void releasePlan(cie_hip_assembly plan) {
    if (!plan) return;
    if (plan->nodeBegin) (void)hipFree(plan->nodeBegin);
    if (plan->nodeElements) (void)hipFree(plan->nodeElements);
    if (plan->dofs) (void)hipFree(plan->dofs);
    if (plan->slotBegin) (void)hipFree(plan->slotBegin);
    if (plan->pairElements) (void)hipFree(plan->pairElements);
    if (plan->pairLocals) (void)hipFree(plan->pairLocals);
    if (plan->tiles) (void)hipFree(plan->tiles);
    cie_hip_csr matrix = plan->matrix;
    delete plan;
    if (matrix) dropMatrix(matrix);
}

}  // namespace

extern "C" {

int cie_hip_assembly_create(cie_hip_csr matrix, int64_t elements, int32_t nodesPerElement, const int32_t* dofs,
                            cie_hip_assembly* plan) {
    if (!plan) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_assembly_create: null output");
    *plan = nullptr;
    if (!matrix || elements < 0 || nodesPerElement < 1 || nodesPerElement > 46340 || (elements && !dofs))
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
    holdMatrix(matrix);
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
    // The contribution lists of every slot: count, prefix on the host, fill.
    const int64_t nnz = matrix->nnz;
    p->wideLocals = int64_t(nodesPerElement) * nodesPerElement > 65536;
    int32_t* counts = nullptr;
    if (hipMalloc(&counts, sizeof(int32_t) * size_t(std::max<int64_t>(1, nnz))) != hipSuccess ||
        hipMalloc(&p->slotBegin, sizeof(int32_t) * size_t(nnz + 1)) != hipSuccess) {
        if (counts) (void)hipFree(counts);
        return bail(CIE_HIP_ERROR_OUT_OF_MEMORY, "cie_hip_assembly_create: slot list allocation failed");
    }
    const unsigned rowBlocks = unsigned((rows + kAssemblyBlock - 1) / kAssemblyBlock);
    auto walk = [&](auto local, int32_t* countOut, int32_t* elementsOut, void* localsOut) {
        using TLocal = decltype(local);
        hipLaunchKernelGGL((slotPairsKernel<TLocal>), dim3(rowBlocks), dim3(kAssemblyBlock), 0, ctx->stream, int(rows),
                           matrix->rowExtents, matrix->columnIndices, p->nodeBegin, p->nodeElements, nodesPerElement,
                           countOut, static_cast<const int32_t*>(p->slotBegin), elementsOut,
                           static_cast<TLocal*>(localsOut));
    };
    std::vector<int32_t> hostBegin(static_cast<size_t>(nnz) + 1, 0);
    if (rows) {
        if (p->wideLocals) walk(int32_t(0), counts, nullptr, nullptr);
        else walk(uint16_t(0), counts, nullptr, nullptr);
        ok = hipGetLastError() == hipSuccess &&
             (!nnz || hipMemcpyAsync(hostBegin.data() + 1, counts, sizeof(int32_t) * size_t(nnz),
                                     hipMemcpyDeviceToHost, ctx->stream) == hipSuccess) &&
             hipStreamSynchronize(ctx->stream) == hipSuccess;
    }
    (void)hipFree(counts);
    if (!ok) return bail(CIE_HIP_ERROR_HIP, "cie_hip_assembly_create: counting slot contributions failed");
    int64_t pairCount = 0;
    for (int64_t k = 0; k < nnz; ++k) {
        pairCount += hostBegin[size_t(k) + 1];
        if (pairCount > INT32_MAX)
            return bail(CIE_HIP_ERROR_SHAPE, "cie_hip_assembly_create: more than 2^31 element contributions");
        hostBegin[size_t(k) + 1] = int32_t(pairCount);
    }
    const size_t localBytes = p->wideLocals ? sizeof(int32_t) : sizeof(uint16_t);
    ok = hipMalloc(&p->pairElements, sizeof(int32_t) * size_t(std::max<int64_t>(1, pairCount))) == hipSuccess &&
         hipMalloc(&p->pairLocals, localBytes * size_t(std::max<int64_t>(1, pairCount))) == hipSuccess &&
         hipMemcpyAsync(p->slotBegin, hostBegin.data(), sizeof(int32_t) * size_t(nnz + 1), hipMemcpyHostToDevice,
                        ctx->stream) == hipSuccess;
    if (ok && rows) {
        if (p->wideLocals) walk(int32_t(0), nullptr, p->pairElements, p->pairLocals);
        else walk(uint16_t(0), nullptr, p->pairElements, p->pairLocals);
        ok = hipGetLastError() == hipSuccess;
    }
    ok = ok && hipStreamSynchronize(ctx->stream) == hipSuccess;
    if (!ok) return bail(CIE_HIP_ERROR_HIP, "cie_hip_assembly_create: writing slot contributions failed");
    // Tiles of whole consecutive slots: at most kSlotTile contributions and kTileSlots slots.
    std::vector<int2> tiles;
    for (int64_t slot = 0; slot < nnz;) {
        tiles.push_back(make_int2(int(slot), hostBegin[size_t(slot)]));
        int64_t last = slot;
        while (last < nnz && last - slot < kTileSlots &&
               int64_t(hostBegin[size_t(last) + 1]) - hostBegin[size_t(slot)] <= kSlotTile)
            ++last;
        if (last == slot)
            return bail(CIE_HIP_ERROR_SHAPE, "cie_hip_assembly_create: a slot has more than 2048 contributions");
        slot = last;
    }
    tiles.push_back(make_int2(int(nnz), hostBegin[size_t(nnz)]));
    p->tileCount = int(tiles.size()) - 1;
    if (hipMalloc(&p->tiles, sizeof(int2) * tiles.size()) != hipSuccess ||
        hipMemcpy(p->tiles, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice) != hipSuccess)
        return bail(CIE_HIP_ERROR_OUT_OF_MEMORY, "cie_hip_assembly_create: tile upload failed");
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
    if (m->nnz && plan->tileCount) {
        const int64_t stride = perElement ? int64_t(plan->nodes) * plan->nodes : 0;
        const dim3 grid(unsigned(plan->tileCount)), block(kAssemblyBlock);
        const KernelTimer timer(ctx);
        if (plan->wideLocals)
            hipLaunchKernelGGL((assembleSlotsKernel<int32_t>), grid, block, 0, ctx->stream,
                               static_cast<const int2*>(plan->tiles), static_cast<const int32_t*>(plan->slotBegin),
                               static_cast<const int32_t*>(plan->pairElements),
                               static_cast<const int32_t*>(plan->pairLocals), elementMatrices, stride, scale,
                               m->entries64, m->entries32);
        else
            hipLaunchKernelGGL((assembleSlotsKernel<uint16_t>), grid, block, 0, ctx->stream,
                               static_cast<const int2*>(plan->tiles), static_cast<const int32_t*>(plan->slotBegin),
                               static_cast<const int32_t*>(plan->pairElements),
                               static_cast<const uint16_t*>(plan->pairLocals), elementMatrices, stride, scale,
                               m->entries64, m->entries32);
        timer.stop();
        CIE_HIP_CHECK(hipGetLastError());
        timer.finish();
        CIE_HIP_CHECK(hipGetLastError());
    }
    CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return entriesChanged(m);  // the Jacobi diagonal and the column-block values belong to the old values
}

int cie_hip_element_stiffness(cie_hip_context context, int64_t elements, int32_t dim, int32_t nodesPerElement,
                              const double* moments, const double* jacobians, const double* modulus, double* out) {
    if (!context || elements < 0 || (dim != 2 && dim != 3) || nodesPerElement < 1 ||
        (elements && (!moments || !jacobians || !out)))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_element_stiffness: invalid arguments");
    const int64_t entries = int64_t(nodesPerElement) * nodesPerElement;
    if (entries > INT32_MAX / kStiffnessBlock)
        return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_element_stiffness: element matrices too large");
    if (!elements) return CIE_HIP_OK;
    CIE_HIP_CHECK(hipSetDevice(context->device));
    const int perBlock = int(std::max<int64_t>(1, std::min<int64_t>(kStiffnessElements, kStiffnessEntries / entries)));
    const int64_t blocks = (elements + perBlock - 1) / perBlock;
    const unsigned grid = unsigned(std::min<int64_t>(blocks, int64_t(context->computeUnits) * 64));
    const KernelTimer timer(context);
    hipLaunchKernelGGL(elementStiffnessKernel, dim3(std::max(1u, grid)), dim3(kStiffnessBlock), 0, context->stream,
                       elements, int(dim), int(entries), perBlock, moments, jacobians, modulus, out);
    timer.stop();
    CIE_HIP_CHECK(hipGetLastError());
    timer.finish();
    CIE_HIP_CHECK(hipStreamSynchronize(context->stream));
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

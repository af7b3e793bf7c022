// Cuthill-McKee orderings of a CSR pattern, and the permutation of a device
// CSR matrix and of device vectors by such an ordering
// (linalg/packages/utilities/src/reorder.cpp).
//
// The ordering is host work, a breadth-first traversal (reorder.cpp:23-173):
// it starts at the first row of least degree (row length); each visited row
// queues its unvisited columns in column order, and that batch is sorted by
// degree with std::sort, whose order for equal degrees is the C++ library's
// (this file is built against the same GCC libstdc++ the reference builds
// with). When the queue runs dry the first unvisited row of least degree
// starts the next component; reverse Cuthill-McKee reverses the result.
// map[new] = old.
//
// The matrix permutation (reorder.cpp:215-365) runs on the device. The map
// is uploaded, inverted and checked there, and the new row extents are
// scanned there. Then one wavefront per row of the result gathers the old row
// and renumbers its columns; a row of up to 64 (128) entries is sorted by a
// bitonic network in registers, one (two) 32-bit keys per lane (64-bit keys
// from 2^25 rows on), a longer one places every entry at its rank among the
// row's new columns (counted out of LDS up to kRankEntries entries, out of L2
// beyond). Both value planes follow their entries, so the result is bitwise
// the reordered input.
#include "dispatch.hpp"

#include <algorithm>
#include <numeric>
#include <vector>

using namespace ciehip;

namespace {

constexpr int kPermuteBlock = 256;
constexpr int kPermuteWaves = kPermuteBlock / 64;
constexpr int kRankEntries = 1024;

// Sorts the 128 keys held two per lane (positions lane and lane + 64) of one
// wavefront ascending: a bitonic network, partners across lanes exchanged by
// __shfl_xor, the stride-64 step inside each lane.
template <class Key>
__device__ __forceinline__ void bitonicSort128(Key& low, Key& high, int lane) {
    for (int k = 2; k <= 128; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j == 64) {
                // positions lane and lane + 64: ascending iff ((lane & k) == 0), here k == 128
                const Key a = low, b = high;
                low = a < b ? a : b;
                high = a < b ? b : a;
            } else {
                const Key otherLow = __shfl_xor(low, j);
                const Key otherHigh = __shfl_xor(high, j);
                const bool lower = (lane & j) == 0;  // this lane holds the smaller position of the pair
                const bool ascendingLow = (lane & k) == 0, ascendingHigh = ((lane + 64) & k) == 0;
                const bool keepMinLow = lower == ascendingLow, keepMinHigh = lower == ascendingHigh;
                low = keepMinLow ? (low < otherLow ? low : otherLow) : (low < otherLow ? otherLow : low);
                high = keepMinHigh ? (high < otherHigh ? high : otherHigh) : (high < otherHigh ? otherHigh : high);
            }
        }
    }
}

// Sorts the 64 keys held one per lane of one wavefront ascending (bitonic).
template <class Key>
__device__ __forceinline__ void bitonicSort64(Key& key, int lane) {
    for (int k = 2; k <= 64; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            const Key other = __shfl_xor(key, j);
            const bool keepMin = ((lane & j) == 0) == ((lane & k) == 0);
            key = keepMin ? (key < other ? key : other) : (key < other ? other : key);
        }
    }
}

// A row of up to 128 entries: sorts (new column, position) keys in registers
// and writes the row in that order. Key is uint32_t ((column << 7) | position,
// for fewer than 2^25 rows) or uint64_t ((column << 32) | position).
template <class Key>
__device__ __forceinline__ void permuteShortRow(int lane, int begin, int length, int out,
                                                const int32_t* __restrict__ target,
                                                const int32_t* __restrict__ oldColumns,
                                                const double* __restrict__ oldEntries64,
                                                const float* __restrict__ oldEntries32,
                                                int32_t* __restrict__ newColumns, double* __restrict__ newEntries64,
                                                float* __restrict__ newEntries32) {
    constexpr int kShift = sizeof(Key) == 4 ? 7 : 32;
    constexpr Key kPad = ~Key(0);
    const auto key = [&](int k) { return (Key(uint32_t(target[oldColumns[begin + k]])) << kShift) | Key(k); };
    const auto place = [&](int k, Key sorted) {
        const int from = begin + int(sorted & ((Key(1) << kShift) - 1));
        newColumns[out + k] = int32_t(sorted >> kShift);
        newEntries64[out + k] = oldEntries64[from];
        if (newEntries32) newEntries32[out + k] = oldEntries32[from];
    };
    if (length <= 64) {  // uniform over the wavefront
        Key one = lane < length ? key(lane) : kPad;
        bitonicSort64(one, lane);
        if (lane < length) place(lane, one);
        return;
    }
    const int k0 = lane, k1 = lane + 64;
    Key low = k0 < length ? key(k0) : kPad;
    Key high = k1 < length ? key(k1) : kPad;
    bitonicSort128(low, high, lane);
    if (k0 < length) place(k0, low);
    if (k1 < length) place(k1, high);
}

__global__ __launch_bounds__(kPermuteBlock) void permuteRowsKernel(
    int rows, const int32_t* __restrict__ source, const int32_t* __restrict__ target,
    const int32_t* __restrict__ oldExtents, const int32_t* __restrict__ oldColumns,
    const double* __restrict__ oldEntries64, const float* __restrict__ oldEntries32,
    const int32_t* __restrict__ newExtents, int32_t* __restrict__ newColumns, double* __restrict__ newEntries64,
    float* __restrict__ newEntries32) {
    __shared__ int32_t staged[kPermuteWaves][kRankEntries];
    const int wave = int(threadIdx.x) / 64, lane = int(threadIdx.x) % 64;
    const int row = int(blockIdx.x) * kPermuteWaves + wave;
    if (row >= rows) return;  // whole wavefronts leave; no block barrier follows
    const int old = source[row];
    const int begin = oldExtents[old];
    const int length = oldExtents[old + 1] - begin;
    const int out = newExtents[row];
    if (length <= 128) {  // uniform over the wavefront
        // keys (new column, position in the old row): distinct, and equal columns keep their order
        if (rows < (1 << 25))
            permuteShortRow<uint32_t>(lane, begin, length, out, target, oldColumns, oldEntries64, oldEntries32,
                                      newColumns, newEntries64, newEntries32);
        else
            permuteShortRow<uint64_t>(lane, begin, length, out, target, oldColumns, oldEntries64, oldEntries32,
                                      newColumns, newEntries64, newEntries32);
        return;
    }
    int32_t* lds = staged[wave];
    const bool inLds = length <= kRankEntries;  // uniform over the wavefront
    if (inLds) {
        for (int k = lane; k < length; k += 64) lds[k] = target[oldColumns[begin + k]];
        // The row is shared inside this wavefront only: order its LDS writes before the reads.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    for (int k = lane; k < length; k += 64) {
        const int column = inLds ? lds[k] : target[oldColumns[begin + k]];
        int rank = 0;
        for (int j = 0; j < length; ++j) {
            const int other = inLds ? lds[j] : target[oldColumns[begin + j]];
            rank += (other < column) || (other == column && j < k);
        }
        newColumns[out + rank] = column;
        newEntries64[out + rank] = oldEntries64[begin + k];
        if (newEntries32) newEntries32[out + rank] = oldEntries32[begin + k];
    }
}

// Preparation of a matrix permutation on the device, from the uploaded map:
// the inverse map, a check that the map is a permutation, and the new row
// extents (each new row's length, then an inclusive scan).

// inverse[map[i]] = i; an index outside [0, rows) is flagged. inverse starts
// at -1, so after this kernel a slot still at -1 means the map repeats an
// index (rows indices in range that miss a slot must hit another twice).
__global__ __launch_bounds__(kReduceBlock) void invertMapKernel(int rows, const int32_t* __restrict__ map,
                                                                int32_t* __restrict__ inverse,
                                                                int* __restrict__ invalid) {
    for (int i = int(blockIdx.x) * kReduceBlock + int(threadIdx.x); i < rows; i += int(gridDim.x) * kReduceBlock) {
        const int32_t j = map[i];
        if (j < 0 || j >= rows) {
            atomicOr(invalid, 1);
            continue;
        }
        inverse[j] = i;
    }
}

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;  // consecutive rows per thread
constexpr int kScanRows = kScanThreads * kScanItems;

// Inclusive scan of the kScanThreads values of a block (Hillis-Steele in LDS).
__device__ __forceinline__ int blockInclusiveScan(int value, int* lds) {
    const int tid = threadIdx.x;
    lds[tid] = value;
    __syncthreads();
    for (int offset = 1; offset < kScanThreads; offset *= 2) {
        const int add = tid >= offset ? lds[tid - offset] : 0;
        __syncthreads();
        lds[tid] += add;
        __syncthreads();
    }
    const int result = lds[tid];
    __syncthreads();  // lds is reused by the caller's next scan
    return result;
}

// New row i is old row source[i]: its length goes to extents[i + 1], scanned
// within the block's kScanRows rows; the block's total goes to totals[block].
// Flags a slot of `inverse` left at -1 and a source index outside [0, rows)
// (then the length counts as 0).
__global__ __launch_bounds__(kScanThreads) void newExtentsKernel(int rows, const int32_t* __restrict__ source,
                                                                 const int32_t* __restrict__ inverse,
                                                                 const int32_t* __restrict__ oldExtents,
                                                                 int32_t* __restrict__ extents,
                                                                 int32_t* __restrict__ totals,
                                                                 int* __restrict__ invalid) {
    __shared__ int lds[kScanThreads];
    const int first = int(blockIdx.x) * kScanRows + int(threadIdx.x) * kScanItems;
    int lengths[kScanItems];
    int sum = 0;
    bool bad = false;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int i = first + k;
        int length = 0;
        if (i < rows) {
            bad |= inverse[i] < 0;
            const int32_t old = source[i];
            if (old >= 0 && old < rows) length = oldExtents[old + 1] - oldExtents[old];
            else bad = true;
        }
        sum += length;
        lengths[k] = sum;
    }
    if (bad) atomicOr(invalid, 1);
    const int before = blockInclusiveScan(sum, lds) - sum;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (first + k < rows) extents[first + k + 1] = before + lengths[k];
    if (threadIdx.x == kScanThreads - 1) totals[blockIdx.x] = before + sum;
    if (blockIdx.x == 0 && threadIdx.x == 0) extents[0] = 0;
}

// One block: totals[b] becomes the exclusive prefix of the block totals.
__global__ __launch_bounds__(kScanThreads) void scanTotalsKernel(int blocks, int32_t* __restrict__ totals) {
    __shared__ int lds[kScanThreads];
    const int per = (blocks + kScanThreads - 1) / kScanThreads;
    const int first = int(threadIdx.x) * per;
    int sum = 0;
    for (int b = first; b < min(blocks, first + per); ++b) sum += totals[b];
    int running = blockInclusiveScan(sum, lds) - sum;
    for (int b = first; b < min(blocks, first + per); ++b) {
        const int total = totals[b];
        totals[b] = running;
        running += total;
    }
}

// Adds the prefix of the blocks before to each block's scanned rows.
__global__ __launch_bounds__(kScanThreads) void addTotalsKernel(int rows, const int32_t* __restrict__ totals,
                                                                int32_t* __restrict__ extents) {
    const int offset = totals[blockIdx.x];
    if (offset == 0) return;
    const int first = int(blockIdx.x) * kScanRows;
    for (int i = first + int(threadIdx.x); i < min(rows, first + kScanRows); i += kScanThreads)
        extents[i + 1] += offset;
}

// out[i] = in[map[i]] (gather) or out[map[i]] = in[i] (scatter); an index
// outside [0, n) is skipped and flagged.
template <typename T>
__global__ __launch_bounds__(kReduceBlock) void permuteVectorKernel(int64_t n, const int32_t* __restrict__ map,
                                                                    int scatter, const T* __restrict__ in,
                                                                    T* __restrict__ out, int* __restrict__ invalid) {
    for (int64_t i = int64_t(blockIdx.x) * kReduceBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kReduceBlock) {
        const int32_t j = map[i];
        if (j < 0 || int64_t(j) >= n) {
            atomicOr(invalid, 1);
            continue;
        }
        if (scatter) out[j] = in[i];
        else out[i] = in[j];
    }
}

template <typename T>
int permuteVector(cie_hip_context ctx, const int32_t* map, int reverse, const T* in, T* out, int64_t n,
                  const char* name) {
    if (!ctx || n < 0 || (n && (!map || !in || !out)))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, std::string(name) + ": invalid arguments");
    if (n && in == out) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, std::string(name) + ": in and out must differ");
    if (!n) return CIE_HIP_OK;
    CIE_HIP_CHECK(hipSetDevice(ctx->device));
    int* invalid = reinterpret_cast<int*>(ctx->scratch.scalars + 24);
    int* pinned = reinterpret_cast<int*>(ctx->scratch.pinned + 24);
    CIE_HIP_CHECK(hipMemsetAsync(invalid, 0, sizeof(int), ctx->stream));
    hipLaunchKernelGGL((permuteVectorKernel<T>), dim3(vectorGrid(n)), dim3(kReduceBlock), 0, ctx->stream, n, map,
                       reverse ? 1 : 0, in, out, invalid);
    CIE_HIP_CHECK(hipGetLastError());
    CIE_HIP_CHECK(hipMemcpyAsync(pinned, invalid, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (*pinned) return setError(CIE_HIP_ERROR_SHAPE, std::string(name) + ": map index out of range");
    return CIE_HIP_OK;
}

}  // namespace

extern "C" {

int cie_hip_make_reordering(int64_t rows, const int32_t* rowExtents, const int32_t* columnIndices, int strategy,
                            int32_t* map) {
    if (rows < 0 || (rows && (!rowExtents || !map)) || strategy < CIE_HIP_REORDER_NONE ||
        strategy > CIE_HIP_REORDER_REVERSE_CUTHILL_MCKEE)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_make_reordering: invalid arguments");
    if (rows > INT32_MAX - 1) return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_make_reordering: too many rows");
    for (int64_t r = 0; r < rows; ++r) map[r] = int32_t(r);
    if (rows == 0) return CIE_HIP_OK;
    // reorder.cpp:33-55: sorted extents, sentinel = entry count.
    if (rowExtents[0] != 0) return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_make_reordering: row extents must start at 0");
    for (int64_t r = 0; r < rows; ++r)
        if (rowExtents[r + 1] < rowExtents[r])
            return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_make_reordering: row extents must be sorted");
    const int64_t nnz = rowExtents[rows];
    if (strategy == CIE_HIP_REORDER_NONE || nnz == 0) return CIE_HIP_OK;
    if (!columnIndices) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_make_reordering: null column indices");
    for (int64_t k = 0; k < nnz; ++k)
        if (columnIndices[k] < 0 || columnIndices[k] >= rows)
            return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_make_reordering: column index out of range (square pattern expected)");

    std::vector<int32_t> degree(static_cast<size_t>(rows));
    for (int64_t r = 0; r < rows; ++r) degree[size_t(r)] = rowExtents[r + 1] - rowExtents[r];
    // Rows by (degree, index): the first unvisited one is where a component starts.
    std::vector<int32_t> starts(static_cast<size_t>(rows));
    std::iota(starts.begin(), starts.end(), 0);
    std::stable_sort(starts.begin(), starts.end(), [&](int32_t a, int32_t b) { return degree[a] < degree[b]; });
    size_t nextStart = 0;
    std::vector<char> seen(static_cast<size_t>(rows), 0);
    std::vector<int32_t> queue;
    queue.reserve(static_cast<size_t>(rows));
    size_t head = 0;
    int64_t placed = 0;
    auto byDegree = [&](int32_t left, int32_t right) { return degree[left] < degree[right]; };
    while (placed < rows) {
        if (head == queue.size()) {
            while (seen[size_t(starts[nextStart])]) ++nextStart;
            const int32_t start = starts[nextStart];
            seen[size_t(start)] = 1;
            queue.push_back(start);
        }
        const int32_t row = queue[head++];
        const size_t batch = queue.size();
        for (int32_t k = rowExtents[row]; k < rowExtents[row + 1]; ++k) {
            const int32_t column = columnIndices[k];
            if (!seen[size_t(column)]) {
                seen[size_t(column)] = 1;
                queue.push_back(column);
            }
        }
        std::sort(queue.begin() + std::ptrdiff_t(batch), queue.end(), byDegree);
        map[placed++] = row;
    }
    if (strategy == CIE_HIP_REORDER_REVERSE_CUTHILL_MCKEE) std::reverse(map, map + rows);
    return CIE_HIP_OK;
}

int cie_hip_csr_reorder(cie_hip_csr matrix, const int32_t* map, int reverse, cie_hip_csr* reordered) {
    if (!reordered) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_reorder: null output");
    *reordered = nullptr;
    if (!matrix || (matrix->rows && !map))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_reorder: invalid arguments");
    if (matrix->rows != matrix->columns)
        return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_csr_reorder: the matrix must be square");
    const int64_t rows = matrix->rows, nnz = matrix->nnz;
    cie_hip_context ctx = matrix->context;
    CIE_HIP_CHECK(hipSetDevice(ctx->device));
    // On the device: the map, its inverse, a check that the map is a
    // permutation, and the new row extents. source[new] = old row,
    // target[old] = new index (reorder.cpp:273-274).
    const int scanBlocks = int((rows + kScanRows - 1) / kScanRows);
    int32_t *dExtents = nullptr, *dColumns = nullptr, *dMap = nullptr, *dInverse = nullptr, *dTotals = nullptr;
    double* dEntries64 = nullptr;
    float* dEntries32 = nullptr;
    auto release = [&]() {
        if (dExtents) (void)hipFree(dExtents);
        if (dColumns) (void)hipFree(dColumns);
        if (dEntries64) (void)hipFree(dEntries64);
        if (dEntries32) (void)hipFree(dEntries32);
    };
    auto releaseMaps = [&]() {
        if (dMap) (void)hipFree(dMap);
        if (dInverse) (void)hipFree(dInverse);
        if (dTotals) (void)hipFree(dTotals);
    };
    const size_t count = size_t(std::max<int64_t>(1, nnz));
    const size_t mapCount = size_t(std::max<int64_t>(1, rows));
    bool ok = hipMalloc(&dExtents, sizeof(int32_t) * (size_t(rows) + 1)) == hipSuccess &&
              hipMalloc(&dColumns, sizeof(int32_t) * count) == hipSuccess &&
              hipMalloc(&dEntries64, sizeof(double) * count) == hipSuccess &&
              (!matrix->entries32 || hipMalloc(&dEntries32, sizeof(float) * count) == hipSuccess) &&
              hipMalloc(&dMap, sizeof(int32_t) * mapCount) == hipSuccess &&
              hipMalloc(&dInverse, sizeof(int32_t) * mapCount) == hipSuccess &&
              hipMalloc(&dTotals, sizeof(int32_t) * size_t(std::max(1, scanBlocks))) == hipSuccess;
    if (!ok) {
        release();
        releaseMaps();
        return setError(CIE_HIP_ERROR_OUT_OF_MEMORY, "cie_hip_csr_reorder: device allocation failed");
    }
    std::vector<int32_t> newExtents(size_t(rows) + 1, 0);
    int* invalid = reinterpret_cast<int*>(ctx->scratch.scalars + 24);
    int* pinned = reinterpret_cast<int*>(ctx->scratch.pinned + 24);
    const int32_t* dSource = reverse ? dInverse : dMap;
    const int32_t* dTarget = reverse ? dMap : dInverse;
    ok = hipMemsetAsync(invalid, 0, sizeof(int), ctx->stream) == hipSuccess;
    if (ok && rows) {
        ok = hipMemcpyAsync(dMap, map, sizeof(int32_t) * size_t(rows), hipMemcpyHostToDevice, ctx->stream) ==
                 hipSuccess &&
             hipMemsetAsync(dInverse, 0xff, sizeof(int32_t) * size_t(rows), ctx->stream) == hipSuccess;  // -1
        if (ok) {
            hipLaunchKernelGGL(invertMapKernel, dim3(unsigned(vectorGrid(rows))), dim3(kReduceBlock), 0, ctx->stream,
                               int(rows), dMap, dInverse, invalid);
            hipLaunchKernelGGL(newExtentsKernel, dim3(unsigned(scanBlocks)), dim3(kScanThreads), 0, ctx->stream,
                               int(rows), dSource, dInverse, matrix->rowExtents, dExtents, dTotals, invalid);
            hipLaunchKernelGGL(scanTotalsKernel, dim3(1), dim3(kScanThreads), 0, ctx->stream, scanBlocks, dTotals);
            hipLaunchKernelGGL(addTotalsKernel, dim3(unsigned(scanBlocks)), dim3(kScanThreads), 0, ctx->stream,
                               int(rows), dTotals, dExtents);
            ok = hipGetLastError() == hipSuccess;
        }
        ok = ok && hipMemcpyAsync(newExtents.data(), dExtents, sizeof(int32_t) * (size_t(rows) + 1),
                                  hipMemcpyDeviceToHost, ctx->stream) == hipSuccess;
    } else if (ok) {
        ok = hipMemcpyAsync(dExtents, newExtents.data(), sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream) ==
             hipSuccess;
    }
    ok = ok && hipMemcpyAsync(pinned, invalid, sizeof(int), hipMemcpyDeviceToHost, ctx->stream) == hipSuccess &&
         hipStreamSynchronize(ctx->stream) == hipSuccess;
    if (!ok) {
        release();
        releaseMaps();
        return setError(CIE_HIP_ERROR_HIP, "cie_hip_csr_reorder: preparing the permutation failed");
    }
    if (*pinned) {
        release();
        releaseMaps();
        return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_csr_reorder: the map is not a permutation");
    }
    if (rows) {
        const KernelTimer timer(ctx);
        hipLaunchKernelGGL(permuteRowsKernel, dim3(unsigned((rows + kPermuteWaves - 1) / kPermuteWaves)),
                           dim3(kPermuteBlock), 0, ctx->stream, int(rows), dSource, dTarget, matrix->rowExtents,
                           matrix->columnIndices, matrix->entries64, matrix->entries32, dExtents, dColumns,
                           dEntries64, dEntries32);
        timer.stop();
        ok = hipGetLastError() == hipSuccess;
        if (ok) timer.finish();
    }
    ok = ok && hipStreamSynchronize(ctx->stream) == hipSuccess;
    releaseMaps();
    if (!ok) {
        release();
        return setError(CIE_HIP_ERROR_HIP, "cie_hip_csr_reorder: permutation failed");
    }
    return adoptMatrix(ctx, rows, rows, dExtents, dColumns, dEntries64, dEntries32, std::move(newExtents), reordered);
}

int cie_hip_reorder_f64(cie_hip_context context, const int32_t* map, int reverse, const double* in, double* out,
                        int64_t size) {
    return permuteVector(context, map, reverse, in, out, size, "cie_hip_reorder_f64");
}

int cie_hip_reorder_f32(cie_hip_context context, const int32_t* map, int reverse, const float* in, float* out,
                        int64_t size) {
    return permuteVector(context, map, reverse, in, out, size, "cie_hip_reorder_f32");
}

}  // extern "C"

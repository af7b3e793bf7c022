// Device CSR matrices, SpMV (LinearOperator::product), the Jacobi inverse
// diagonal and the SpMV launch autotuner.
#include "dispatch.hpp"

#include <climits>
#include "kernels.hpp"

#include <limits>
#include <mutex>
#include <unordered_map>
#include <vector>

using namespace ciehip;

namespace {

// Smallest column of each narrow run (one block per run) and the largest
// column span of any run (an atomic max of integers: order-independent).
__global__ __launch_bounds__(kReduceBlock) void narrowBoundsKernel(int64_t nnz, const int32_t* __restrict__ columns,
                                                                   int32_t* __restrict__ bases,
                                                                   unsigned* __restrict__ maxSpan) {
    __shared__ int ldsLow[kReduceBlock / kWave];
    __shared__ int ldsHigh[kReduceBlock / kWave];
    const int64_t begin = int64_t(blockIdx.x) << kNarrowShift;
    const int64_t end = min(nnz, begin + kNarrowRun);
    int low = INT32_MAX, high = INT32_MIN;
    for (int64_t k = begin + threadIdx.x; k < end; k += kReduceBlock) {
        const int c = columns[k];
        low = min(low, c);
        high = max(high, c);
    }
#pragma unroll
    for (int offset = kWave / 2; offset > 0; offset /= 2) {
        low = min(low, __shfl_xor(low, offset, kWave));
        high = max(high, __shfl_xor(high, offset, kWave));
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
        ldsLow[threadIdx.x / kWave] = low;
        ldsHigh[threadIdx.x / kWave] = high;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kReduceBlock / kWave; ++w) {
            low = min(low, ldsLow[w]);
            high = max(high, ldsHigh[w]);
        }
        bases[blockIdx.x] = low;
        atomicMax(maxSpan, unsigned(high - low));
    }
}

__global__ __launch_bounds__(kReduceBlock) void narrowColumnsKernel(int64_t nnz, const int32_t* __restrict__ columns,
                                                                    const int32_t* __restrict__ bases,
                                                                    uint16_t* __restrict__ narrow) {
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < nnz; k += stride)
        narrow[k] = uint16_t(columns[k] - bases[k >> kNarrowShift]);
}

// ---- column-block form ------------------------------------------------------
// A row's entries are covered by 1 x width blocks: each block starts at the
// column of the row's first entry not yet covered (or at columns - width, when
// that would pass the last column) and covers [start, start + width). A group
// of kCoverLanes lanes walks each row: it loads the row's columns
// kCoverLanes at a time (coalesced), and within each load the next block
// starts at the first lane whose column lies past the current block (a ballot
// and a find-first-set), so the walk visits each entry once. The count and
// structure passes run the same walk, so they agree block for block; the
// structure pass also records each block's first entry, so refilling the
// values after the entries change is one coalesced pass with a thread per block.

constexpr int kCoverLanes = 16;

__device__ __forceinline__ int blockStart(int column, int width, int columns) {
    return column + width > columns ? columns - width : column;
}

// Calls emit(block index within the row, start column, first entry) for every
// block of `row`, in order, from the group's first lane; returns the count.
template <class Emit>
__device__ __forceinline__ int coverRow(int row, const int32_t* __restrict__ extents,
                                        const int32_t* __restrict__ columnIndices, int width, int columns, int lane,
                                        int groupShift, Emit&& emit) {
    const int begin = extents[row], end = extents[row + 1];
    int limit = INT_MIN;  // columns below this are covered (start + width of the last block)
    int count = 0;
    for (int base = begin; base < end; base += kCoverLanes) {
        const int k = base + lane;
        const int column = k < end ? columnIndices[k] : INT_MAX;
        for (;;) {
            const uint64_t open = __ballot(k < end && column >= limit);
            const uint32_t mine = uint32_t(open >> groupShift) & ((1u << kCoverLanes) - 1);
            if (!mine) break;
            const int first = __ffs(int(mine)) - 1;
            const int next = __shfl(column, (groupShift & (kWave - 1)) + first, kWave);
            const int start = blockStart(next, width, columns);
            if (lane == 0) emit(count, start, base + first);
            ++count;
            limit = start + width;
        }
    }
    return count;
}

__global__ __launch_bounds__(kReduceBlock) void columnBlockCountKernel(int rows, const int32_t* __restrict__ extents,
                                                                       const int32_t* __restrict__ columnIndices,
                                                                       int width, int columns,
                                                                       int32_t* __restrict__ counts) {
    const int lane = int(threadIdx.x) % kCoverLanes;
    const int groupShift = int(threadIdx.x) % kWave - lane;  // the group's first lane in the wavefront
    const int groupsPerBlock = kReduceBlock / kCoverLanes;
    const int stride = int(gridDim.x) * groupsPerBlock;
    for (int row = int(blockIdx.x) * groupsPerBlock + int(threadIdx.x) / kCoverLanes; row < rows; row += stride) {
        const int count = coverRow(row, extents, columnIndices, width, columns, lane, groupShift,
                                   [](int, int, int) {});
        if (lane == 0) counts[row] = count;
    }
}

// The first column of every block, and its first CSR entry (blocks hold
// consecutive runs of entries, so block b's entries are
// [entryBegin[b], entryBegin[b + 1]); the last slot, nnz, is set by the host).
__global__ __launch_bounds__(kReduceBlock) void columnBlockStartsKernel(int rows, const int32_t* __restrict__ extents,
                                                                        const int32_t* __restrict__ columnIndices,
                                                                        int width, int columns,
                                                                        const int32_t* __restrict__ blockExtents,
                                                                        int32_t* __restrict__ blockColumns,
                                                                        int32_t* __restrict__ entryBegin) {
    const int lane = int(threadIdx.x) % kCoverLanes;
    const int groupShift = int(threadIdx.x) % kWave - lane;
    const int groupsPerBlock = kReduceBlock / kCoverLanes;
    const int stride = int(gridDim.x) * groupsPerBlock;
    for (int row = int(blockIdx.x) * groupsPerBlock + int(threadIdx.x) / kCoverLanes; row < rows; row += stride) {
        const int b = blockExtents[row];
        coverRow(row, extents, columnIndices, width, columns, lane, groupShift, [&](int i, int start, int entry) {
            entryBegin[b + i] = entry;
            blockColumns[b + i] = start;
        });
    }
}

// The width values of every block, one thread per block: its entries at their
// columns, zero elsewhere (a column listed twice in a row adds up).
template <class T>
__global__ __launch_bounds__(kReduceBlock) void columnBlockValuesKernel(
    int64_t blocks, const int32_t* __restrict__ columnIndices, const T* __restrict__ entries, int width,
    const int32_t* __restrict__ blockColumns, const int32_t* __restrict__ entryBegin, T* __restrict__ values) {
    for (int64_t b = int64_t(blockIdx.x) * kReduceBlock + threadIdx.x; b < blocks;
         b += int64_t(gridDim.x) * kReduceBlock) {
        const int start = blockColumns[b];
        T v[cie_hip_csr_s::kMaxBlockWidth] = {};
        for (int k = entryBegin[b]; k < entryBegin[b + 1]; ++k) v[columnIndices[k] - start] += entries[k];
        for (int i = 0; i < width; ++i) values[b * width + i] = v[i];
    }
}

// Column-sorted tiles of a column-block matrix, one block per tile of at most
// `capacity` (<= kSortMax) blocks: the tile's blocks sorted by (first column,
// position) with a bitonic network in LDS; writes the tile's smallest column
// and each block's key ((column - smallest) << kSortedPositionBits | its
// position in the tile). A tile of one row longer than the capacity is
// skipped (it is walked unsorted). A tile whose columns span more than the
// keys hold sets *failed.
constexpr int kSortThreads = 256;
constexpr int kSortMax = 1 << kSortedPositionBits;

__global__ __launch_bounds__(kSortThreads) void sortedKeysKernel(const int2* __restrict__ tiles, int capacity,
                                                                 const int32_t* __restrict__ blockColumns,
                                                                 int32_t* __restrict__ bases,
                                                                 uint32_t* __restrict__ keys, int* __restrict__ failed) {
    __shared__ uint64_t items[kSortMax];
    const int t = blockIdx.x;
    const int begin = tiles[t].y, count = tiles[t + 1].y - begin;
    if (count > capacity || count == 0) {
        if (threadIdx.x == 0) bases[t] = 0;
        return;
    }
    int n = 1;
    while (n < count) n <<= 1;
    for (int i = threadIdx.x; i < n; i += kSortThreads)
        items[i] = i < count ? (uint64_t(uint32_t(blockColumns[begin + i])) << kSortedPositionBits) | uint64_t(i)
                             : ~uint64_t(0);
    __syncthreads();
    for (int k = 2; k <= n; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n; i += kSortThreads) {
                const int partner = i ^ j;
                if (partner > i) {
                    const uint64_t a = items[i], b = items[partner];
                    if ((a > b) == ((i & k) == 0)) {
                        items[i] = b;
                        items[partner] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
    const int64_t base = int64_t(items[0] >> kSortedPositionBits);
    if (threadIdx.x == 0) {
        bases[t] = int32_t(base);
        if (int64_t(items[count - 1] >> kSortedPositionBits) - base > kSortedMaxSpan) atomicOr(failed, 1);
    }
    for (int i = threadIdx.x; i < count; i += kSortThreads) {
        const uint64_t item = items[i];
        const uint32_t offset = uint32_t(int64_t(item >> kSortedPositionBits) - base);
        keys[begin + i] = (offset << kSortedPositionBits) | uint32_t(item & kSortedPositionMask);
    }
}

// The width values of each sorted block, copied from its block: one block of
// threads per tile, one thread per value in sorted order (coalesced writes
// within each plane, reads within the tile's own span); tiles walked unsorted
// are skipped. `blocks` sorted blocks in all: runs longer than 16 bytes go to
// two planes (sortedHead).
template <class T>
__global__ __launch_bounds__(kReduceBlock) void sortedValuesKernel(const int2* __restrict__ tiles, int capacity,
                                                                   int width, int64_t blocks,
                                                                   const uint32_t* __restrict__ keys,
                                                                   const T* __restrict__ values,
                                                                   T* __restrict__ sorted) {
    const int t = blockIdx.x;
    const int begin = tiles[t].y, end = tiles[t + 1].y;
    if (end - begin > capacity) return;
    const int head = sortedHead(width, int(sizeof(T)));
    T* __restrict__ tail = sorted + blocks * head;
    const int64_t first = int64_t(begin) * width, last = int64_t(end) * width;
    for (int64_t v = first + threadIdx.x; v < last; v += kReduceBlock) {
        const int64_t i = v / width, w = v - i * width;
        const int64_t source = int64_t(begin) + int64_t(keys[i] & kSortedPositionMask);
        const T value = values[source * width + w];
        if (w < head) sorted[i * head + w] = value;
        else tail[i * (width - head) + (w - head)] = value;
    }
}

// Grid of the block-cover kernels: a group of lanes per row, rows strided over a few rounds of blocks.
int coverGrid(cie_hip_context ctx, int rows) {
    constexpr int groupsPerBlock = kReduceBlock / kCoverLanes;
    const int64_t blocks = (int64_t(rows) + groupsPerBlock - 1) / groupsPerBlock;
    return int(std::max<int64_t>(1, std::min<int64_t>(blocks, int64_t(std::max(1, ctx->computeUnits)) * 32)));
}

template <class T>
int launchProduct(cie_hip_csr m, const CsrLaunch& launch, T inScale, const T* in, T outScale, T* out,
                  const T* entries) {
    int status = CIE_HIP_OK;
    const bool ok = withWalker<T, T>(launch, [&](auto walker) {
        using Walker = decltype(walker);
        RowSource source;
        int grid = 0;
        status = prepareLaunch(m, launch, csrProductKernel<Walker>, source, grid);
        if (status != CIE_HIP_OK) return;
        hipLaunchKernelGGL((csrProductKernel<Walker>), dim3(grid), dim3(Walker::kBlock), 0, m->context->stream,
                           source, entries, in, out, inScale, outScale);
    });
    if (!ok) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "unsupported SpMV launch shape");
    if (status != CIE_HIP_OK) return status;
    CIE_HIP_CHECK(hipGetLastError());
    return CIE_HIP_OK;
}

CsrLaunch& launchOf(cie_hip_csr m, int precision) { return precision == 32 ? m->launch32 : m->launch64; }

void releaseMatrix(cie_hip_csr m) {
    if (!m) return;
    if (m->context) {
        (void)hipSetDevice(m->context->device);
        (void)hipStreamSynchronize(m->context->stream);
    }
    if (m->ownsIndices) {
        (void)hipFree(m->rowExtents);
        (void)hipFree(m->columnIndices);
    }
    if (m->ownsEntries64) (void)hipFree(m->entries64);
    if (m->entries32) (void)hipFree(m->entries32);
    if (m->inverseDiagonal64) (void)hipFree(m->inverseDiagonal64);
    if (m->inverseDiagonal32) (void)hipFree(m->inverseDiagonal32);
    if (m->rowCorrection32) (void)hipFree(m->rowCorrection32);
    if (m->workspace) (void)hipFree(m->workspace);
    for (TileSet& set : m->tileSets) {
        if (set.tiles) (void)hipFree(set.tiles);
        releaseSortedCopy(m->context, set.sorted);
    }
    if (m->narrowColumns) (void)hipFree(m->narrowColumns);
    if (m->narrowBases) (void)hipFree(m->narrowBases);
    if (m->blockEntryBegin) (void)hipFree(m->blockEntryBegin);
    for (cie_hip_csr blocks : m->columnBlocks) releaseMatrix(blocks);
    for (PairedForm* form : m->pairedForms) releasePairedForm(m->context, form);
    cie_hip_context context = m->context;
    delete m;
    if (context) dropContext(context);
}

int rowStatistics(cie_hip_csr m, const int32_t* hostExtents) {
    int maxLength = 0;
    for (int64_t row = 0; row < m->rows; ++row)
        maxLength = std::max(maxLength, hostExtents[row + 1] - hostExtents[row]);
    m->maxRowLength = maxLength;
    m->meanRowLength = m->rows ? double(m->nnz) / double(m->rows) : 0.0;
    return CIE_HIP_OK;
}

int buildPlane32(cie_hip_csr m) {
    if (m->entries32) return CIE_HIP_OK;
    CIE_HIP_CHECK(hipMalloc(&m->entries32, sizeof(float) * std::max<int64_t>(1, m->nnz)));
    if (m->nnz) {
        hipLaunchKernelGGL((convertKernel<double, float>), dim3(vectorGrid(m->nnz)), dim3(kReduceBlock), 0,
                           m->context->stream, m->nnz, m->entries64, m->entries32);
        CIE_HIP_CHECK(hipGetLastError());
    }
    return CIE_HIP_OK;
}

// Lanes per row of the Jacobi set-up kernel: FEM rows hold tens to hundreds of entries.
constexpr int kDiagonalLanes = 16;

template <class T>
int inverseDiagonal(cie_hip_csr m, T* inverse, int64_t* failedRow, float* correction = nullptr) {
    cie_hip_context ctx = m->context;
    unsigned long long* failure = reinterpret_cast<unsigned long long*>(ctx->scratch.scalars);
    unsigned long long* pinned = reinterpret_cast<unsigned long long*>(ctx->scratch.pinned);
    pinned[0] = pinned[1] = std::numeric_limits<unsigned long long>::max();
    CIE_HIP_CHECK(hipMemcpyAsync(failure, pinned, 2 * sizeof(unsigned long long), hipMemcpyHostToDevice, ctx->stream));
    if (m->rows) {
        constexpr int groupsPerBlock = kReduceBlock / kDiagonalLanes;
        const int64_t blocks = (m->rows + groupsPerBlock - 1) / groupsPerBlock;
        const int grid = int(std::min<int64_t>(blocks, int64_t(ctx->computeUnits) * 32));
        hipLaunchKernelGGL((diagonalKernel<T, kDiagonalLanes>), dim3(std::max(grid, 1)), dim3(kReduceBlock), 0,
                           ctx->stream, int(m->rows), m->rowExtents, m->columnIndices, m->entries64, inverse,
                           correction, failure);
        CIE_HIP_CHECK(hipGetLastError());
    }
    CIE_HIP_CHECK(hipMemcpyAsync(pinned, failure, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
    CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    const unsigned long long zero = pinned[0], missing = pinned[1];
    if (zero == std::numeric_limits<unsigned long long>::max() && missing == zero) return CIE_HIP_OK;
    // Report the first failing row, as the CPU loop does (DiagonalOperator.cpp:64-83).
    const bool missingFirst = missing < zero;
    const unsigned long long row = missingFirst ? missing : zero;
    if (failedRow) *failedRow = int64_t(row);
    return missingFirst
        ? setError(CIE_HIP_ERROR_MISSING_DIAGONAL, "row " + std::to_string(row) + " has no diagonal entry")
        : setError(CIE_HIP_ERROR_ZERO_DIAGONAL, "row " + std::to_string(row) + " has an explicit zero on its diagonal");
}

}  // namespace

namespace ciehip {

void holdMatrix(cie_hip_csr m) { ++m->holds.holders; }

void dropMatrix(cie_hip_csr m) {
    if (--m->holds.holders == 0 && m->holds.destroyRequested) releaseMatrix(m);
}

int residentBlocks(const void* kernel, int block) {
    static std::mutex mutex;
    static std::unordered_map<const void*, int> cache;
    std::lock_guard<std::mutex> lock(mutex);
    const auto found = cache.find(kernel);
    if (found != cache.end()) return found->second;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, block, 0) != hipSuccess || blocks < 1) blocks = 1;
    cache.emplace(kernel, blocks);
    return blocks;
}

int matrixWorkspace(cie_hip_csr m, size_t bytes, void** pointer) {
    if (m->workspaceBytes < bytes) {
        if (m->workspace) {
            CIE_HIP_CHECK(hipStreamSynchronize(m->context->stream));
            CIE_HIP_CHECK(hipFree(m->workspace));
            m->workspace = nullptr;
            m->workspaceBytes = 0;
        }
        CIE_HIP_CHECK(hipMalloc(&m->workspace, bytes));
        m->workspaceBytes = bytes;
    }
    *pointer = m->workspace;
    return CIE_HIP_OK;
}

// Cuts the rows into tiles of whole consecutive rows with at most `capacity`
// entries and at most kSpmvBlock rows (a longer row is a tile of its own).
int ensureTiles(cie_hip_csr m, int capacity, const TileSet** tiles) {
    for (const TileSet& set : m->tileSets)
        if (set.capacity == capacity) {
            *tiles = &set;
            return CIE_HIP_OK;
        }
    const int32_t* extents = m->hostExtents.data();
    const int rows = int(m->rows);
    std::vector<int2> cut;
    cut.reserve(size_t(m->nnz / std::max(1, capacity / 2) + rows / kSpmvBlock + 2));
    int row = 0;
    while (row < rows) {
        const int first = row;
        if (extents[row + 1] - extents[row] > capacity) {
            ++row;
        } else {
            while (row < rows && row - first < kSpmvBlock && extents[row + 1] - extents[first] <= capacity) ++row;
        }
        cut.push_back(make_int2(first, extents[first]));
    }
    cut.push_back(make_int2(rows, extents[rows]));
    TileSet set;
    set.capacity = capacity;
    set.count = int(cut.size()) - 1;
    CIE_HIP_CHECK(hipMalloc(&set.tiles, sizeof(int2) * cut.size()));
    const hipError_t copied = hipMemcpy(set.tiles, cut.data(), sizeof(int2) * cut.size(), hipMemcpyHostToDevice);
    if (copied != hipSuccess) {
        (void)hipFree(set.tiles);
        return setError(CIE_HIP_ERROR_HIP, std::string("tile upload failed: ") + hipGetErrorString(copied));
    }
    m->tileSets.push_back(set);
    *tiles = &m->tileSets.back();
    return CIE_HIP_OK;
}

void releaseSortedCopy(cie_hip_context context, SortedCopy& copy) {
    if (!copy.bases && !copy.keys && !copy.values64 && !copy.values32) {
        copy.state = copy.state < 0 ? copy.state : 0;
        return;
    }
    if (context) (void)hipStreamSynchronize(context->stream);
    for (void* p : {static_cast<void*>(copy.bases), static_cast<void*>(copy.keys), static_cast<void*>(copy.values64),
                    static_cast<void*>(copy.values32)})
        if (p) (void)hipFree(p);
    copy = SortedCopy{};
}

int ensureSortedCopy(cie_hip_csr blocks, const int2* tiles, int tileCount, int capacity, SortedCopy& copy) {
    if (!blocks->blockWidth) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "sorted tiles of a matrix that is not a column-block form");
    if (capacity > kSortMax)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "sorted tiles hold at most " + std::to_string(kSortMax) + " blocks");
    cie_hip_context ctx = blocks->context;
    const int width = blocks->blockWidth;
    const size_t count = size_t(std::max<int64_t>(1, blocks->nnz));
    const auto spanError = [&]() {
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "sorted tiles: a tile's block columns span more than " +
                                                            std::to_string(kSortedMaxSpan) + " columns");
    };
    if (copy.state < 0) return spanError();
    if (copy.state == 0) {
        int* failed = reinterpret_cast<int*>(ctx->scratch.scalars);
        int* pinned = reinterpret_cast<int*>(ctx->scratch.pinned);
        bool ok = hipMalloc(&copy.bases, sizeof(int32_t) * size_t(std::max(1, tileCount))) == hipSuccess &&
                  hipMalloc(&copy.keys, sizeof(uint32_t) * count) == hipSuccess &&
                  hipMalloc(&copy.values64, sizeof(double) * count * width) == hipSuccess &&
                  hipMemsetAsync(failed, 0, sizeof(int), ctx->stream) == hipSuccess;
        if (ok && tileCount) {
            hipLaunchKernelGGL(sortedKeysKernel, dim3(tileCount), dim3(kSortThreads), 0, ctx->stream, tiles, capacity,
                               blocks->columnIndices, copy.bases, copy.keys, failed);
            ok = hipGetLastError() == hipSuccess;
        }
        ok = ok && hipMemcpyAsync(pinned, failed, sizeof(int), hipMemcpyDeviceToHost, ctx->stream) == hipSuccess &&
             hipStreamSynchronize(ctx->stream) == hipSuccess;
        if (!ok || pinned[0]) {
            releaseSortedCopy(ctx, copy);
            if (!ok) return setError(CIE_HIP_ERROR_OUT_OF_MEMORY, "sorted tiles: building the keys failed");
            copy.state = -1;
            return spanError();
        }
        copy.state = 1;
        copy.version = std::numeric_limits<uint64_t>::max();  // values not written yet
    }
    if (blocks->entries32 && !copy.values32) {
        CIE_HIP_CHECK(hipMalloc(&copy.values32, sizeof(float) * count * width));
        copy.version = std::numeric_limits<uint64_t>::max();
    }
    if (copy.version != blocks->filledVersion) {
        if (tileCount && blocks->nnz) {
            hipLaunchKernelGGL((sortedValuesKernel<double>), dim3(tileCount), dim3(kReduceBlock), 0, ctx->stream, tiles,
                               capacity, width, int64_t(blocks->nnz), copy.keys, blocks->entries64, copy.values64);
            CIE_HIP_CHECK(hipGetLastError());
            if (copy.values32) {
                hipLaunchKernelGGL((sortedValuesKernel<float>), dim3(tileCount), dim3(kReduceBlock), 0, ctx->stream,
                                   tiles, capacity, width, int64_t(blocks->nnz), copy.keys,
                                   static_cast<const float*>(blocks->entries32), copy.values32);
                CIE_HIP_CHECK(hipGetLastError());
            }
        }
        copy.version = blocks->filledVersion;
    }
    return CIE_HIP_OK;
}

int ensureSortedTiles(cie_hip_csr blocks, int capacity, const TileSet** tiles) {
    const TileSet* found = nullptr;
    int status = ensureTiles(blocks, capacity, &found);
    if (status != CIE_HIP_OK) return status;
    TileSet& set = *const_cast<TileSet*>(found);
    status = ensureSortedCopy(blocks, set.tiles, set.count, capacity, set.sorted);
    if (status != CIE_HIP_OK) return status;
    *tiles = &set;
    return CIE_HIP_OK;
}

int ensureNarrowColumns(cie_hip_csr m) {
    if (m->narrowState != 0) return CIE_HIP_OK;
#ifdef CIEHIP_NO_NARROW_COLUMNS
    m->narrowState = -1;  // build option: always read the int32 columns
    return CIE_HIP_OK;
#endif
    if (!m->nnz) {
        m->narrowState = -1;
        return CIE_HIP_OK;
    }
    cie_hip_context ctx = m->context;
    const int64_t runs = (m->nnz + kNarrowRun - 1) / kNarrowRun;
    int32_t* bases = nullptr;
    uint16_t* narrow = nullptr;
    CIE_HIP_CHECK(hipMalloc(&bases, sizeof(int32_t) * runs));
    unsigned* span = reinterpret_cast<unsigned*>(ctx->scratch.scalars + 8);
    unsigned* pinned = reinterpret_cast<unsigned*>(ctx->scratch.pinned + 8);
    bool fits = false;
    hipError_t error = hipMemsetAsync(span, 0, sizeof(unsigned), ctx->stream);
    if (error == hipSuccess) {
        hipLaunchKernelGGL(narrowBoundsKernel, dim3(unsigned(runs)), dim3(kReduceBlock), 0, ctx->stream, m->nnz,
                           m->columnIndices, bases, span);
        error = hipGetLastError();
    }
    if (error == hipSuccess) error = hipMemcpyAsync(pinned, span, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream);
    if (error == hipSuccess) error = hipStreamSynchronize(ctx->stream);
    if (error == hipSuccess) fits = *pinned <= 0xffffu;
    if (error == hipSuccess && fits) {
        error = hipMalloc(&narrow, sizeof(uint16_t) * size_t(m->nnz));
        if (error == hipSuccess) {
            hipLaunchKernelGGL(narrowColumnsKernel, dim3(vectorGrid(m->nnz)), dim3(kReduceBlock), 0, ctx->stream,
                               m->nnz, m->columnIndices, bases, narrow);
            error = hipGetLastError();
        }
        if (error == hipSuccess) error = hipStreamSynchronize(ctx->stream);
    }
    if (error != hipSuccess) {
        (void)hipFree(bases);
        if (narrow) (void)hipFree(narrow);
        return setError(CIE_HIP_ERROR_HIP, std::string("narrow column plane: ") + hipGetErrorString(error));
    }
    if (!fits) {
        (void)hipFree(bases);
        m->narrowState = -1;
        return CIE_HIP_OK;
    }
    m->narrowBases = bases;
    m->narrowColumns = narrow;
    m->narrowRuns = int(runs);
    m->narrowState = 1;
    return CIE_HIP_OK;
}

int ensureInverseDiagonal(cie_hip_csr m, int precision) {
    if (m->rows != m->columns)
        return setError(CIE_HIP_ERROR_SHAPE, "the Jacobi preconditioner needs a square matrix");
    if (precision == 32) {
        // The FP32 inverse diagonal and, from the same pass, the row-sum
        // correction of the FP32 plane (both used by the mixed solvers).
        if (m->inverseDiagonal32 && m->rowCorrection32) return CIE_HIP_OK;
        float *inverse = nullptr, *correction = nullptr;
        const size_t bytes = sizeof(float) * size_t(std::max<int64_t>(1, m->rows));
        CIE_HIP_CHECK(hipMalloc(&inverse, bytes));
        if (hipMalloc(&correction, bytes) != hipSuccess) {
            (void)hipFree(inverse);
            return setError(CIE_HIP_ERROR_OUT_OF_MEMORY, "row-sum correction: out of device memory");
        }
        const int status = inverseDiagonal<float>(m, inverse, nullptr, correction);
        if (status != CIE_HIP_OK) {
            (void)hipFree(inverse);
            (void)hipFree(correction);
            return status;
        }
        if (m->inverseDiagonal32) (void)hipFree(m->inverseDiagonal32);
        if (m->rowCorrection32) (void)hipFree(m->rowCorrection32);
        m->inverseDiagonal32 = inverse;
        m->rowCorrection32 = correction;
        return CIE_HIP_OK;
    }
    if (m->inverseDiagonal64) return CIE_HIP_OK;
    double* inverse = nullptr;
    CIE_HIP_CHECK(hipMalloc(&inverse, sizeof(double) * std::max<int64_t>(1, m->rows)));
    const int status = inverseDiagonal<double>(m, inverse, nullptr);
    if (status != CIE_HIP_OK) {
        (void)hipFree(inverse);
        return status;
    }
    m->inverseDiagonal64 = inverse;
    return CIE_HIP_OK;
}

int adoptMatrix(cie_hip_context context, int64_t rows, int64_t columns, int32_t* rowExtents, int32_t* columnIndices,
                double* entries64, float* entries32, std::vector<int32_t> hostExtents, cie_hip_csr* matrix) {
    auto* m = new cie_hip_csr_s;
    m->context = context;
    holdContext(context);
    m->rows = rows;
    m->columns = columns;
    m->nnz = hostExtents[size_t(rows)];
    m->rowExtents = rowExtents;
    m->columnIndices = columnIndices;
    m->entries64 = entries64;
    m->entries32 = entries32;
    m->ownsIndices = m->ownsEntries64 = true;
    rowStatistics(m, hostExtents.data());
    m->hostExtents = std::move(hostExtents);
    m->launch64 = m->launch32 = defaultLaunch();
    *matrix = m;
    return CIE_HIP_OK;
}

int entriesChanged(cie_hip_csr m) {
    ++m->valuesVersion;
    if (m->inverseDiagonal64) {
        CIE_HIP_CHECK(hipFree(m->inverseDiagonal64));
        m->inverseDiagonal64 = nullptr;
    }
    if (m->inverseDiagonal32) {
        CIE_HIP_CHECK(hipFree(m->inverseDiagonal32));
        m->inverseDiagonal32 = nullptr;
    }
    if (m->rowCorrection32) {
        CIE_HIP_CHECK(hipFree(m->rowCorrection32));
        m->rowCorrection32 = nullptr;
    }
    return CIE_HIP_OK;
}

int ensureColumnBlocks(cie_hip_csr m, int width, cie_hip_csr* out) {
    if (m->blockWidth) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "column blocks of a column-block form");
    if (width < 2 || width > cie_hip_csr_s::kMaxBlockWidth)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "column blocks: width must be 2.." +
                                                            std::to_string(cie_hip_csr_s::kMaxBlockWidth));
    if (m->columns < width)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "column blocks: the matrix has fewer columns than the block width");
    cie_hip_context ctx = m->context;
    const int rows = int(m->rows), columns = int(m->columns);
    cie_hip_csr blocks = m->columnBlocks[width];
    if (!blocks) {
        // Structure: blocks per row, their prefix on the host, then the first column of each block.
        int32_t *counts = nullptr, *extents = nullptr, *starts = nullptr, *entryBegin = nullptr;
        double* values64 = nullptr;
        float* values32 = nullptr;
        auto release = [&]() {
            for (void* p : {static_cast<void*>(counts), static_cast<void*>(extents), static_cast<void*>(starts),
                            static_cast<void*>(entryBegin), static_cast<void*>(values64), static_cast<void*>(values32)})
                if (p) (void)hipFree(p);
        };
        std::vector<int32_t> hostCounts(static_cast<size_t>(rows)), hostExtents(static_cast<size_t>(rows) + 1, 0);
        bool ok = hipMalloc(&counts, sizeof(int32_t) * size_t(std::max(1, rows))) == hipSuccess &&
                  hipMalloc(&extents, sizeof(int32_t) * (size_t(rows) + 1)) == hipSuccess;
        if (ok && rows) {
            hipLaunchKernelGGL(columnBlockCountKernel, dim3(coverGrid(ctx, rows)), dim3(kReduceBlock), 0, ctx->stream, rows,
                               m->rowExtents, m->columnIndices, width, columns, counts);
            ok = hipGetLastError() == hipSuccess &&
                 hipMemcpyAsync(hostCounts.data(), counts, sizeof(int32_t) * size_t(rows), hipMemcpyDeviceToHost,
                                ctx->stream) == hipSuccess &&
                 hipStreamSynchronize(ctx->stream) == hipSuccess;
        }
        if (!ok) {
            release();
            return setError(CIE_HIP_ERROR_HIP, "column blocks: counting the blocks failed");
        }
        int64_t total = 0;
        for (int row = 0; row < rows; ++row) {
            total += hostCounts[size_t(row)];
            hostExtents[size_t(row) + 1] = int32_t(total);  // total <= nnz <= INT32_MAX
        }
        const size_t count = size_t(std::max<int64_t>(1, total));
        const int32_t entryCount = int32_t(m->nnz);
        ok = hipMalloc(&starts, sizeof(int32_t) * count) == hipSuccess &&
             hipMalloc(&entryBegin, sizeof(int32_t) * (count + 1)) == hipSuccess &&
             hipMemcpyAsync(entryBegin + total, &entryCount, sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream) ==
                 hipSuccess &&
             hipMalloc(&values64, sizeof(double) * count * width) == hipSuccess &&
             (!m->entries32 || hipMalloc(&values32, sizeof(float) * count * width) == hipSuccess) &&
             hipMemcpyAsync(extents, hostExtents.data(), sizeof(int32_t) * (size_t(rows) + 1), hipMemcpyHostToDevice,
                            ctx->stream) == hipSuccess;
        if (ok && rows) {
            hipLaunchKernelGGL(columnBlockStartsKernel, dim3(coverGrid(ctx, rows)), dim3(kReduceBlock), 0, ctx->stream,
                               rows, m->rowExtents, m->columnIndices, width, columns, extents, starts, entryBegin);
            ok = hipGetLastError() == hipSuccess;
        }
        ok = ok && hipStreamSynchronize(ctx->stream) == hipSuccess;
        (void)hipFree(counts);
        counts = nullptr;
        if (!ok) {
            release();
            return setError(CIE_HIP_ERROR_OUT_OF_MEMORY, "column blocks: building the block structure failed");
        }
        adoptMatrix(ctx, rows, columns, extents, starts, values64, values32, std::move(hostExtents), &blocks);
        blocks->blockWidth = width;
        blocks->blockEntryBegin = entryBegin;
        blocks->filledVersion = std::numeric_limits<uint64_t>::max();  // values not written yet
        m->columnBlocks[width] = blocks;
    }
    if (m->entries32 && !blocks->entries32) {  // an FP32 plane the block form does not mirror yet
        CIE_HIP_CHECK(hipMalloc(&blocks->entries32, sizeof(float) * size_t(std::max<int64_t>(1, blocks->nnz)) * width));
        blocks->filledVersion = std::numeric_limits<uint64_t>::max();
    }
    if (blocks->filledVersion != m->valuesVersion) {
        if (blocks->nnz) {
            const dim3 grid(unsigned(vectorGrid(blocks->nnz)));
            hipLaunchKernelGGL((columnBlockValuesKernel<double>), grid, dim3(kReduceBlock), 0, ctx->stream,
                               blocks->nnz, m->columnIndices, static_cast<const double*>(m->entries64), width,
                               blocks->columnIndices, blocks->blockEntryBegin, blocks->entries64);
            CIE_HIP_CHECK(hipGetLastError());
            if (blocks->entries32 && m->entries32) {
                hipLaunchKernelGGL((columnBlockValuesKernel<float>), grid, dim3(kReduceBlock), 0, ctx->stream,
                                   blocks->nnz, m->columnIndices, static_cast<const float*>(m->entries32), width,
                                   blocks->columnIndices, blocks->blockEntryBegin, blocks->entries32);
                CIE_HIP_CHECK(hipGetLastError());
            }
        }
        blocks->filledVersion = m->valuesVersion;
    }
    *out = blocks;
    return CIE_HIP_OK;
}

// Frees the column-block forms no launch of the matrix uses, the sorted
// copies of tile capacities no sorted launch uses and the paired forms no
// paired launch uses (all rebuilt if asked for again; paired forms that do not
// fit the matrix are kept, they hold no memory).
void releaseUnusedColumnBlocks(cie_hip_csr m) {
    std::vector<PairedForm*> kept;
    for (PairedForm* form : m->pairedForms) {
        const auto uses = [form](const CsrLaunch& launch) {
            return launch.kernel == CIE_HIP_SPMV_PAIRED && pairedWide(launch.width) == form->wide &&
                   pairedNarrow(launch.width) == form->narrow && launch.unroll == form->unroll;
        };
        if (form->state < 0 || uses(m->launch64) || uses(m->launch32)) kept.push_back(form);
        else releasePairedForm(m->context, form);
    }
    m->pairedForms = std::move(kept);
    for (int width = 2; width <= cie_hip_csr_s::kMaxBlockWidth; ++width) {
        cie_hip_csr blocks = m->columnBlocks[width];
        if (!blocks) continue;
        for (TileSet& set : blocks->tileSets) {
            if (set.sorted.state <= 0) continue;
            const auto sortedUse = [&](const CsrLaunch& launch) {
                return launch.kernel == CIE_HIP_SPMV_SORTED && launch.width == width &&
                       launch.block * launch.unroll == set.capacity;
            };
            if (!sortedUse(m->launch64) && !sortedUse(m->launch32)) releaseSortedCopy(m->context, set.sorted);
        }
    }
    for (int width = 2; width <= cie_hip_csr_s::kMaxBlockWidth; ++width) {
        const auto uses = [width](const CsrLaunch& launch) {
            return (launch.kernel == CIE_HIP_SPMV_BLOCKS || launch.kernel == CIE_HIP_SPMV_SORTED) &&
                   launch.width == width;
        };
        const bool used = uses(m->launch64) || uses(m->launch32);
        if (!used && m->columnBlocks[width] && !m->columnBlocks[width]->blockUsers) {
            releaseMatrix(m->columnBlocks[width]);
            m->columnBlocks[width] = nullptr;
        }
    }
}

}  // namespace ciehip

extern "C" {

int cie_hip_csr_create(cie_hip_context context, int64_t rows, int64_t columns, int64_t nnz,
                       const int32_t* rowExtents, const int32_t* columnIndices, const double* entries,
                       int location, int planes, cie_hip_csr* matrix) {
    if (!matrix) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_create: null output");
    *matrix = nullptr;
    if (!context || rows < 0 || columns < 0 || nnz < 0 || !rowExtents || (nnz && (!columnIndices || !entries)))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_create: invalid arguments");
    if (rows > INT32_MAX - 1 || columns > INT32_MAX || nnz > INT32_MAX)
        return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_csr_create: sizes exceed int32 indexing");
    if (location != CIE_HIP_HOST && location != CIE_HIP_DEVICE)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_create: location must be host or device");
    if (!(planes & (CIE_HIP_PLANE_F64 | CIE_HIP_PLANE_F32)))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_create: no value plane requested");
    CIE_HIP_CHECK(hipSetDevice(context->device));

    std::vector<int32_t> hostExtents;
    const int32_t* extents = rowExtents;
    if (location == CIE_HIP_HOST) {
        // CSROperator.cpp:21-29 checks; here also monotone extents and column range.
        if (rowExtents[0] != 0 || rowExtents[rows] != nnz)
            return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_csr_create: row extents must start at 0 and end at nnz");
        for (int64_t row = 0; row < rows; ++row)
            if (rowExtents[row + 1] < rowExtents[row])
                return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_csr_create: row extents decrease at row " + std::to_string(row));
        for (int64_t k = 0; k < nnz; ++k)
            if (columnIndices[k] < 0 || columnIndices[k] >= columns)
                return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_csr_create: column index out of range at entry " + std::to_string(k));
    } else {
        hostExtents.resize(size_t(rows) + 1);
        CIE_HIP_CHECK(hipMemcpy(hostExtents.data(), rowExtents, sizeof(int32_t) * (rows + 1), hipMemcpyDeviceToHost));
        if (hostExtents[0] != 0 || hostExtents[rows] != nnz)
            return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_csr_create: row extents must start at 0 and end at nnz");
        extents = hostExtents.data();
    }

    auto* m = new cie_hip_csr_s;
    m->context = context;
    holdContext(context);
    m->rows = rows;
    m->columns = columns;
    m->nnz = nnz;
    auto bail = [&](int status, const std::string& message) {
        releaseMatrix(m);
        return setError(status, message);
    };
    const size_t extentBytes = sizeof(int32_t) * size_t(rows + 1);
    const size_t indexBytes = sizeof(int32_t) * size_t(std::max<int64_t>(1, nnz));
    const size_t entryBytes = sizeof(double) * size_t(std::max<int64_t>(1, nnz));
    if (location == CIE_HIP_HOST) {
        m->ownsIndices = m->ownsEntries64 = true;
        if (hipMalloc(&m->rowExtents, extentBytes) != hipSuccess || hipMalloc(&m->columnIndices, indexBytes) != hipSuccess
            || hipMalloc(&m->entries64, entryBytes) != hipSuccess)
            return bail(CIE_HIP_ERROR_OUT_OF_MEMORY, "cie_hip_csr_create: device allocation failed");
        if (hipMemcpyAsync(m->rowExtents, rowExtents, extentBytes, hipMemcpyHostToDevice, context->stream) != hipSuccess
            || (nnz && hipMemcpyAsync(m->columnIndices, columnIndices, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, context->stream) != hipSuccess)
            || (nnz && hipMemcpyAsync(m->entries64, entries, sizeof(double) * nnz, hipMemcpyHostToDevice, context->stream) != hipSuccess))
            return bail(CIE_HIP_ERROR_HIP, "cie_hip_csr_create: upload failed");
    } else {
        m->rowExtents = const_cast<int32_t*>(rowExtents);
        m->columnIndices = const_cast<int32_t*>(columnIndices);
        m->entries64 = const_cast<double*>(entries);
    }
    rowStatistics(m, extents);
    m->hostExtents.assign(extents, extents + rows + 1);
    if (planes & CIE_HIP_PLANE_F32) {
        const int status = buildPlane32(m);
        if (status != CIE_HIP_OK) return bail(status, cie_hip_last_error());
    }
    if (hipStreamSynchronize(context->stream) != hipSuccess)
        return bail(CIE_HIP_ERROR_HIP, "cie_hip_csr_create: stream synchronisation failed");
    m->launch64 = m->launch32 = defaultLaunch();
    *matrix = m;
    return CIE_HIP_OK;
}

int cie_hip_csr_copy_out(cie_hip_csr matrix, int32_t* rowExtents, int32_t* columnIndices, double* entries) {
    if (!matrix) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_copy_out: null matrix");
    cie_hip_context ctx = matrix->context;
    CIE_HIP_CHECK(hipSetDevice(ctx->device));
    CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (rowExtents)
        CIE_HIP_CHECK(hipMemcpy(rowExtents, matrix->rowExtents, sizeof(int32_t) * (matrix->rows + 1), hipMemcpyDeviceToHost));
    if (columnIndices && matrix->nnz)
        CIE_HIP_CHECK(hipMemcpy(columnIndices, matrix->columnIndices, sizeof(int32_t) * matrix->nnz, hipMemcpyDeviceToHost));
    if (entries && matrix->nnz)
        CIE_HIP_CHECK(hipMemcpy(entries, matrix->entries64, sizeof(double) * matrix->nnz, hipMemcpyDeviceToHost));
    return CIE_HIP_OK;
}

int cie_hip_csr_destroy(cie_hip_csr matrix) {
    if (!matrix || matrix->holds.destroyRequested) return CIE_HIP_OK;
    if (matrix->holds.holders > 0) {
        matrix->holds.destroyRequested = true;  // released with the last batch or plan over it
        return CIE_HIP_OK;
    }
    releaseMatrix(matrix);
    return CIE_HIP_OK;
}

int cie_hip_csr_shape(cie_hip_csr matrix, int64_t* rows, int64_t* columns, int64_t* nnz) {
    if (!matrix) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_shape: null matrix");
    if (rows) *rows = matrix->rows;
    if (columns) *columns = matrix->columns;
    if (nnz) *nnz = matrix->nnz;
    return CIE_HIP_OK;
}

int cie_hip_csr_update_entries(cie_hip_csr matrix, const double* entries) {
    if (!matrix || (!entries && matrix->nnz)) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_update_entries: null argument");
    cie_hip_context ctx = matrix->context;
    CIE_HIP_CHECK(hipSetDevice(ctx->device));
    if (!matrix->nnz) return CIE_HIP_OK;
    if (matrix->ownsEntries64) {
        if (entries != matrix->entries64)
            CIE_HIP_CHECK(hipMemcpyAsync(matrix->entries64, entries, sizeof(double) * matrix->nnz, hipMemcpyDeviceToDevice, ctx->stream));
    } else {
        matrix->entries64 = const_cast<double*>(entries);
    }
    if (matrix->entries32) {
        hipLaunchKernelGGL((convertKernel<double, float>), dim3(vectorGrid(matrix->nnz)), dim3(kReduceBlock), 0,
                           ctx->stream, matrix->nnz, matrix->entries64, matrix->entries32);
        CIE_HIP_CHECK(hipGetLastError());
    }
    CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return entriesChanged(matrix);
}

int cie_hip_csr_set_launch(cie_hip_csr matrix, int precision, int kernel, int width, int unroll) {
    if (!matrix || (precision != 32 && precision != 64))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_set_launch: invalid matrix or precision");
    if (kernel == CIE_HIP_SPMV_STAGED) {
        if (!validStage(width, unroll))
            return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_set_launch: staged width must be 512 or 1024, unroll 4 or 8");
    } else if (kernel == CIE_HIP_SPMV_VECTOR) {
        if (!validShape(width, unroll))
            return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_set_launch: vector width must be 1..64 (power of two), unroll 1, 2, 4 or 8");
    } else if (kernel == CIE_HIP_SPMV_TILED) {
        if (!validTiles(width, unroll))
            return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_set_launch: tiled width must be 256, unroll 4, 8 or 16");
    } else if (kernel == CIE_HIP_SPMV_PAIRED) {
        if (!validPaired(width, unroll))
            return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_set_launch: paired width must be 53 or 32, unroll 2 or 4");
        if (matrix->columns < pairedWide(width))
            return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_set_launch: the matrix has fewer columns than the wide block");
    } else if (kernel == CIE_HIP_SPMV_BLOCKS || kernel == CIE_HIP_SPMV_SORTED) {
        if (!validBlocks(width, unroll))
            return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_set_launch: block width must be 2, 3 or 4, unroll 4, 8 or 16");
        if (matrix->columns < width)
            return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_set_launch: the matrix has fewer columns than the block width");
    } else {
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_set_launch: unknown kernel");
    }
    CsrLaunch& launch = launchOf(matrix, precision);
    launch.kernel = kernel;
    launch.width = width;
    launch.unroll = unroll;
    launch.block = kSpmvBlock;
    return CIE_HIP_OK;
}

int cie_hip_csr_column_bytes(cie_hip_csr matrix, int* bytes) {
    if (!matrix || !bytes) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_column_bytes: null argument");
    CIE_HIP_CHECK(hipSetDevice(matrix->context->device));
    const int status = ensureNarrowColumns(matrix);
    if (status != CIE_HIP_OK) return status;
    *bytes = matrix->narrowState > 0 ? int(sizeof(uint16_t)) : int(sizeof(int32_t));
    return CIE_HIP_OK;
}

int cie_hip_csr_column_blocks(cie_hip_csr matrix, int width, int64_t* blocks, int* bytes) {
    if (!matrix) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_column_blocks: null matrix");
    CIE_HIP_CHECK(hipSetDevice(matrix->context->device));
    cie_hip_csr form = nullptr;
    int status = ensureColumnBlocks(matrix, width, &form);
    if (status != CIE_HIP_OK) return status;
    status = ensureNarrowColumns(form);
    if (status != CIE_HIP_OK) return status;
    if (blocks) *blocks = form->nnz;
    if (bytes) *bytes = form->narrowState > 0 ? int(sizeof(uint16_t)) : int(sizeof(int32_t));
    return CIE_HIP_OK;
}

int cie_hip_csr_paired_blocks(cie_hip_csr matrix, int width, int64_t* wideBlocks, int64_t* narrowBlocks) {
    if (!matrix) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_paired_blocks: null matrix");
    if (!validPaired(width, 4)) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_paired_blocks: paired width must be 53 or 32");
    CIE_HIP_CHECK(hipSetDevice(matrix->context->device));
    int unroll = 4;
    for (const CsrLaunch* launch : {&matrix->launch64, &matrix->launch32})
        if (launch->kernel == CIE_HIP_SPMV_PAIRED && launch->width == width) unroll = launch->unroll;
    const PairedForm* form = nullptr;
    const int status = ensurePairedForm(matrix, pairedWide(width), pairedNarrow(width), unroll, &form);
    if (status != CIE_HIP_OK) return status;
    if (wideBlocks) *wideBlocks = form->blocks[0];
    if (narrowBlocks) *narrowBlocks = form->blocks[1];
    return CIE_HIP_OK;
}

int cie_hip_csr_get_launch(cie_hip_csr matrix, int precision, int* kernel, int* width, int* unroll) {
    if (!matrix || (precision != 32 && precision != 64))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_get_launch: invalid matrix or precision");
    const CsrLaunch& launch = launchOf(matrix, precision);
    if (kernel) *kernel = launch.kernel;
    if (width) *width = launch.width;
    if (unroll) *unroll = launch.unroll;
    return CIE_HIP_OK;
}

int cie_hip_csr_autotune(cie_hip_csr matrix, int precision, int kernels, int repetitions, double* bestSeconds) {
    if (!matrix || (precision != 32 && precision != 64))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_autotune: invalid matrix or precision");
    if (precision == 32 && !matrix->entries32)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_autotune: matrix has no FP32 plane");
    if (!(kernels & ((1 << CIE_HIP_SPMV_VECTOR) | (1 << CIE_HIP_SPMV_STAGED) | (1 << CIE_HIP_SPMV_TILED) |
                     (1 << CIE_HIP_SPMV_BLOCKS) | (1 << CIE_HIP_SPMV_SORTED) | (1 << CIE_HIP_SPMV_PAIRED))))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_autotune: empty kernel mask");
    cie_hip_context ctx = matrix->context;
    CIE_HIP_CHECK(hipSetDevice(ctx->device));
    repetitions = std::max(1, repetitions);

    std::vector<CsrLaunch> candidates;
    if ((kernels & (1 << CIE_HIP_SPMV_STAGED)) && matrix->maxRowLength <= kStagedMaxRow) {
        for (int width : {512, 1024})
            for (int unroll : {4, 8}) {
                CsrLaunch c;
                c.kernel = CIE_HIP_SPMV_STAGED;
                c.width = width;
                c.unroll = unroll;
                c.block = kSpmvBlock;
                candidates.push_back(c);
            }
    }
    if (kernels & (1 << CIE_HIP_SPMV_TILED)) {
        for (int unroll : {4, 8, 16}) {
            CsrLaunch c;
            c.kernel = CIE_HIP_SPMV_TILED;
            c.width = kSpmvBlock;
            c.unroll = unroll;
            c.block = kSpmvBlock;
            candidates.push_back(c);
        }
    }
    if (kernels & ((1 << CIE_HIP_SPMV_BLOCKS) | (1 << CIE_HIP_SPMV_SORTED))) {
        // Widths whose blocks hold at most 2 entries' worth of padding per 3 entries (fill <= 5/3).
        for (int width = 2; width <= cie_hip_csr_s::kMaxBlockWidth && width <= matrix->columns; ++width) {
            cie_hip_csr blocks = nullptr;
            const int status = ensureColumnBlocks(matrix, width, &blocks);
            if (status != CIE_HIP_OK) return status;
            if (3 * blocks->nnz * width > 5 * std::max<int64_t>(1, matrix->nnz)) continue;
            for (int kernel : {CIE_HIP_SPMV_BLOCKS, CIE_HIP_SPMV_SORTED}) {
                if (!(kernels & (1 << kernel))) continue;
                for (int unroll : {4, 8, 16}) {
                    if (kernel == CIE_HIP_SPMV_SORTED) {  // sorted tiles only where every tile's columns fit the keys
                        const TileSet* set = nullptr;
                        const int sorted = ensureSortedTiles(blocks, kSpmvBlock * unroll, &set);
                        if (sorted != CIE_HIP_OK) {
                            const TileSet* plain = nullptr;
                            if (ensureTiles(blocks, kSpmvBlock * unroll, &plain) == CIE_HIP_OK && plain->sorted.state < 0)
                                continue;  // too wide for the keys: not a candidate
                            return sorted;
                        }
                    }
                    CsrLaunch c;
                    c.kernel = kernel;
                    c.width = width;
                    c.unroll = unroll;
                    c.block = kSpmvBlock;
                    candidates.push_back(c);
                }
            }
        }
    }
    if (kernels & (1 << CIE_HIP_SPMV_PAIRED)) {
        // Each supported pair and unroll whose form fits the matrix.
        for (const auto& [width, unroll] : std::vector<std::pair<int, int>>{
#define CIEHIP_PAIRED_ENTRY(W1, W2, U) {10 * W1 + W2, U},
                 CIEHIP_PAIRED(CIEHIP_PAIRED_ENTRY)
#undef CIEHIP_PAIRED_ENTRY
             }) {
            if (matrix->columns < pairedWide(width)) continue;
            const PairedForm* form = nullptr;
            const int status = ensurePairedForm(matrix, pairedWide(width), pairedNarrow(width), unroll, &form);
            if (status == CIE_HIP_ERROR_INVALID_ARGUMENT) continue;  // the matrix does not fit this form
            if (status != CIE_HIP_OK) return status;
            CsrLaunch c;
            c.kernel = CIE_HIP_SPMV_PAIRED;
            c.width = width;
            c.unroll = unroll;
            c.block = kSpmvBlock;
            candidates.push_back(c);
        }
    }
    if (kernels & (1 << CIE_HIP_SPMV_VECTOR)) {
        // Lanes within a factor 4 of about one pass of 4 entries per lane per row.
        const int seed = std::min(64, nextPow2(int(std::ceil(std::max(1.0, matrix->meanRowLength) / 4))));
        for (int lanes = std::max(1, seed / 4); lanes <= std::min(64, seed * 4); lanes *= 2)
            for (int unroll : {1, 2, 4, 8}) {
                CsrLaunch c;
                c.kernel = CIE_HIP_SPMV_VECTOR;
                c.width = lanes;
                c.unroll = unroll;
                c.block = kSpmvBlock;
                candidates.push_back(c);
            }
    }
    if (candidates.empty()) {
        releaseUnusedColumnBlocks(matrix);
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_autotune: no candidate for this matrix");
    }

    const size_t valueSize = precision == 32 ? sizeof(float) : sizeof(double);
    void* buffer = nullptr;
    const int64_t inCount = std::max<int64_t>(1, matrix->columns);
    const size_t bytes = valueSize * size_t(inCount + std::max<int64_t>(1, matrix->rows));
    CIE_HIP_CHECK(hipMalloc(&buffer, bytes));
    ScopedEvents<2> events;
    if (const hipError_t error = events.create(); error != hipSuccess || hipMemsetAsync(buffer, 0, bytes, ctx->stream) != hipSuccess) {
        (void)hipFree(buffer);
        return setError(CIE_HIP_ERROR_HIP, "cie_hip_csr_autotune: timing set-up failed");
    }
    const hipEvent_t start = events.event[0], stop = events.event[1];

    double best = std::numeric_limits<double>::infinity();
    CsrLaunch bestLaunch = candidates.front();
    int status = CIE_HIP_OK;
    for (const CsrLaunch& c : candidates) {
        for (int pass = 0; pass < 2 && status == CIE_HIP_OK; ++pass) {  // pass 0 warms up
            if (hipEventRecord(start, ctx->stream) != hipSuccess) { status = CIE_HIP_ERROR_HIP; break; }
            for (int r = 0; r < (pass ? repetitions : 1) && status == CIE_HIP_OK; ++r) {
                if (precision == 32) {
                    float* in = static_cast<float*>(buffer);
                    status = launchProduct<float>(matrix, c, 0.0f, in, 1.0f, in + inCount, matrix->entries32);
                } else {
                    double* in = static_cast<double*>(buffer);
                    status = launchProduct<double>(matrix, c, 0.0, in, 1.0, in + inCount, matrix->entries64);
                }
            }
            if (hipEventRecord(stop, ctx->stream) != hipSuccess || hipEventSynchronize(stop) != hipSuccess) {
                status = CIE_HIP_ERROR_HIP;
                break;
            }
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, start, stop);
            if (pass && ms * 1e-3 / repetitions < best) {
                best = ms * 1e-3 / repetitions;
                bestLaunch = c;
            }
        }
        if (status != CIE_HIP_OK) break;
    }
    (void)hipFree(buffer);
    if (status != CIE_HIP_OK) {
        releaseUnusedColumnBlocks(matrix);
        return setError(status, std::string("cie_hip_csr_autotune: ") + cie_hip_last_error());
    }
    launchOf(matrix, precision) = bestLaunch;
    releaseUnusedColumnBlocks(matrix);
    if (bestSeconds) *bestSeconds = best;
    return CIE_HIP_OK;
}

int cie_hip_csr_product_f64(cie_hip_context context, cie_hip_csr matrix, double inScale, const double* in,
                            int64_t inSize, double outScale, double* out, int64_t outSize) {
    if (!context || !matrix) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_product_f64: null handle");
    CIE_HIP_REQUIRE_CONTEXT(context, matrix->context, "cie_hip_csr_product_f64");
    if (matrix->rows != outSize || matrix->columns != inSize)
        return setError(CIE_HIP_ERROR_SHAPE, "Incompatible matrix-vector product: [" + std::to_string(matrix->rows) + "x"
                                                 + std::to_string(matrix->columns) + "] * [" + std::to_string(inSize)
                                                 + "] = [" + std::to_string(outSize) + "]");
    if (!matrix->rows) return CIE_HIP_OK;
    if (!in || !out) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_product_f64: null vector");
    if (!matrix->entries64) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_product_f64: matrix has no FP64 plane");
    CIE_HIP_CHECK(hipSetDevice(context->device));
    return launchProduct<double>(matrix, matrix->launch64, inScale, in, outScale, out, matrix->entries64);
}

int cie_hip_csr_product_f32(cie_hip_context context, cie_hip_csr matrix, float inScale, const float* in,
                            int64_t inSize, float outScale, float* out, int64_t outSize) {
    if (!context || !matrix) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_product_f32: null handle");
    CIE_HIP_REQUIRE_CONTEXT(context, matrix->context, "cie_hip_csr_product_f32");
    if (matrix->rows != outSize || matrix->columns != inSize)
        return setError(CIE_HIP_ERROR_SHAPE, "Incompatible matrix-vector product: [" + std::to_string(matrix->rows) + "x"
                                                 + std::to_string(matrix->columns) + "] * [" + std::to_string(inSize)
                                                 + "] = [" + std::to_string(outSize) + "]");
    if (!matrix->rows) return CIE_HIP_OK;
    if (!in || !out) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_product_f32: null vector");
    if (!matrix->entries32) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_product_f32: matrix has no FP32 plane");
    CIE_HIP_CHECK(hipSetDevice(context->device));
    return launchProduct<float>(matrix, matrix->launch32, inScale, in, outScale, out, matrix->entries32);
}

int cie_hip_inverse_diagonal_f64(cie_hip_context context, cie_hip_csr matrix, double* inverse, int64_t* failedRow) {
    if (!context || !matrix || (!inverse && matrix->rows))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_inverse_diagonal_f64: null argument");
    CIE_HIP_REQUIRE_CONTEXT(context, matrix->context, "cie_hip_inverse_diagonal_f64");
    if (matrix->rows != matrix->columns) return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_inverse_diagonal_f64: matrix is not square");
    CIE_HIP_CHECK(hipSetDevice(context->device));
    return inverseDiagonal<double>(matrix, inverse, failedRow);
}

int cie_hip_inverse_diagonal_f32(cie_hip_context context, cie_hip_csr matrix, float* inverse, int64_t* failedRow) {
    if (!context || !matrix || (!inverse && matrix->rows))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_inverse_diagonal_f32: null argument");
    CIE_HIP_REQUIRE_CONTEXT(context, matrix->context, "cie_hip_inverse_diagonal_f32");
    if (matrix->rows != matrix->columns) return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_inverse_diagonal_f32: matrix is not square");
    CIE_HIP_CHECK(hipSetDevice(context->device));
    return inverseDiagonal<float>(matrix, inverse, failedRow);
}

int cie_hip_csr_row_sum_correction(cie_hip_context context, cie_hip_csr matrix, float* correction) {
    if (!context || !matrix || (!correction && matrix->rows))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_row_sum_correction: null argument");
    CIE_HIP_REQUIRE_CONTEXT(context, matrix->context, "cie_hip_csr_row_sum_correction");
    if (!matrix->entries64)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_row_sum_correction: matrix has no FP64 plane");
    CIE_HIP_CHECK(hipSetDevice(context->device));
    if (const int status = ensureInverseDiagonal(matrix, 32)) return status;
    if (!matrix->rows) return CIE_HIP_OK;
    CIE_HIP_CHECK(hipMemcpyAsync(correction, matrix->rowCorrection32, sizeof(float) * size_t(matrix->rows),
                                 hipMemcpyDeviceToDevice, context->stream));
    CIE_HIP_CHECK(hipStreamSynchronize(context->stream));
    return CIE_HIP_OK;
}

}  // extern "C"

// The paired column-block form of a CSR matrix, read by PairedBlockTiles:
// rows covered by 1 x wide and 1 x narrow blocks of adjacent columns, cut into
// tiles of whole rows, each width's blocks column-sorted within its tile.
//
// Cover. A row is walked in column order by a group of kCoverLanes lanes (the
// row's columns loaded kCoverLanes at a time with the next kCoverLanes as a
// look-ahead, the next block starting at the first lane whose column lies
// past the blocks so far: a ballot and a find-first-set; its width and entry
// count are ballots over the look-ahead window). A block starts at the column of the first entry not yet
// covered (moved left to columns - width at the last columns) and is wide
// when an entry of [start + narrow, start + wide) follows, that is when a
// narrow block would leave part of the wide span to a further block. The
// count and structure passes run the same walk, so they agree block for block.
//
// Tiles. Consecutive rows, at most kSpmvBlock of them and at most
// kSpmvBlock * unroll blocks of each width; a row with more blocks of one
// width than that does not fit the form.
//
// Sort. One workgroup per tile and width sorts that width's blocks by (first
// column, position) with a bitonic network in LDS and writes each block's key
// (column offset from the width's smallest column in the tile, position among
// all the tile's blocks in row order), its first column, first CSR entry and
// entry count (the refill reads the CSR through them).
//
// Values. One workgroup per tile visits the tile's blocks in row order (so
// the CSR is read in order), gathers each block's entries into its width
// values (zero where the row has no entry), stages them in LDS at the block's
// sorted slot and writes each plane's run of the tile coalesced (PlaneRun
// layout); once for the FP64 plane and, when the matrix has one, once for
// the FP32 plane; rerun whenever the entries change.
#include "dispatch.hpp"
#include "kernels.hpp"

#include <climits>
#include <limits>
#include <vector>

using namespace ciehip;

namespace {

constexpr int kCoverLanes = 16;
constexpr int kSortThreads = 256;
constexpr int kPairedSortMax = 2048;  // blocks of one width per tile at most (kSpmvBlock * 8)
constexpr int kMaxPairedWidth = 8;
constexpr int kValuesThreads = 256;
constexpr int kStageValues = 5120;  // values staged per pass: 1024 blocks of 1 x 5 (40 KB in FP64)

// Calls emit(wide, index in row, start column, first entry, entries) for
// every block of `row`, in column order, from the group's first lane. The
// group holds a window of 2 * kCoverLanes columns (lane l: entries base + l
// and base + kCoverLanes + l); a block's entries lie within `wide` <=
// kMaxPairedWidth positions of its first one, inside the window, so the
// width and entry count come from ballots over the window, not loads.
template <class Emit>
__device__ __forceinline__ void pairedCover(int row, const int32_t* __restrict__ extents,
                                            const int32_t* __restrict__ columnIndices, int wide, int narrow,
                                            int columns, int lane, int groupShift, Emit&& emit) {
    static_assert(kMaxPairedWidth <= kCoverLanes, "a block's entries must fit the window past its first lane");
    const int begin = extents[row], end = extents[row + 1];
    int limit = INT_MIN;  // columns below this are covered
    int index = 0;
    const uint32_t groupMask = (1u << kCoverLanes) - 1;
    auto groupBits = [&](bool predicate) { return uint32_t(__ballot(predicate) >> groupShift) & groupMask; };
    for (int base = begin; base < end; base += kCoverLanes) {
        const int k = base + lane;
        const int column = k < end ? columnIndices[k] : INT_MAX;
        const int ahead = k + kCoverLanes < end ? columnIndices[k + kCoverLanes] : INT_MAX;
        const int high = lane + kCoverLanes;  // window position of `ahead`
        for (;;) {
            const uint32_t mine = groupBits(k < end && column >= limit);
            if (!mine) break;
            const int firstLane = __ffs(int(mine)) - 1;
            const int next = __shfl(column, (groupShift & (kWave - 1)) + firstLane, kWave);
            const int entry = base + firstLane;
            // Columns rise along the row, so the entries of [next, next + wide) are the next few positions.
            auto inWideTail = [&](int position, int c) {
                return position > firstLane && position < firstLane + wide && c >= next + narrow && c < next + wide;
            };
            const bool isWide = (groupBits(inWideTail(lane, column)) | groupBits(inWideTail(high, ahead))) != 0;
            const int width = isWide ? wide : narrow;
            const int start = next + width > columns ? columns - width : next;
            auto inBlock = [&](int position, int c) {
                return position >= firstLane && position < firstLane + width && c < start + width;
            };
            const int count = __popc(groupBits(inBlock(lane, column))) + __popc(groupBits(inBlock(high, ahead)));
            if (lane == 0) emit(isWide, index, start, entry, count);
            ++index;
            limit = start + width;
        }
    }
}

// Blocks of each width per row: counts[2 row] wide, counts[2 row + 1] narrow.
__global__ __launch_bounds__(kReduceBlock) void pairedCountKernel(int rows, const int32_t* __restrict__ extents,
                                                                  const int32_t* __restrict__ columnIndices, int wide,
                                                                  int narrow, int columns, int32_t* __restrict__ counts) {
    const int lane = int(threadIdx.x) % kCoverLanes;
    const int groupShift = int(threadIdx.x) % kWave - lane;  // the group's first lane in the wavefront
    const int groupsPerBlock = kReduceBlock / kCoverLanes;
    const int stride = int(gridDim.x) * groupsPerBlock;
    for (int row = int(blockIdx.x) * groupsPerBlock + int(threadIdx.x) / kCoverLanes; row < rows; row += stride) {
        int wideCount = 0, narrowCount = 0;
        pairedCover(row, extents, columnIndices, wide, narrow, columns, lane, groupShift,
                    [&](bool isWide, int, int, int, int) { ++(isWide ? wideCount : narrowCount); });
        if (lane == 0) {
            counts[2 * row] = wideCount;
            counts[2 * row + 1] = narrowCount;
        }
    }
}

// The blocks of one width, in row order: first column, first CSR entry,
// entries, and position among all blocks of the matrix in row order.
struct WidthBlocks {
    int32_t* start;
    int32_t* entry;
    uint8_t* count;
    int32_t* merged;
};

__global__ __launch_bounds__(kReduceBlock) void pairedStartsKernel(int rows, const int32_t* __restrict__ extents,
                                                                   const int32_t* __restrict__ columnIndices, int wide,
                                                                   int narrow, int columns,
                                                                   const int32_t* __restrict__ wideExtents,
                                                                   const int32_t* __restrict__ narrowExtents,
                                                                   WidthBlocks wideBlocks, WidthBlocks narrowBlocks) {
    const int lane = int(threadIdx.x) % kCoverLanes;
    const int groupShift = int(threadIdx.x) % kWave - lane;
    const int groupsPerBlock = kReduceBlock / kCoverLanes;
    const int stride = int(gridDim.x) * groupsPerBlock;
    for (int row = int(blockIdx.x) * groupsPerBlock + int(threadIdx.x) / kCoverLanes; row < rows; row += stride) {
        int w = wideExtents[row], n = narrowExtents[row];
        const int merged = w + n;
        pairedCover(row, extents, columnIndices, wide, narrow, columns, lane, groupShift,
                    [&](bool isWide, int index, int start, int entry, int count) {
                        WidthBlocks& to = isWide ? wideBlocks : narrowBlocks;
                        const int b = isWide ? w++ : n++;
                        to.start[b] = start;
                        to.entry[b] = entry;
                        to.count[b] = uint8_t(count);
                        to.merged[b] = merged + index;
                    });
    }
}

// One workgroup per (tile, width): that width's blocks of the tile sorted by
// (first column, position); keys and refill data in sorted order. A tile
// whose blocks of one width span more columns than the keys hold sets *failed.
__global__ __launch_bounds__(kSortThreads) void pairedSortKernel(const int2* __restrict__ tiles,
                                                                 const int2* __restrict__ begins,
                                                                 WidthBlocks wideBlocks, WidthBlocks narrowBlocks,
                                                                 int64_t wideCount, int32_t* __restrict__ bases,
                                                                 uint32_t* __restrict__ keys,
                                                                 int32_t* __restrict__ sortedStart,
                                                                 int32_t* __restrict__ sortedEntry,
                                                                 uint8_t* __restrict__ sortedCount,
                                                                 int* __restrict__ failed) {
    __shared__ uint64_t items[kPairedSortMax];
    const int t = blockIdx.x;
    const bool isWide = blockIdx.y == 0;
    const WidthBlocks& from = isWide ? wideBlocks : narrowBlocks;
    const int begin = isWide ? begins[t].x : begins[t].y;
    const int count = (isWide ? begins[t + 1].x : begins[t + 1].y) - begin;
    const int mergedBegin = tiles[t].y;
    const int64_t offset = isWide ? 0 : wideCount;
    if (count == 0) {
        if (threadIdx.x == 0) bases[2 * t + blockIdx.y] = 0;
        return;
    }
    int n = 1;
    while (n < count) n <<= 1;
    for (int i = threadIdx.x; i < n; i += kSortThreads)
        items[i] = i < count ? (uint64_t(uint32_t(from.start[begin + i])) << kSortedPositionBits) | uint64_t(i)
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
        bases[2 * t + blockIdx.y] = int32_t(base);
        if (int64_t(items[count - 1] >> kSortedPositionBits) - base > kSortedMaxSpan) atomicOr(failed, 1);
    }
    for (int i = threadIdx.x; i < count; i += kSortThreads) {
        const uint64_t item = items[i];
        const int block = begin + int(item & kSortedPositionMask);
        const uint32_t column = uint32_t(int64_t(item >> kSortedPositionBits) - base);
        const int64_t slot = offset + begin + i;
        keys[slot] = (column << kSortedPositionBits) | uint32_t(from.merged[block] - mergedBegin);
        sortedStart[slot] = from.start[block];
        sortedEntry[slot] = from.entry[block];
        sortedCount[slot] = from.count[block];
    }
}

// The values of one tile's blocks, one precision, into their planes: one
// workgroup per tile. The blocks are visited in row order (their positions in
// the tile, found through the sorted keys), so consecutive lanes read
// consecutive CSR entries. Each width's values are staged in LDS at their
// sorted slots, kStageValues values per pass, and each plane's run of the
// tile is then written out coalesced.
template <class T>
__global__ __launch_bounds__(kValuesThreads) void pairedValuesKernel(
    const int2* __restrict__ begins, const uint32_t* __restrict__ keys, const int32_t* __restrict__ sortedStart,
    const int32_t* __restrict__ sortedEntry, const uint8_t* __restrict__ sortedCount, int64_t wideBlocks,
    int64_t narrowBlocks, int wide, int narrow, const int32_t* __restrict__ columnIndices,
    const T* __restrict__ entries, T* __restrict__ planes) {
    constexpr int kChunk = 16 / int(sizeof(T));  // values per plane (PlaneRun)
    __shared__ int16_t order[2 * kPairedSortMax];  // position in the tile -> the tile's wide, then narrow, slots
    __shared__ T stage[kStageValues];
    const int t = blockIdx.x;
    const int tid = threadIdx.x;
    const int2 head = begins[t], next = begins[t + 1];
    const int wideCount = next.x - head.x, total = wideCount + (next.y - head.y);
    for (int i = tid; i < total; i += kValuesThreads) {
        const int64_t slot = i < wideCount ? int64_t(head.x) + i : wideBlocks + head.y + (i - wideCount);
        order[keys[slot] & kSortedPositionMask] = int16_t(i);
    }
    __syncthreads();
    for (int w = 0; w < 2; ++w) {
        const bool isWide = w == 0;
        const int width = isWide ? wide : narrow;
        const int64_t blocks = isWide ? wideBlocks : narrowBlocks;
        const int first = isWide ? 0 : wideCount;             // the width's slots among the tile's
        const int count = isWide ? wideCount : total - wideCount;
        const int64_t begin = isWide ? head.x : head.y;       // the tile's first slot within the width
        T* __restrict__ widthPlanes = planes + (isWide ? 0 : wideBlocks * wide);
        const int perPass = kStageValues / width;
        for (int pass = 0; pass < count; pass += perPass) {
            const int n = min(perPass, count - pass);
            for (int position = tid; position < total; position += kValuesThreads) {
                const int k = order[position] - first - pass;  // slot within this pass
                if (k < 0 || k >= n) continue;
                const int64_t slot = (isWide ? 0 : wideBlocks) + begin + pass + k;
                const int start = sortedStart[slot], entry = sortedEntry[slot], held = sortedCount[slot];
                T v[kMaxPairedWidth] = {};
                for (int j = 0; j < held; ++j) {
                    const int c = columnIndices[entry + j] - start;
                    const T e = entries[entry + j];
#pragma unroll
                    for (int u = 0; u < kMaxPairedWidth; ++u)  // register indices only
                        if (c == u) v[u] += e;
                }
#pragma unroll
                for (int u = 0; u < kMaxPairedWidth; ++u)
                    if (u < width) stage[k * width + u] = v[u];
            }
            __syncthreads();
            for (int p = 0; p * kChunk < width; ++p) {
                const int length = min(kChunk, width - p * kChunk);
                T* __restrict__ run = widthPlanes + blocks * (int64_t(p) * kChunk) + (begin + pass) * length;
                for (int i = tid; i < n * length; i += kValuesThreads) {
                    const int k = i / length, j = i - k * length;
                    run[i] = stage[k * width + p * kChunk + j];
                }
            }
            __syncthreads();  // the stage is reused by the next pass
        }
    }
}

// Device arrays freed at scope exit unless kept.
struct Allocations {
    std::vector<void*> pointers;
    ~Allocations() {
        for (void* p : pointers)
            if (p) (void)hipFree(p);
    }
    template <class T>
    hipError_t allocate(T** pointer, size_t count) {
        *pointer = nullptr;
        const hipError_t error = hipMalloc(pointer, sizeof(T) * std::max<size_t>(1, count));
        if (error == hipSuccess) pointers.push_back(*pointer);
        return error;
    }
    void keep(const void* pointer) {
        for (void*& p : pointers)
            if (p == pointer) p = nullptr;
    }
};

int coverGrid(cie_hip_context ctx, int rows) {
    constexpr int groupsPerBlock = kReduceBlock / kCoverLanes;
    const int64_t blocks = (int64_t(rows) + groupsPerBlock - 1) / groupsPerBlock;
    return int(std::max<int64_t>(1, std::min<int64_t>(blocks, int64_t(std::max(1, ctx->computeUnits)) * 32)));
}

int refill(cie_hip_csr m, PairedForm& form) {
    cie_hip_context ctx = m->context;
    if (form.tileCount && form.blocks[0] + form.blocks[1]) {
        const dim3 grid(unsigned(form.tileCount)), block(kValuesThreads);
        hipLaunchKernelGGL((pairedValuesKernel<double>), grid, block, 0, ctx->stream, form.begins, form.keys,
                           form.sortedStart, form.sortedEntry, form.sortedCount, form.blocks[0], form.blocks[1],
                           form.wide, form.narrow, m->columnIndices, static_cast<const double*>(m->entries64),
                           form.values64);
        CIE_HIP_CHECK(hipGetLastError());
        if (form.values32) {
            hipLaunchKernelGGL((pairedValuesKernel<float>), grid, block, 0, ctx->stream, form.begins, form.keys,
                               form.sortedStart, form.sortedEntry, form.sortedCount, form.blocks[0], form.blocks[1],
                               form.wide, form.narrow, m->columnIndices, static_cast<const float*>(m->entries32),
                               form.values32);
            CIE_HIP_CHECK(hipGetLastError());
        }
    }
    form.version = m->valuesVersion;
    return CIE_HIP_OK;
}

}  // namespace

namespace ciehip {

// Builds the structure of `form` (cover, tiles cut as `cut` asks, sort); on a
// matrix that does not fit, leaves state -1 and the reason in form.error.
int buildPairedForm(cie_hip_csr m, PairedForm& form, const PairedCut& cut) {
    cie_hip_context ctx = m->context;
    const int rows = int(m->rows), columns = int(m->columns);
    const int perWidth = kSpmvBlock * form.unroll;
    Allocations scratch;
    int32_t* counts = nullptr;
    CIE_HIP_CHECK(scratch.allocate(&counts, 2 * size_t(rows)));
    if (rows) {
        hipLaunchKernelGGL(pairedCountKernel, dim3(coverGrid(ctx, rows)), dim3(kReduceBlock), 0, ctx->stream, rows,
                           m->rowExtents, m->columnIndices, form.wide, form.narrow, columns, counts);
        CIE_HIP_CHECK(hipGetLastError());
    }
    std::vector<int32_t> hostCounts(2 * size_t(rows));
    if (rows)
        CIE_HIP_CHECK(hipMemcpyAsync(hostCounts.data(), counts, sizeof(int32_t) * hostCounts.size(),
                                     hipMemcpyDeviceToHost, ctx->stream));
    CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    std::vector<int32_t> extents[2] = {std::vector<int32_t>(size_t(rows) + 1, 0),
                                       std::vector<int32_t>(size_t(rows) + 1, 0)};
    std::vector<int32_t> merged(size_t(rows) + 1, 0);
    for (int row = 0; row < rows; ++row) {
        for (int w = 0; w < 2; ++w) {
            const int32_t c = hostCounts[2 * size_t(row) + w];
            if (c > perWidth) {
                form.state = -1;
                form.error = "paired tiles: row " + std::to_string(row) + " holds " + std::to_string(c) +
                             " blocks of one width, more than a tile's " + std::to_string(perWidth);
                return CIE_HIP_OK;
            }
            extents[w][size_t(row) + 1] = extents[w][size_t(row)] + c;  // <= nnz <= INT32_MAX
        }
        merged[size_t(row) + 1] = extents[0][size_t(row) + 1] + extents[1][size_t(row) + 1];
    }
    form.blocks[0] = extents[0][size_t(rows)];
    form.blocks[1] = extents[1][size_t(rows)];
    const int64_t total = form.blocks[0] + form.blocks[1];

    // Tiles: consecutive rows of one system, within the row cap and the per-width block capacities.
    std::vector<int2> tiles, begins;
    if (cut.systemTiles) cut.systemTiles->assign(size_t(cut.systems) + 1, 0);
    for (int system = 0; system < cut.systems; ++system) {
        const int lo = cut.offsets ? int(cut.offsets[system]) : 0;
        const int hi = cut.offsets ? int(cut.offsets[system + 1]) : rows;
        const int cap = cut.rowCaps ? std::min(cut.rowCaps[system], kSpmvBlock) : kSpmvBlock;
        if (cut.systemTiles) (*cut.systemTiles)[size_t(system)] = int32_t(tiles.size());
        for (int row = lo; row < hi;) {
            const int first = row;
            while (row < hi && row - first < cap &&
                   extents[0][size_t(row) + 1] - extents[0][size_t(first)] <= perWidth &&
                   extents[1][size_t(row) + 1] - extents[1][size_t(first)] <= perWidth)
                ++row;
            tiles.push_back(make_int2(first, merged[size_t(first)]));
            begins.push_back(make_int2(extents[0][size_t(first)], extents[1][size_t(first)]));
        }
    }
    if (cut.systemTiles) (*cut.systemTiles)[size_t(cut.systems)] = int32_t(tiles.size());
    tiles.push_back(make_int2(rows, merged[size_t(rows)]));
    begins.push_back(make_int2(extents[0][size_t(rows)], extents[1][size_t(rows)]));
    form.tileCount = int(tiles.size()) - 1;

    int32_t *wideExtents = nullptr, *narrowExtents = nullptr;
    WidthBlocks blocks[2] = {};
    CIE_HIP_CHECK(scratch.allocate(&wideExtents, size_t(rows) + 1));
    CIE_HIP_CHECK(scratch.allocate(&narrowExtents, size_t(rows) + 1));
    for (int w = 0; w < 2; ++w) {
        const size_t n = size_t(form.blocks[w]);
        CIE_HIP_CHECK(scratch.allocate(&blocks[w].start, n));
        CIE_HIP_CHECK(scratch.allocate(&blocks[w].entry, n));
        CIE_HIP_CHECK(scratch.allocate(&blocks[w].count, n));
        CIE_HIP_CHECK(scratch.allocate(&blocks[w].merged, n));
    }
    CIE_HIP_CHECK(scratch.allocate(&form.extents, size_t(rows) + 1));
    CIE_HIP_CHECK(scratch.allocate(&form.tiles, tiles.size()));
    CIE_HIP_CHECK(scratch.allocate(&form.begins, begins.size()));
    CIE_HIP_CHECK(scratch.allocate(&form.bases, size_t(std::max(1, form.tileCount))));
    CIE_HIP_CHECK(scratch.allocate(&form.keys, size_t(total)));
    CIE_HIP_CHECK(scratch.allocate(&form.sortedStart, size_t(total)));
    CIE_HIP_CHECK(scratch.allocate(&form.sortedEntry, size_t(total)));
    CIE_HIP_CHECK(scratch.allocate(&form.sortedCount, size_t(total)));
    CIE_HIP_CHECK(scratch.allocate(&form.values64, size_t(form.blocks[0] * form.wide + form.blocks[1] * form.narrow)));
    if (m->entries32)
        CIE_HIP_CHECK(scratch.allocate(&form.values32, size_t(form.blocks[0] * form.wide + form.blocks[1] * form.narrow)));
    CIE_HIP_CHECK(hipMemcpyAsync(wideExtents, extents[0].data(), sizeof(int32_t) * (size_t(rows) + 1),
                                 hipMemcpyHostToDevice, ctx->stream));
    CIE_HIP_CHECK(hipMemcpyAsync(narrowExtents, extents[1].data(), sizeof(int32_t) * (size_t(rows) + 1),
                                 hipMemcpyHostToDevice, ctx->stream));
    CIE_HIP_CHECK(hipMemcpyAsync(form.extents, merged.data(), sizeof(int32_t) * (size_t(rows) + 1),
                                 hipMemcpyHostToDevice, ctx->stream));
    CIE_HIP_CHECK(hipMemcpyAsync(form.tiles, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice,
                                 ctx->stream));
    CIE_HIP_CHECK(hipMemcpyAsync(form.begins, begins.data(), sizeof(int2) * begins.size(), hipMemcpyHostToDevice,
                                 ctx->stream));
    if (rows) {
        hipLaunchKernelGGL(pairedStartsKernel, dim3(coverGrid(ctx, rows)), dim3(kReduceBlock), 0, ctx->stream, rows,
                           m->rowExtents, m->columnIndices, form.wide, form.narrow, columns, wideExtents,
                           narrowExtents, blocks[0], blocks[1]);
        CIE_HIP_CHECK(hipGetLastError());
    }
    int* failed = reinterpret_cast<int*>(ctx->scratch.scalars);
    int* pinned = reinterpret_cast<int*>(ctx->scratch.pinned);
    CIE_HIP_CHECK(hipMemsetAsync(failed, 0, sizeof(int), ctx->stream));
    if (form.tileCount) {
        hipLaunchKernelGGL(pairedSortKernel, dim3(unsigned(form.tileCount), 2), dim3(kSortThreads), 0, ctx->stream,
                           form.tiles, form.begins, blocks[0], blocks[1], form.blocks[0],
                           reinterpret_cast<int32_t*>(form.bases), form.keys, form.sortedStart, form.sortedEntry,
                           form.sortedCount, failed);
        CIE_HIP_CHECK(hipGetLastError());
    }
    CIE_HIP_CHECK(hipMemcpyAsync(pinned, failed, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (pinned[0]) {  // the form's arrays are freed with the scratch
        form = PairedForm{form.wide, form.narrow, form.unroll};
        form.state = -1;
        form.error = "paired tiles: a tile's blocks of one width span more than " + std::to_string(kSortedMaxSpan) +
                     " columns";
        return CIE_HIP_OK;
    }
    for (const void* p : {static_cast<const void*>(form.extents), static_cast<const void*>(form.tiles),
                          static_cast<const void*>(form.begins), static_cast<const void*>(form.bases),
                          static_cast<const void*>(form.keys), static_cast<const void*>(form.sortedStart),
                          static_cast<const void*>(form.sortedEntry), static_cast<const void*>(form.sortedCount),
                          static_cast<const void*>(form.values64), static_cast<const void*>(form.values32)})
        scratch.keep(p);
    form.state = 1;
    form.version = std::numeric_limits<uint64_t>::max();  // values not written yet
    return CIE_HIP_OK;
}

int refreshPairedValues(cie_hip_csr m, PairedForm& form) {
    if (m->entries32 && !form.values32) {  // an FP32 plane added after the form was built
        CIE_HIP_CHECK(hipMalloc(&form.values32, sizeof(float) * size_t(std::max<int64_t>(
                                                    1, form.blocks[0] * form.wide + form.blocks[1] * form.narrow))));
        form.version = std::numeric_limits<uint64_t>::max();
    }
    return form.version == m->valuesVersion ? CIE_HIP_OK : refill(m, form);
}

void releasePairedForm(cie_hip_context context, PairedForm* form) {
    if (!form) return;
    if (context) (void)hipStreamSynchronize(context->stream);
    for (void* p : {static_cast<void*>(form->extents), static_cast<void*>(form->tiles),
                    static_cast<void*>(form->begins), static_cast<void*>(form->bases), static_cast<void*>(form->keys),
                    static_cast<void*>(form->sortedStart), static_cast<void*>(form->sortedEntry),
                    static_cast<void*>(form->sortedCount), static_cast<void*>(form->values64),
                    static_cast<void*>(form->values32)})
        if (p) (void)hipFree(p);
    delete form;
}

int ensurePairedForm(cie_hip_csr m, int wide, int narrow, int unroll, const PairedForm** out) {
    if (m->blockWidth) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "paired tiles of a column-block form");
    if (!(narrow >= 1 && narrow < wide && wide <= kMaxPairedWidth) || kSpmvBlock * unroll > kPairedSortMax ||
        unroll < 1)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "paired tiles: unsupported widths or unroll");
    if (m->columns < wide)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "paired tiles: the matrix has fewer columns than the wide block");
    PairedForm* form = nullptr;
    for (PairedForm* f : m->pairedForms)
        if (f->wide == wide && f->narrow == narrow && f->unroll == unroll) form = f;
    if (!form) {
        form = new PairedForm;
        form->wide = wide;
        form->narrow = narrow;
        form->unroll = unroll;
        m->pairedForms.push_back(form);
    }
    if (form->state < 0) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, form->error);
    if (form->state == 0) {
        const int status = buildPairedForm(m, *form, PairedCut{});
        if (status != CIE_HIP_OK) {
            *form = PairedForm{wide, narrow, unroll};  // nothing kept: built again on the next use
            return status;
        }
        if (form->state < 0) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, form->error);
    }
    const int status = refreshPairedValues(m, *form);
    if (status != CIE_HIP_OK) return status;
    *out = form;
    return CIE_HIP_OK;
}

}  // namespace ciehip

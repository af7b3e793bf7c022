// Device kernels of libcie_hip (CDNA4 / gfx950, wave64).
//
// Determinism: every sum below has a fixed association order that depends
// only on the problem size and the launch shape, never on scheduling:
//   * a CSR row is split over L lanes; lane l sums entries l, l+L, l+2L, ...
//     in increasing order, then an xor-butterfly combines the L lane sums
//     (every lane of the group ends with the same bits, since a+b == b+a);
//   * a grid reduction writes one partial per block (in-block order fixed),
//     and a one-block kernel launched after it sums the partials in block
//     order; no value is ever accumulated with atomics.
#pragma once

#include "common.hpp"

#include <cfloat>
#include <type_traits>

namespace ciehip {

// Narrow column plane: the columns of each run of kNarrowRun consecutive
// entries are stored as 16-bit offsets from the run's smallest column, when
// every run's columns span at most 65535 (banded FEM matrices in their natural
// node order do: the 3D Q2 64^3 matrix spans at most 65,160 per run). A block
// tile of at most kNarrowRun entries touches at most two runs.
constexpr int kNarrowShift = 12;
constexpr int kNarrowRun = 1 << kNarrowShift;

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------

template <int W, class T>
__device__ __forceinline__ T groupSum(T value) {
#pragma unroll
    for (int offset = W / 2; offset > 0; offset /= 2) value += __shfl_xor(value, offset, W);
    return value;
}

// Sum of `value` over the block, valid in thread 0. `lds` holds BS/64 entries.
template <int BS, class T>
__device__ __forceinline__ T blockSum(T value, T* lds) {
    value = groupSum<kWave>(value);
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = value;
    __syncthreads();
    T total = T(0);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int w = 0; w < BS / kWave; ++w) total += lds[w];
    }
    return total;
}

// Sum partials[0..count) with stride `stride` in a fixed order; valid in thread 0.
template <int BS, class T>
__device__ __forceinline__ T sumPartials(const T* partials, int count, int stride, T* lds) {
    T value = T(0);
    for (int i = threadIdx.x; i < count; i += BS) value += partials[size_t(i) * stride];
    return blockSum<BS>(value, lds);
}

// cieutils/packages/maths/impl/Comparison_impl.hpp:49-63, default tolerances.
__device__ __forceinline__ bool floatEqual(double left, double right) {
    const double norm = fmin(fabs(left) + fabs(right), DBL_MAX);
    return fabs(left - right) < fmax(DBL_MIN, DBL_EPSILON * norm);
}

__device__ __forceinline__ bool floatLess(double left, double right) {
    return left < right && !floatEqual(left, right);
}

// Matrix entries and column indices are read once per product. Plain loads
// measured faster than non-temporal ones on gfx950 (the streaming hint is
// kept as a build option).
#ifdef CIEHIP_STREAMING_MATRIX_LOADS
#define CIEHIP_STREAM_LOAD(pointer) __builtin_nontemporal_load(pointer)
#else
#define CIEHIP_STREAM_LOAD(pointer) (*(pointer))
#endif

// Blocks are dealt to the 8 XCDs round-robin; give each XCD a contiguous run
// of chunks so that its L2 sees a contiguous window of the gathered vector.
__device__ __forceinline__ int xcdChunk() {
    const int grid = gridDim.x;
#ifdef CIEHIP_NO_XCD_REMAP
    return blockIdx.x;
#endif
    if (grid % kXcds) return blockIdx.x;
    return (blockIdx.x % kXcds) * (grid / kXcds) + blockIdx.x / kXcds;
}

// ---------------------------------------------------------------------------
// CSR row product: L lanes per row, U entries per lane loaded before use.
// ---------------------------------------------------------------------------

template <class TV, class TM, int L, int U>
__device__ __forceinline__ TV rowProduct(int row, int lane, const int32_t* __restrict__ rowExtents,
                                         const int32_t* __restrict__ columnIndices,
                                         const TM* __restrict__ entries, const TV* __restrict__ in) {
    const int begin = rowExtents[row];
    const int end = rowExtents[row + 1];
    TV sum = TV(0);
    for (int base = begin; base < end; base += L * U) {
        int column[U];
        TM entry[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = base + u * L + lane;
            if (k < end) {
                column[u] = CIEHIP_STREAM_LOAD(columnIndices + k);
                entry[u] = CIEHIP_STREAM_LOAD(entries + k);
            } else {
                // Inside the loop the row has at least one entry, so the
                // matrix has a column 0 to gather harmlessly from.
                column[u] = 0;
                entry[u] = TM(0);
            }
        }
        TV gathered[U];
#pragma unroll
        for (int u = 0; u < U; ++u) gathered[u] = in[column[u]];
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (base + u * L + lane < end) sum = fma(TV(entry[u]), gathered[u], sum);
    }
    return groupSum<L>(sum);
}

// ---------------------------------------------------------------------------
// Row walkers. run() visits every row of the block's share of the matrix and
// calls epilogue(row, sum) once per row, from one lane, where sum = A[row,:] in.
// Kernels keep per-thread partials (dot products) inside their epilogues.
// ---------------------------------------------------------------------------

// The sparsity pattern as the walkers read it. param is the block's share:
// rows per block (VectorRows), 64-row tiles per block (StagedRows) or block
// tiles per block (BlockTiles, ColumnBlockTiles). tiles/tileCount describe the
// block tiles (tileCount + 1 entries {first row, first entry}, or {first row,
// first column block}); unused otherwise. For ColumnBlockTiles the extents,
// columns and narrow plane are those of the matrix's column-block form.
struct RowSource {
    int rows;
    int param;
    const int32_t* __restrict__ rowExtents;
    const int32_t* __restrict__ columnIndices;
    const int2* __restrict__ tiles;
    int tileCount;
    int narrowRuns;                            // runs of the narrow column plane (0: none)
    const uint16_t* __restrict__ narrowColumns;  // column - narrowBases[entry >> kNarrowShift]
    const int32_t* __restrict__ narrowBases;     // smallest column of each run of kNarrowRun entries
    const void* blockValues64;                 // column-block walker: w values per block (FP64, FP32)
    const void* blockValues32;
    // Column-sorted block tiles: per tile the smallest block column, per block
    // (in the tile's column order) its column offset and tile position, and
    // its w values in that order.
    const int32_t* __restrict__ tileBases;
    const uint32_t* __restrict__ sortedKeys;
    const void* sortedValues64;
    const void* sortedValues32;
    int64_t sortedBlocks;  // blocks of the sorted values (the offset of their second plane, see sortedHead)
    // Paired block tiles (PairedBlockTiles): per tile its first wide and first
    // narrow block and the smallest column of each width's blocks; the count
    // of wide blocks, after which the narrow blocks' keys and values follow.
    const int2* __restrict__ pairedBegins;
    const int2* __restrict__ pairedBases;
    int64_t pairedWide;
};

// Sorted-tile keys: (column - tile base) << kSortedPositionBits | position of
// the block in the tile's column-block order. Tiles hold at most 4096 blocks,
// and a tile's block columns may span at most 2^20 - 1 columns.
constexpr int kSortedPositionBits = 12;
constexpr uint32_t kSortedPositionMask = (1u << kSortedPositionBits) - 1;
constexpr int64_t kSortedMaxSpan = (int64_t(1) << (32 - kSortedPositionBits)) - 1;

// Sorted block values are stored in two planes when a block's run is longer
// than one 16-byte load (FP64 1 x 3): the first sortedHead values of every
// block, then the remaining ones of every block. Each load instruction then
// reads whole cache lines of one plane and can be non-temporal.
__host__ __device__ constexpr int sortedHead(int width, int valueBytes) {
    return width * valueBytes <= 16 ? width : 16 / valueBytes;
}

// L lanes per row, BS / L rows per block step; rows split into one contiguous
// chunk per block. Sum order: per-lane strided FMAs, then an xor butterfly.
template <class TV, class TM, int L, int U, int BS>
struct VectorRows {
    using Value = TV;
    using Matrix = TM;
    static constexpr int kBlock = BS;

    // Index of this block's partial in a grid reduction.
    __device__ __forceinline__ static int slot() { return xcdChunk(); }

    template <class Epilogue>
    __device__ __forceinline__ static void run(const RowSource& source, const TM* __restrict__ entries,
                                               const TV* __restrict__ in, Epilogue&& epilogue) {
        const int32_t* __restrict__ rowExtents = source.rowExtents;
        const int32_t* __restrict__ columnIndices = source.columnIndices;
        const int chunk = xcdChunk();
        const int rowBegin = int(min<int64_t>(source.rows, int64_t(chunk) * source.param));
        const int rowEnd = int(min<int64_t>(source.rows, int64_t(rowBegin) + source.param));
        const int group = threadIdx.x / L;
        const int lane = threadIdx.x % L;
        for (int row0 = rowBegin; row0 < rowEnd; row0 += BS / L) {
            const int row = row0 + group;
            if (row < rowEnd) {
                const TV sum = rowProduct<TV, TM, L, U>(row, lane, rowExtents, columnIndices, entries, in);
                if (lane == 0) epilogue(row, sum);
            }
        }
    }
};

// Walkers over block tiles of whole consecutive rows (BlockTiles,
// ColumnBlockTiles, SortedBlockTiles) expose tile(t, ...) and kCapacity.
template <class W, class = void>
struct IsTileWalker : std::false_type {};
template <class W>
struct IsTileWalker<W, std::void_t<decltype(W::kTiles)>> : std::bool_constant<W::kTiles> {};

// Orders a wavefront's LDS writes before its later LDS reads of other lanes' data.
__device__ __forceinline__ void stageFence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// One wavefront per tile of 64 consecutive rows. The tile's entries are
// streamed in coalesced passes (lane l loads entry base + l, U loads in
// flight per lane), each product a_ij * x_j is rounded and staged in LDS
// (CAP entries per segment), and lane r then adds the products of row r in
// CSR entry order. That is the summation of CSROperator.cpp:50-55
// (contribution += entry * in[column], no fused multiply-add), so every row
// comes out bit-identical to a sequential CPU loop over the same CSR.
// Sum of row first + lane (lane < count <= 64) of one tile, by one wavefront;
// `stage` holds CAP values private to the wavefront. Every lane must call it.
// The tile's entries are consumed in passes of 64 * U (CAP is a multiple of a
// pass, so passes never straddle a segment). Software pipelined: the column
// and entry loads of pass i + 1 are issued before the gathers of pass i are
// consumed and before the row sums of a finished segment run, so the matrix
// stream stays in flight while the wavefront does its LDS work.
template <class TV, class TM, int CAP, int U>
__device__ __forceinline__ TV stagedTileSum(int first, int count, int lane, const int32_t* __restrict__ rowExtents,
                                            const int32_t* __restrict__ columnIndices,
                                            const TM* __restrict__ entries, const TV* __restrict__ in, TV* stage) {
    constexpr int kPass = kWave * U;
    static_assert(CAP % kPass == 0, "a segment holds whole passes");
    const bool active = lane < count;
    const int rowBegin = active ? rowExtents[first + lane] : 0;
    const int rowEnd = active ? rowExtents[first + lane + 1] : 0;
    const int begin = __shfl(rowBegin, 0);
    const int end = __shfl(rowEnd, count - 1);
    int column[U], nextColumn[U];
    TM entry[U], nextEntry[U];
    auto load = [&](int base, int* c, TM* e) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = base + u * kWave + lane;
            if (k < end) {
                c[u] = CIEHIP_STREAM_LOAD(columnIndices + k);
                e[u] = CIEHIP_STREAM_LOAD(entries + k);
            } else {
                c[u] = 0;  // the tile is non-empty, so column 0 exists
                e[u] = TM(0);
            }
        }
    };
    TV sum = TV(0);
    if (begin < end) load(begin, column, entry);
    int base = begin;
    for (int segment = begin; segment < end; segment += CAP) {
        const int segmentEnd = min(segment + CAP, end);
        for (; base < segmentEnd; base += kPass) {
            TV gathered[U];
#pragma unroll
            for (int u = 0; u < U; ++u) gathered[u] = in[column[u]];
            const bool more = base + kPass < end;
            if (more) load(base + kPass, nextColumn, nextEntry);
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = base + u * kWave + lane;
                if (k < segmentEnd) stage[k - segment] = TV(entry[u]) * gathered[u];
            }
            if (more) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    column[u] = nextColumn[u];
                    entry[u] = nextEntry[u];
                }
            }
        }
        stageFence();
        const int lo = max(rowBegin, segment);
        const int hi = min(rowEnd, segmentEnd);
        for (int k = lo; k < hi; ++k) sum += stage[k - segment];
        stageFence();
    }
    return sum;
}

template <class TV, class TM, int CAP, int U, int BS>
struct StagedRows {
    using Value = TV;
    using Matrix = TM;
    static constexpr int kBlock = BS;
    static constexpr int kTile = kWave;

    // Index of this block's partial in a grid reduction.
    __device__ __forceinline__ static int slot() { return xcdChunk(); }

    template <class Epilogue>
    __device__ __forceinline__ static void run(const RowSource& source, const TM* __restrict__ entries,
                                               const TV* __restrict__ in, Epilogue&& epilogue) {
        __shared__ TV stage[BS / kWave][CAP];
        const int rows = source.rows;
        const int32_t* __restrict__ rowExtents = source.rowExtents;
        const int32_t* __restrict__ columnIndices = source.columnIndices;
        const int tiles = (rows + kTile - 1) / kTile;
        const int chunk = xcdChunk();
        const int tileBegin = int(min<int64_t>(tiles, int64_t(chunk) * source.param));
        const int tileEnd = int(min<int64_t>(tiles, int64_t(tileBegin) + source.param));
        const int wave = threadIdx.x / kWave;
        const int lane = threadIdx.x % kWave;
        for (int tile = tileBegin + wave; tile < tileEnd; tile += BS / kWave) {
            const int first = tile * kTile;
            const int count = min(kTile, rows - first);
            const TV sum = stagedTileSum<TV, TM, CAP, U>(first, count, lane, rowExtents, columnIndices, entries, in,
                                                          stage[wave]);
            if (lane < count) epilogue(first + lane, sum);
        }
    }
};

// Block tiles (the CSR-stream scheme of Greathouse and Daga, SC14). The rows
// are cut once, on the host, into tiles of whole consecutive rows with at most
// BS * U entries and at most BS rows; tiles[t] = {first row, first entry}.
// Block t handles tile t (one block per tile, so the hardware balances tiles
// over the compute units as blocks retire). It streams the tile's columns and
// entries with every lane (entry first + lane + BS * u, all U loads in flight
// before any is used), gathers in[column], stages the rounded products
// a_ij * x_j in LDS, and then sums each row with a group of G lanes, G the
// largest power of two (<= 64) that gives every row of the tile its own group:
// lane g of a group adds products g, g + G, ... in order, then an xor butterfly
// combines the group. The order depends only on the tile cut, so results are
// bitwise reproducible. A row longer than a tile is a tile of its own, summed
// in passes of BS * U. With fewer blocks than tiles (a launch sized to one
// resident round), block b handles tiles b, b + grid, ... in turn.
template <class TV, class TM, int U, int BS>
struct BlockTiles {
    static constexpr bool kTiles = true;  // walks tiles of whole rows: tile(t, ...) is public
    using Value = TV;
    using Matrix = TM;
    static constexpr int kBlock = BS;
    static constexpr int kCapacity = BS * U;
    static_assert(kCapacity <= kNarrowRun, "a tile must touch at most two narrow runs");

    // Index of this block's partial in a grid reduction.
    __device__ __forceinline__ static int slot() { return blockIdx.x; }

    template <class Epilogue>
    __device__ __forceinline__ static void run(const RowSource& source, const TM* __restrict__ entries,
                                               const TV* __restrict__ in, Epilogue&& epilogue) {
        __shared__ TV stage[kCapacity];
        __shared__ int32_t extents[BS + 1];
        for (int t = blockIdx.x; t < source.tileCount; t += gridDim.x) {
            tile(t, source, entries, in, epilogue, stage, extents);
            if (t + int(gridDim.x) < source.tileCount) __syncthreads();  // stage and extents are reused
        }
    }

    template <class Epilogue>
    __device__ __forceinline__ static void tile(int t, const RowSource& source, const TM* __restrict__ entries,
                                                const TV* __restrict__ in, Epilogue& epilogue, TV* stage,
                                                int32_t* extents) {
        const int32_t* __restrict__ columnIndices = source.columnIndices;
        const int tid = threadIdx.x;
        const int2 head = source.tiles[t];
        const int2 next = source.tiles[t + 1];
        const int first = head.x, rows = next.x - head.x, begin = head.y, end = next.y;
        if (end - begin > kCapacity) {
            const TV sum = longRow(begin, end, columnIndices, entries, in, stage);
            if (tid == 0) epilogue(first, sum);
            return;
        }
        // Every load is issued unconditionally (indices clamped into the tile,
        // so lanes past its end re-read its last entry) and masked afterwards:
        // a load consumed inside a per-entry branch would make the compiler wait
        // for it there, one entry at a time.
        int column[U];
        TM entry[U];
        const int last = max(begin, end - 1);  // an empty tile loads nothing below
        if (end > begin && source.narrowRuns) {
            const int run = begin >> kNarrowShift;
            const int base0 = source.narrowBases[run];
            const int base1 = source.narrowBases[min(run + 1, source.narrowRuns - 1)];
            uint32_t offset[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = min(begin + tid + BS * u, last);
                offset[u] = CIEHIP_STREAM_LOAD(source.narrowColumns + k);
                entry[u] = CIEHIP_STREAM_LOAD(entries + k);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = begin + tid + BS * u;
                const bool live = k < end;
                column[u] = live ? ((k >> kNarrowShift) == run ? base0 : base1) + int(offset[u]) : base0;
                entry[u] = live ? entry[u] : TM(0);
            }
        } else if (end > begin) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = min(begin + tid + BS * u, last);
                column[u] = CIEHIP_STREAM_LOAD(columnIndices + k);
                entry[u] = CIEHIP_STREAM_LOAD(entries + k);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool live = begin + tid + BS * u < end;
                entry[u] = live ? entry[u] : TM(0);  // the re-read column is a real one: harmless to gather
            }
        } else {
            // Only empty rows: each sums to +0, as below, and nothing is
            // gathered (a matrix without columns has no in[0]).
            if (tid < rows) epilogue(first + tid, TV(0));
            return;
        }
        if (tid < rows) extents[tid] = source.rowExtents[first + tid] - begin;
        if (tid == 0) extents[rows] = end - begin;
        TV gathered[U];
#pragma unroll
        for (int u = 0; u < U; ++u) gathered[u] = in[column[u]];
#pragma unroll
        for (int u = 0; u < U; ++u) stage[tid + BS * u] = TV(entry[u]) * gathered[u];
        __syncthreads();
        int group = kWave;
        while (group > 1 && rows * group > BS) group >>= 1;
        const int row = tid / group;
        const int g = tid % group;
        TV sum = TV(0);
        if (row < rows) {
            const int rowEnd = extents[row + 1];
            for (int k = extents[row] + g; k < rowEnd; k += group) sum += stage[k];
        }
        for (int offset = group / 2; offset > 0; offset /= 2) sum += __shfl_xor(sum, offset, kWave);
        if (row < rows && g == 0) epilogue(first + row, sum);
    }

    // Sum of one row of more than kCapacity entries, valid in thread 0: per
    // pass, lane sums over u, wavefront butterflies, then waves in order.
    __device__ __forceinline__ static TV longRow(int begin, int end, const int32_t* __restrict__ columnIndices,
                                                 const TM* __restrict__ entries, const TV* __restrict__ in,
                                                 TV* stage) {
        const int tid = threadIdx.x;
        TV total = TV(0);
        for (int base = begin; base < end; base += kCapacity) {
            int column[U];
            TM entry[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {  // clamped, unconditional loads (see tile())
                const int k = min(base + tid + BS * u, end - 1);
                column[u] = CIEHIP_STREAM_LOAD(columnIndices + k);
                entry[u] = CIEHIP_STREAM_LOAD(entries + k);
            }
            TV gathered[U];
#pragma unroll
            for (int u = 0; u < U; ++u) gathered[u] = in[column[u]];
            TV part = TV(0);
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (base + tid + BS * u < end) part += TV(entry[u]) * gathered[u];
            part = groupSum<kWave>(part);
            if ((tid & (kWave - 1)) == 0) stage[tid / kWave] = part;
            __syncthreads();
            if (tid == 0) {
#pragma unroll
                for (int w = 0; w < BS / kWave; ++w) total += stage[w];
            }
            __syncthreads();
        }
        return total;
    }
};

// ---------------------------------------------------------------------------
// Column-block tiles: the block-tile walker over the matrix's column-block
// form, 1 x W blocks of adjacent columns (W values each, zero where the row
// has no entry). A lane's unit is a block: its W values arrive in one load
// and its W x values in one gather, so a row whose entries come in runs of
// adjacent columns (FEM rows in node order) costs about a third of the
// gathers at W = 3; the gather instructions, not the bytes, bound the tiled
// product in FP32. A block's products are added in column order, a zero value
// adding nothing (so padding never turns an infinite x into NaN); then each
// row sums its blocks as BlockTiles sums entries. The `entries` argument (the
// CSR plane) is not read: the block values come with the row source.
// ---------------------------------------------------------------------------

template <class T, int W>
struct ValueRun {  // W consecutive values; loads of it need only T's alignment
    T v[W];
};

template <class TV, class TM, int U, int BS, int W>
struct ColumnBlockTiles {
    static constexpr bool kTiles = true;  // walks tiles of whole rows: tile(t, ...) is public
    using Value = TV;
    using Matrix = TM;
    static constexpr int kBlock = BS;
    static constexpr int kCapacity = BS * U;  // blocks per tile
    static_assert(kCapacity <= kNarrowRun, "a tile must touch at most two narrow runs");

    __device__ __forceinline__ static int slot() { return blockIdx.x; }

    __device__ __forceinline__ static const TM* values(const RowSource& source) {
        if constexpr (sizeof(TM) == sizeof(double)) return static_cast<const TM*>(source.blockValues64);
        else return static_cast<const TM*>(source.blockValues32);
    }

    // A block's values are read once per product. When one load instruction
    // fetches them (at most 16 bytes) it is marked non-temporal, so the
    // stream does not displace the gathered vector from the caches (FP32
    // 1 x 3 blocks on 3D Q2 64^3: 143 -> 130 us per product). A larger run
    // takes two instructions over shared cache lines, and the second one
    // misses when the first did not keep them (FP64 1 x 3 measured slower).
    __device__ __forceinline__ static ValueRun<TM, W> loadBlock(const TM* __restrict__ values, int k) {
        ValueRun<TM, W> run;
        if constexpr (sizeof(run) <= 16) {
#pragma unroll
            for (int w = 0; w < W; ++w) run.v[w] = __builtin_nontemporal_load(values + int64_t(k) * W + w);
        } else {
            __builtin_memcpy(&run, values + int64_t(k) * W, sizeof(run));
        }
        return run;
    }

    __device__ __forceinline__ static ValueRun<TV, W> gather(const TV* __restrict__ in, int column) {
        ValueRun<TV, W> run;
        __builtin_memcpy(&run, in + column, sizeof(run));
        return run;
    }

    __device__ __forceinline__ static TV blockProduct(const ValueRun<TM, W>& a, const ValueRun<TV, W>& x) {
        TV sum = TV(0);
#pragma unroll
        for (int i = 0; i < W; ++i)
            if (a.v[i] != TM(0)) sum += TV(a.v[i]) * x.v[i];
        return sum;
    }

    template <class Epilogue>
    __device__ __forceinline__ static void run(const RowSource& source, const TM* __restrict__,
                                               const TV* __restrict__ in, Epilogue&& epilogue) {
        __shared__ TV stage[kCapacity];
        __shared__ int32_t extents[BS + 1];
        for (int t = blockIdx.x; t < source.tileCount; t += gridDim.x) {
            tile(t, source, in, epilogue, stage, extents);
            if (t + int(gridDim.x) < source.tileCount) __syncthreads();  // stage and extents are reused
        }
    }

    // BlockTiles::tile's signature (the CSR plane argument is not read).
    template <class Epilogue>
    __device__ __forceinline__ static void tile(int t, const RowSource& source, const TM* __restrict__,
                                                const TV* __restrict__ in, Epilogue& epilogue, TV* stage,
                                                int32_t* extents) {
        tile(t, source, in, epilogue, stage, extents);
    }

    template <class Epilogue>
    __device__ __forceinline__ static void tile(int t, const RowSource& source, const TV* __restrict__ in,
                                                Epilogue& epilogue, TV* stage, int32_t* extents) {
        const TM* __restrict__ blockValues = values(source);
        const int tid = threadIdx.x;
        const int2 head = source.tiles[t];
        const int2 next = source.tiles[t + 1];
        const int first = head.x, rows = next.x - head.x, begin = head.y, end = next.y;
        if (end - begin > kCapacity) {
            const TV sum = longRow(begin, end, source.columnIndices, blockValues, in, stage);
            if (tid == 0) epilogue(first, sum);
            return;
        }
        if (end == begin) {  // only empty rows: nothing to gather
            if (tid < rows) epilogue(first + tid, TV(0));
            return;
        }
        // Unconditional loads, indices clamped into the tile, masked afterwards (see BlockTiles).
        int column[U];
        ValueRun<TM, W> block[U];
        const int last = end - 1;
        if (source.narrowRuns) {
            const int run = begin >> kNarrowShift;
            const int base0 = source.narrowBases[run];
            const int base1 = source.narrowBases[min(run + 1, source.narrowRuns - 1)];
            uint32_t offset[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = min(begin + tid + BS * u, last);
                offset[u] = CIEHIP_STREAM_LOAD(source.narrowColumns + k);
                block[u] = loadBlock(blockValues, k);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = begin + tid + BS * u;
                column[u] = ((min(k, last) >> kNarrowShift) == run ? base0 : base1) + int(offset[u]);
            }
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = min(begin + tid + BS * u, last);
                column[u] = CIEHIP_STREAM_LOAD(source.columnIndices + k);
                block[u] = loadBlock(blockValues, k);
            }
        }
        if (tid < rows) extents[tid] = source.rowExtents[first + tid] - begin;
        if (tid == 0) extents[rows] = end - begin;
        ValueRun<TV, W> gathered[U];
#pragma unroll
        for (int u = 0; u < U; ++u) gathered[u] = gather(in, column[u]);  // clamped lanes re-read a real block
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (begin + tid + BS * u < end) stage[tid + BS * u] = blockProduct(block[u], gathered[u]);
        __syncthreads();
        int group = kWave;
        while (group > 1 && rows * group > BS) group >>= 1;
        const int row = tid / group;
        const int g = tid % group;
        TV sum = TV(0);
        if (row < rows) {
            const int rowEnd = extents[row + 1];
            for (int k = extents[row] + g; k < rowEnd; k += group) sum += stage[k];
        }
        for (int offset = group / 2; offset > 0; offset /= 2) sum += __shfl_xor(sum, offset, kWave);
        if (row < rows && g == 0) epilogue(first + row, sum);
    }

    // Sum of one row of more than kCapacity blocks, valid in thread 0 (as BlockTiles::longRow).
    __device__ __forceinline__ static TV longRow(int begin, int end, const int32_t* __restrict__ columnIndices,
                                                 const TM* __restrict__ blockValues, const TV* __restrict__ in,
                                                 TV* stage) {
        const int tid = threadIdx.x;
        TV total = TV(0);
        for (int base = begin; base < end; base += kCapacity) {
            int column[U];
            ValueRun<TM, W> block[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = min(base + tid + BS * u, end - 1);
                column[u] = CIEHIP_STREAM_LOAD(columnIndices + k);
                block[u] = loadBlock(blockValues, k);
            }
            ValueRun<TV, W> gathered[U];
#pragma unroll
            for (int u = 0; u < U; ++u) gathered[u] = gather(in, column[u]);
            TV part = TV(0);
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (base + tid + BS * u < end) part += blockProduct(block[u], gathered[u]);
            part = groupSum<kWave>(part);
            if ((tid & (kWave - 1)) == 0) stage[tid / kWave] = part;
            __syncthreads();
            if (tid == 0) {
#pragma unroll
                for (int w = 0; w < BS / kWave; ++w) total += stage[w];
            }
            __syncthreads();
        }
        return total;
    }
};

// ---------------------------------------------------------------------------
// Column-sorted block tiles: the column-block walker with each tile's blocks
// stored in the order of their first column (ties in block order) and a key
// per block holding its column offset from the tile's smallest column and its
// position among the tile's blocks. The lanes of a wavefront then gather
// adjacent x values (the blocks of a tile's rows that reach one node plane lie
// together) instead of one run per row across all the planes a row couples
// to, so each gather instruction touches a few cache lines instead of tens.
// Each product is staged in LDS at the block's position, so the row sums, and
// every result bit, are those of ColumnBlockTiles over the same tiles. A row
// longer than a tile is walked as ColumnBlockTiles walks it (its blocks are
// already in column order).
// ---------------------------------------------------------------------------

template <class TV, class TM, int U, int BS, int W>
struct SortedBlockTiles {
    static constexpr bool kTiles = true;  // walks tiles of whole rows: tile(t, ...) is public
    using Value = TV;
    using Matrix = TM;
    using Plain = ColumnBlockTiles<TV, TM, U, BS, W>;
    static constexpr int kBlock = BS;
    static constexpr int kWidth = W;
    static constexpr int kCapacity = BS * U;  // blocks per tile
    static_assert(kCapacity <= (1 << kSortedPositionBits), "a tile position must fit the key");

    __device__ __forceinline__ static int slot() { return blockIdx.x; }

    __device__ __forceinline__ static const TM* sortedValues(const RowSource& source) {
        if constexpr (sizeof(TM) == sizeof(double)) return static_cast<const TM*>(source.sortedValues64);
        else return static_cast<const TM*>(source.sortedValues32);
    }

    // Block k's values: kHead of them in the first plane, the rest in `tail`;
    // non-temporal loads of at most 16 bytes each (see ColumnBlockTiles::loadBlock).
    static constexpr int kHead = sortedHead(W, int(sizeof(TM)));
    __device__ __forceinline__ static ValueRun<TM, W> loadSorted(const TM* __restrict__ values,
                                                                 const TM* __restrict__ tail, int k) {
        ValueRun<TM, W> run;
#pragma unroll
        for (int w = 0; w < kHead; ++w) run.v[w] = __builtin_nontemporal_load(values + int64_t(k) * kHead + w);
#pragma unroll
        for (int w = kHead; w < W; ++w)
            run.v[w] = __builtin_nontemporal_load(tail + int64_t(k) * (W - kHead) + (w - kHead));
        return run;
    }

    template <class Epilogue>
    __device__ __forceinline__ static void run(const RowSource& source, const TM* __restrict__,
                                               const TV* __restrict__ in, Epilogue&& epilogue) {
        __shared__ TV stage[kCapacity];
        __shared__ int32_t extents[BS + 1];
        for (int t = blockIdx.x; t < source.tileCount; t += gridDim.x) {
            tile(t, source, in, epilogue, stage, extents);
            if (t + int(gridDim.x) < source.tileCount) __syncthreads();  // stage and extents are reused
        }
    }

    // BlockTiles::tile's signature (the CSR plane argument is not read).
    template <class Epilogue>
    __device__ __forceinline__ static void tile(int t, const RowSource& source, const TM* __restrict__,
                                                const TV* __restrict__ in, Epilogue& epilogue, TV* stage,
                                                int32_t* extents) {
        tile(t, source, in, epilogue, stage, extents);
    }

    template <class Epilogue>
    __device__ __forceinline__ static void tile(int t, const RowSource& source, const TV* __restrict__ in,
                                                Epilogue& epilogue, TV* stage, int32_t* extents) {
        const int tid = threadIdx.x;
        const int2 head = source.tiles[t];
        const int2 next = source.tiles[t + 1];
        const int first = head.x, rows = next.x - head.x, begin = head.y, end = next.y;
        if (end - begin > kCapacity) {
            const TV sum = Plain::longRow(begin, end, source.columnIndices, Plain::values(source), in, stage);
            if (tid == 0) epilogue(first, sum);
            return;
        }
        if (end == begin) {  // only empty rows: nothing to gather
            if (tid < rows) epilogue(first + tid, TV(0));
            return;
        }
        const TM* __restrict__ values = sortedValues(source);
        const TM* __restrict__ tail = values + source.sortedBlocks * kHead;
        const int base = source.tileBases[t];
        const int last = end - 1;
        // Unconditional loads, indices clamped into the tile, masked afterwards (see BlockTiles).
        uint32_t key[U];
        ValueRun<TM, W> block[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = min(begin + tid + BS * u, last);
            key[u] = CIEHIP_STREAM_LOAD(source.sortedKeys + k);
            block[u] = loadSorted(values, tail, k);
        }
        if (tid < rows) extents[tid] = source.rowExtents[first + tid] - begin;
        if (tid == 0) extents[rows] = end - begin;
        ValueRun<TV, W> gathered[U];
#pragma unroll
        for (int u = 0; u < U; ++u) gathered[u] = Plain::gather(in, base + int(key[u] >> kSortedPositionBits));
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (begin + tid + BS * u < end)
                stage[key[u] & kSortedPositionMask] = Plain::blockProduct(block[u], gathered[u]);
        __syncthreads();
        int group = kWave;
        while (group > 1 && rows * group > BS) group >>= 1;
        const int row = tid / group;
        const int g = tid % group;
        TV sum = TV(0);
        if (row < rows) {
            const int rowEnd = extents[row + 1];
            for (int k = extents[row] + g; k < rowEnd; k += group) sum += stage[k];
        }
        for (int offset = group / 2; offset > 0; offset /= 2) sum += __shfl_xor(sum, offset, kWave);
        if (row < rows && g == 0) epilogue(first + row, sum);
    }
};

// ---------------------------------------------------------------------------
// Paired block tiles: rows covered by 1 x W1 ("wide") and 1 x W2 ("narrow")
// blocks of adjacent columns. A block is wide where a narrow one would leave
// entries of the wide span to a further block, so runs of W1 columns (and of
// W1 + W2, 2 W1, ...) are covered without padding: on a 3D Q2 stiffness
// matrix, whose rows hold runs of 5 and 3 columns, 1 x 5 and 1 x 3 blocks
// store 16.6% fewer FP32 bytes than 1 x 3 blocks alone (14% in FP64). Each
// tile of whole rows keeps its wide and its narrow blocks each in column
// order, with the key of the column-sorted tiles: column offset from the
// width's smallest column in the tile, and the block's position among all the
// tile's blocks in row order. A lane walks U wide and U narrow blocks; each
// product is staged at its position, and each row sums its staged products as
// the other tile walkers do. Values are stored per width in planes of at most
// 16 bytes per block (PlaneRun), so every value load is one non-temporal
// instruction. Tiles never hold a row longer than BS * U blocks of one width
// (the form refuses such matrices).
// ---------------------------------------------------------------------------

template <class T, int W>
struct PlaneRun {
    static constexpr int kChunk = 16 / int(sizeof(T));  // values per plane
    static constexpr int kPlanes = (W + kChunk - 1) / kChunk;
    // Values per block in plane p.
    __host__ __device__ static constexpr int length(int p) { return W - p * kChunk < kChunk ? W - p * kChunk : kChunk; }
    // Values of block k among `blocks` blocks: plane p holds values
    // [p * kChunk, p * kChunk + length(p)) of every block, contiguously, and
    // starts at blocks * p * kChunk.
    __device__ __forceinline__ static ValueRun<T, W> load(const T* __restrict__ values, int64_t blocks, int k) {
        ValueRun<T, W> run;
#pragma unroll
        for (int p = 0; p < kPlanes; ++p) {
            const T* __restrict__ plane = values + blocks * (int64_t(p) * kChunk) + int64_t(k) * length(p);
#pragma unroll
            for (int j = 0; j < length(p); ++j) run.v[p * kChunk + j] = __builtin_nontemporal_load(plane + j);
        }
        return run;
    }
};

template <class TV, class TM, int U, int BS, int W1, int W2>
struct PairedBlockTiles {
    static_assert(W1 > W2 && W2 >= 1, "the wide blocks are the wider ones");
    static constexpr bool kTiles = true;  // walks tiles of whole rows: tile(t, ...) is public
    using Value = TV;
    using Matrix = TM;
    using Wide = ColumnBlockTiles<TV, TM, U, BS, W1>;
    using Narrow = ColumnBlockTiles<TV, TM, U, BS, W2>;
    static constexpr int kBlock = BS;
    static constexpr int kPerWidth = BS * U;        // blocks of each width per tile
    static constexpr int kCapacity = 2 * kPerWidth;  // staged products per tile
    static_assert(kCapacity <= (1 << kSortedPositionBits), "a tile position must fit the key");

    __device__ __forceinline__ static int slot() { return blockIdx.x; }

    __device__ __forceinline__ static const TM* values(const RowSource& source) {
        if constexpr (sizeof(TM) == sizeof(double)) return static_cast<const TM*>(source.sortedValues64);
        else return static_cast<const TM*>(source.sortedValues32);
    }

    template <class Epilogue>
    __device__ __forceinline__ static void run(const RowSource& source, const TM* __restrict__,
                                               const TV* __restrict__ in, Epilogue&& epilogue) {
        __shared__ TV stage[kCapacity];
        __shared__ int32_t extents[BS + 1];
        for (int t = blockIdx.x; t < source.tileCount; t += gridDim.x) {
            tile(t, source, in, epilogue, stage, extents);
            if (t + int(gridDim.x) < source.tileCount) __syncthreads();  // stage and extents are reused
        }
    }

    // BlockTiles::tile's signature (the CSR plane argument is not read).
    template <class Epilogue>
    __device__ __forceinline__ static void tile(int t, const RowSource& source, const TM* __restrict__,
                                                const TV* __restrict__ in, Epilogue& epilogue, TV* stage,
                                                int32_t* extents) {
        tile(t, source, in, epilogue, stage, extents);
    }

    template <class Epilogue>
    __device__ __forceinline__ static void tile(int t, const RowSource& source, const TV* __restrict__ in,
                                                Epilogue& epilogue, TV* stage, int32_t* extents) {
        const int tid = threadIdx.x;
        const int2 head = source.tiles[t];
        const int2 next = source.tiles[t + 1];
        const int first = head.x, rows = next.x - head.x, merged = head.y;
        const int2 begin = source.pairedBegins[t], end = source.pairedBegins[t + 1];
        const bool wide = end.x > begin.x, narrow = end.y > begin.y;  // the same for the whole block
        if (!wide && !narrow) {  // only empty rows: nothing to gather
            if (tid < rows) epilogue(first + tid, TV(0));
            return;
        }
        const int2 base = source.pairedBases[t];
        const int64_t wideBlocks = source.pairedWide;
        const TM* __restrict__ wideValues = values(source);
        const TM* __restrict__ narrowValues = wideValues + wideBlocks * W1;
        const uint32_t* __restrict__ wideKeys = source.sortedKeys;
        const uint32_t* __restrict__ narrowKeys = source.sortedKeys + wideBlocks;
        const int64_t narrowBlocks = source.sortedBlocks - wideBlocks;
        // Unconditional loads, indices clamped into the width's run of the
        // tile (skipped when the tile has none of that width), masked afterwards.
        uint32_t wideKey[U], narrowKey[U];
        ValueRun<TM, W1> wideRun[U];
        ValueRun<TM, W2> narrowRun[U];
        if (wide) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = min(begin.x + tid + BS * u, end.x - 1);
                wideKey[u] = CIEHIP_STREAM_LOAD(wideKeys + k);
                wideRun[u] = PlaneRun<TM, W1>::load(wideValues, wideBlocks, k);
            }
        }
        if (narrow) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int k = min(begin.y + tid + BS * u, end.y - 1);
                narrowKey[u] = CIEHIP_STREAM_LOAD(narrowKeys + k);
                narrowRun[u] = PlaneRun<TM, W2>::load(narrowValues, narrowBlocks, k);
            }
        }
        if (tid < rows) extents[tid] = source.rowExtents[first + tid] - merged;
        if (tid == 0) extents[rows] = next.y - merged;
        if (wide) {
            ValueRun<TV, W1> gathered[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                gathered[u] = Wide::gather(in, base.x + int(wideKey[u] >> kSortedPositionBits));
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (begin.x + tid + BS * u < end.x)
                    stage[wideKey[u] & kSortedPositionMask] = Wide::blockProduct(wideRun[u], gathered[u]);
        }
        if (narrow) {
            ValueRun<TV, W2> gathered[U];
#pragma unroll
            for (int u = 0; u < U; ++u)
                gathered[u] = Narrow::gather(in, base.y + int(narrowKey[u] >> kSortedPositionBits));
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (begin.y + tid + BS * u < end.y)
                    stage[narrowKey[u] & kSortedPositionMask] = Narrow::blockProduct(narrowRun[u], gathered[u]);
        }
        __syncthreads();
        int group = kWave;
        while (group > 1 && rows * group > BS) group >>= 1;
        const int row = tid / group;
        const int g = tid % group;
        TV sum = TV(0);
        if (row < rows) {
            const int rowEnd = extents[row + 1];
            for (int k = extents[row] + g; k < rowEnd; k += group) sum += stage[k];
        }
        for (int offset = group / 2; offset > 0; offset /= 2) sum += __shfl_xor(sum, offset, kWave);
        if (row < rows && g == 0) epilogue(first + row, sum);
    }
};

// out = inScale * out + outScale * A * in. With contraction off this equals
// each special case of CSROperator.cpp:78-110 bit for bit (1 * y, -1 * y and
// 0 + x are exact), except inScale = outScale = -1, where the reference
// computes A*in - out and this follows the documented contract, -A*in - out.
template <class Walker>
__global__ __launch_bounds__(Walker::kBlock) void csrProductKernel(
    RowSource source, const typename Walker::Matrix* __restrict__ entries, const typename Walker::Value* __restrict__ in,
    typename Walker::Value* __restrict__ out, typename Walker::Value inScale, typename Walker::Value outScale) {
    using TV = typename Walker::Value;
    Walker::run(source, entries, in, [&](int row, TV sum) {
        out[row] = inScale == TV(0) ? outScale * sum : inScale * out[row] + outScale * sum;
    });
}

// ---------------------------------------------------------------------------
// Elementwise kernels of the LinalgSpaceLike operations
// ---------------------------------------------------------------------------

enum VectorOp : int { kScale = 0, kScaleUniform = 1, kAdd = 2, kAssign = 3, kFill = 4 };

template <class T>
__global__ __launch_bounds__(kReduceBlock) void vectorKernel(int op, int64_t n, T* __restrict__ target,
                                                             const T* __restrict__ source, T value) {
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        switch (op) {
        case kScale:  // DefaultSpace.cpp:100-105: t = value * t * s
            target[i] = value == T(1) ? target[i] * source[i]
                      : value == T(-1) ? -target[i] * source[i]
                      : value * target[i] * source[i];
            break;
        case kScaleUniform: target[i] *= value; break;
        case kAdd:  // DefaultSpace.cpp:145-150: t = t + value * s
            target[i] = value == T(1) ? target[i] + source[i]
                      : value == T(-1) ? target[i] - source[i]
                      : target[i] + value * source[i];
            break;
        case kAssign: target[i] = source[i]; break;
        case kFill: target[i] = value; break;
        }
    }
}

// DiagonalOperator::product: out = inScale * out + outScale * inverse .* in
template <class T>
__global__ __launch_bounds__(kReduceBlock) void diagonalProductKernel(int64_t n, const T* __restrict__ inverse,
                                                                      T inScale, const T* __restrict__ in,
                                                                      T outScale, T* __restrict__ out) {
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const T scaled = outScale == T(1) ? in[i] * inverse[i] : outScale * in[i] * inverse[i];
        out[i] = inScale == T(0) ? scaled : (inScale == T(1) ? out[i] : inScale * out[i]) + scaled;
    }
}

// Two-stage deterministic inner product: per-block partials in a fixed
// order, then innerProductFinishKernel (one block) sums them in block order.
template <class T>
__global__ __launch_bounds__(kReduceBlock) void innerProductKernel(int64_t n, const T* __restrict__ left,
                                                                   const T* __restrict__ right, T* partials) {
    __shared__ T lds[kReduceBlock / kWave];
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    T sum = T(0);
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) sum += left[i] * right[i];
    sum = blockSum<kReduceBlock>(sum, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = sum;
}

template <class T>
__global__ __launch_bounds__(kReduceBlock) void innerProductFinishKernel(int count, const T* __restrict__ partials,
                                                                         T* result) {
    __shared__ T lds[kReduceBlock / kWave];
    const T total = sumPartials<kReduceBlock>(partials, count, 1, lds);
    if (threadIdx.x == 0) *result = total;
}

// Jacobi set-up (DiagonalOperator.cpp:37-88), a group of G lanes per row: the
// group reads the row's entries G at a time, coalesced, and finds the
// diagonal among them (columns are unique per row). inverse[row] = 1 / a_ii;
// the first row without a diagonal entry or with a zero there is recorded in
// failure[1] / failure[0] (atomicMin of the row; no value is accumulated).
// With `correction`, also the row-sum correction of the FP32 plane:
// correction[row] = fl32(sum_j (a_ij - fl32(a_ij))), each difference exact in
// FP64 (Sterbenz), summed in a fixed order. The FP32 plane plus this term on
// the diagonal keeps the row sums of the FP64 matrix, so the smooth vectors
// that FEM stiffness rows annihilate are annihilated by the FP32 operator too.
template <class T, int G>
__global__ __launch_bounds__(kReduceBlock) void diagonalKernel(int rows, const int32_t* __restrict__ rowExtents,
                                                               const int32_t* __restrict__ columnIndices,
                                                               const double* __restrict__ entries,
                                                               T* __restrict__ inverse, float* __restrict__ correction,
                                                               unsigned long long* failure) {
    static_assert(G >= 1 && G <= kWave && (G & (G - 1)) == 0, "lane groups are powers of two within a wave");
    const int lane = int(threadIdx.x) & (G - 1);
    const int groupsPerBlock = kReduceBlock / G;
    const int stride = int(gridDim.x) * groupsPerBlock;
    for (int row = int(blockIdx.x) * groupsPerBlock + int(threadIdx.x) / G; row < rows; row += stride) {
        const int begin = rowExtents[row], end = rowExtents[row + 1];
        double diagonal = 0.0, residue = 0.0;
        int found = 0;
        for (int k = begin + lane; k < end; k += G) {
            const double a = entries[k];
            if (columnIndices[k] == row) {
                diagonal = a;
                found = 1;
            }
            residue += a - double(float(a));
        }
        found = groupSum<G>(found);          // 0 or 1: columns are unique per row
        diagonal = groupSum<G>(diagonal);    // one lane holds it, the others add zeros
        residue = groupSum<G>(residue);
        if (lane == 0) {
            if (!found) atomicMin(failure + 1, (unsigned long long)row);
            else if (diagonal == 0.0) atomicMin(failure + 0, (unsigned long long)row);
            inverse[row] = found && diagonal != 0.0 ? T(1.0 / diagonal) : T(0);
            if (correction) correction[row] = float(residue);
        }
    }
}

template <class TS, class TD>
__global__ __launch_bounds__(kReduceBlock) void convertKernel(int64_t n, const TS* __restrict__ source,
                                                              TD* __restrict__ target) {
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) target[i] = TD(source[i]);
}

}  // namespace ciehip

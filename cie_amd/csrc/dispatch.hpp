// Compile-time dispatch over CSR launch shapes and launch-size helpers.
#pragma once

#include "common.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace ciehip {

constexpr int kSpmvBlock = 256;
constexpr int kStagedMaxRow = 8192;  // longest row the autotuner tries the staged kernel for

// Lanes per row x entries per lane in flight.
#define CIEHIP_SHAPES(X)                                              \
    X(1, 1) X(1, 2) X(1, 4) X(1, 8)                                   \
    X(2, 1) X(2, 2) X(2, 4) X(2, 8)                                   \
    X(4, 1) X(4, 2) X(4, 4) X(4, 8)                                   \
    X(8, 1) X(8, 2) X(8, 4) X(8, 8)                                   \
    X(16, 1) X(16, 2) X(16, 4) X(16, 8)                               \
    X(32, 1) X(32, 2) X(32, 4) X(32, 8)                               \
    X(64, 1) X(64, 2) X(64, 4) X(64, 8)

// Calls f(integral_constant<L>, integral_constant<U>); false if the shape is unsupported.
template <class F>
bool withShape(int lanes, int unroll, F&& f) {
    switch (lanes * 16 + unroll) {
#define CIEHIP_SHAPE_CASE(L, U)                                                        \
    case L * 16 + U:                                                                   \
        f(std::integral_constant<int, L>{}, std::integral_constant<int, U>{});        \
        return true;
        CIEHIP_SHAPES(CIEHIP_SHAPE_CASE)
#undef CIEHIP_SHAPE_CASE
    default: return false;
    }
}

inline bool validShape(int lanes, int unroll) {
    return withShape(lanes, unroll, [](auto, auto) {});
}

#define CIEHIP_STAGES(X) X(512, 4) X(512, 8) X(1024, 4) X(1024, 8)

inline bool validStage(int capacity, int unroll) {
    switch (capacity * 16 + unroll) {
#define CIEHIP_STAGE_VALID(C, U) case C * 16 + U: return true;
        CIEHIP_STAGES(CIEHIP_STAGE_VALID)
#undef CIEHIP_STAGE_VALID
    default: return false;
    }
}

// Entries per lane in flight of the block-tile kernel (tiles of 256 * U entries).
#define CIEHIP_TILES(X) X(4) X(8) X(16)

inline bool validTiles(int block, int unroll) {
    if (block != kSpmvBlock) return false;
    switch (unroll) {
#define CIEHIP_TILES_VALID(U) case U: return true;
        CIEHIP_TILES(CIEHIP_TILES_VALID)
#undef CIEHIP_TILES_VALID
    default: return false;
    }
}

// Columns per block x blocks per lane in flight of the column-block kernel.
#define CIEHIP_BLOCKS(X) X(2, 4) X(2, 8) X(2, 16) X(3, 4) X(3, 8) X(3, 16) X(4, 4) X(4, 8) X(4, 16)

inline bool validBlocks(int width, int unroll) {
    switch (width * 32 + unroll) {
#define CIEHIP_BLOCKS_VALID(W, U) case W * 32 + U: return true;
        CIEHIP_BLOCKS(CIEHIP_BLOCKS_VALID)
#undef CIEHIP_BLOCKS_VALID
    default: return false;
    }
}

// Wide width, narrow width and blocks per lane and width of the paired-tile
// kernel; its launch width is 10 * wide + narrow.
#define CIEHIP_PAIRED(X) X(5, 3, 2) X(5, 3, 4) X(3, 2, 2) X(3, 2, 4)

inline bool validPaired(int width, int unroll) {
    switch (width * 32 + unroll) {
#define CIEHIP_PAIRED_VALID(W1, W2, U) case (10 * W1 + W2) * 32 + U: return true;
        CIEHIP_PAIRED(CIEHIP_PAIRED_VALID)
#undef CIEHIP_PAIRED_VALID
    default: return false;
    }
}

inline int pairedWide(int width) { return width / 10; }
inline int pairedNarrow(int width) { return width % 10; }

// Calls f(Walker{}) with the row walker the launch selects; false if unsupported.
template <class TV, class TM, class F>
bool withWalker(const CsrLaunch& launch, F&& f) {
    if (launch.kernel == CIE_HIP_SPMV_PAIRED) {
        switch (launch.width * 32 + launch.unroll) {
#define CIEHIP_PAIRED_CASE(W1, W2, U)                                  \
    case (10 * W1 + W2) * 32 + U:                                      \
        f(PairedBlockTiles<TV, TM, U, kSpmvBlock, W1, W2>{});          \
        return true;
            CIEHIP_PAIRED(CIEHIP_PAIRED_CASE)
#undef CIEHIP_PAIRED_CASE
        default: return false;
        }
    }
    if (launch.kernel == CIE_HIP_SPMV_SORTED) {
        switch (launch.width * 32 + launch.unroll) {
#define CIEHIP_SORTED_CASE(W, U)                                  \
    case W * 32 + U:                                              \
        f(SortedBlockTiles<TV, TM, U, kSpmvBlock, W>{});          \
        return true;
            CIEHIP_BLOCKS(CIEHIP_SORTED_CASE)
#undef CIEHIP_SORTED_CASE
        default: return false;
        }
    }
    if (launch.kernel == CIE_HIP_SPMV_BLOCKS) {
        switch (launch.width * 32 + launch.unroll) {
#define CIEHIP_BLOCKS_CASE(W, U)                                  \
    case W * 32 + U:                                              \
        f(ColumnBlockTiles<TV, TM, U, kSpmvBlock, W>{});          \
        return true;
            CIEHIP_BLOCKS(CIEHIP_BLOCKS_CASE)
#undef CIEHIP_BLOCKS_CASE
        default: return false;
        }
    }
    if (launch.kernel == CIE_HIP_SPMV_TILED) {
        if (launch.width != kSpmvBlock) return false;
        switch (launch.unroll) {
#define CIEHIP_TILES_CASE(U)                               \
    case U:                                                \
        f(BlockTiles<TV, TM, U, kSpmvBlock>{});            \
        return true;
            CIEHIP_TILES(CIEHIP_TILES_CASE)
#undef CIEHIP_TILES_CASE
        default: return false;
        }
    }
    if (launch.kernel == CIE_HIP_SPMV_STAGED) {
        switch (launch.width * 16 + launch.unroll) {
#define CIEHIP_STAGE_CASE(C, U)                                   \
    case C * 16 + U:                                              \
        f(StagedRows<TV, TM, C, U, kSpmvBlock>{});                \
        return true;
            CIEHIP_STAGES(CIEHIP_STAGE_CASE)
#undef CIEHIP_STAGE_CASE
        default: return false;
        }
    }
    return withShape(launch.width, launch.unroll, [&](auto lanes, auto unroll) {
        f(VectorRows<TV, TM, decltype(lanes)::value, decltype(unroll)::value, kSpmvBlock>{});
    });
}

inline int nextPow2(int v) {
    int p = 1;
    while (p < v) p *= 2;
    return p;
}

inline int roundUp(int64_t value, int64_t multiple) { return int((value + multiple - 1) / multiple * multiple); }

// Blocks of `kernel` that are resident on one compute unit at `block` threads
// (HIP occupancy query, cached per kernel).
int residentBlocks(const void* kernel, int block);

// Grid and row source of a launch of `kernel`. Vector and staged walkers get
// one resident round of blocks: compute units x blocks per unit (rounded down
// to a multiple of 8 so each XCD gets a contiguous run), but no more blocks
// than there is work for; param is each block's share. The tiled walker gets
// one block per tile (the tiles are cut on first use), or one resident round
// with `residentRound` (for launches that usually exit at once, where tens of
// thousands of empty blocks would cost more than the loop). A grid reduction
// over the launch writes one partial per block (Walker::slot()).
template <class Kernel>
int prepareLaunch(cie_hip_csr m, const CsrLaunch& launch, Kernel kernel, RowSource& source, int& grid,
                  bool residentRound = false) {
    const int perUnit = residentBlocks(reinterpret_cast<const void*>(kernel), launch.block);
    int cap = std::max(1, m->context->computeUnits) * perUnit;
    if (cap >= kXcds) cap -= cap % kXcds;
    cap = std::min(cap, kMaxPartials);
    source.rows = int(m->rows);
    source.rowExtents = m->rowExtents;
    source.columnIndices = m->columnIndices;
    source.tiles = nullptr;
    source.tileCount = 0;
    source.narrowRuns = 0;
    source.narrowColumns = nullptr;
    source.narrowBases = nullptr;
    source.blockValues64 = nullptr;
    source.blockValues32 = nullptr;
    source.tileBases = nullptr;
    source.sortedKeys = nullptr;
    source.sortedValues64 = nullptr;
    source.sortedValues32 = nullptr;
    source.sortedBlocks = 0;
    source.pairedBegins = nullptr;
    source.pairedBases = nullptr;
    source.pairedWide = 0;
    const int64_t rows = m->rows;
    if (launch.kernel == CIE_HIP_SPMV_PAIRED) {  // wide and narrow column blocks in column-sorted tiles
        const PairedForm* form = nullptr;
        const int status = ensurePairedForm(m, pairedWide(launch.width), pairedNarrow(launch.width), launch.unroll,
                                            &form);
        if (status != CIE_HIP_OK) return status;
        source.rowExtents = form->extents;
        source.columnIndices = nullptr;
        source.tiles = form->tiles;
        source.tileCount = form->tileCount;
        source.sortedKeys = form->keys;
        source.sortedValues64 = form->values64;
        source.sortedValues32 = form->values32;
        source.sortedBlocks = form->blocks[0] + form->blocks[1];
        source.pairedBegins = form->begins;
        source.pairedBases = form->bases;
        source.pairedWide = form->blocks[0];
        source.param = 1;
        grid = std::max(1, residentRound ? std::min(form->tileCount, cap) : form->tileCount);
    } else if (launch.kernel == CIE_HIP_SPMV_SORTED) {  // the column-block form with column-sorted tiles
        cie_hip_csr blocks = nullptr;
        int status = ensureColumnBlocks(m, launch.width, &blocks);
        if (status != CIE_HIP_OK) return status;
        const TileSet* set = nullptr;
        status = ensureSortedTiles(blocks, launch.block * launch.unroll, &set);
        if (status != CIE_HIP_OK) return status;
        source.rowExtents = blocks->rowExtents;
        source.columnIndices = blocks->columnIndices;
        source.tiles = set->tiles;
        source.tileCount = set->count;
        source.blockValues64 = blocks->entries64;
        source.blockValues32 = blocks->entries32;
        source.tileBases = set->sorted.bases;
        source.sortedKeys = set->sorted.keys;
        source.sortedValues64 = set->sorted.values64;
        source.sortedValues32 = set->sorted.values32;
        source.sortedBlocks = blocks->nnz;
        source.param = 1;
        grid = std::max(1, residentRound ? std::min(set->count, cap) : set->count);
    } else if (launch.kernel == CIE_HIP_SPMV_BLOCKS) {  // the tiled walker over the column-block form
        cie_hip_csr blocks = nullptr;
        int status = ensureColumnBlocks(m, launch.width, &blocks);
        if (status != CIE_HIP_OK) return status;
        const TileSet* set = nullptr;
        status = ensureTiles(blocks, launch.block * launch.unroll, &set);
        if (status != CIE_HIP_OK) return status;
        status = ensureNarrowColumns(blocks);
        if (status != CIE_HIP_OK) return status;
        source.rowExtents = blocks->rowExtents;
        source.columnIndices = blocks->columnIndices;
        source.tiles = set->tiles;
        source.tileCount = set->count;
        if (blocks->narrowState > 0) {
            source.narrowRuns = blocks->narrowRuns;
            source.narrowColumns = blocks->narrowColumns;
            source.narrowBases = blocks->narrowBases;
        }
        source.blockValues64 = blocks->entries64;
        source.blockValues32 = blocks->entries32;
        source.param = 1;
        grid = std::max(1, residentRound ? std::min(set->count, cap) : set->count);
    } else if (launch.kernel == CIE_HIP_SPMV_TILED) {
        const TileSet* set = nullptr;
        const int status = ensureTiles(m, launch.block * launch.unroll, &set);
        if (status != CIE_HIP_OK) return status;
        source.tiles = set->tiles;
        source.tileCount = set->count;
        const int narrow = ensureNarrowColumns(m);
        if (narrow != CIE_HIP_OK) return narrow;
        if (m->narrowState > 0) {
            source.narrowRuns = m->narrowRuns;
            source.narrowColumns = m->narrowColumns;
            source.narrowBases = m->narrowBases;
        }
        source.param = 1;
        grid = std::max(1, residentRound ? std::min(set->count, cap) : set->count);
    } else if (launch.kernel == CIE_HIP_SPMV_STAGED) {
        const int64_t tiles = std::max<int64_t>(1, (rows + kWave - 1) / kWave);
        const int64_t wavesPerBlock = launch.block / kWave;
        grid = int(std::min<int64_t>((tiles + wavesPerBlock - 1) / wavesPerBlock, cap));
        source.param = int((tiles + grid - 1) / grid);
    } else {
        const int64_t rowsPerStep = launch.block / launch.width;
        const int64_t steps = std::max<int64_t>(1, (rows + rowsPerStep - 1) / rowsPerStep);
        grid = int(std::min<int64_t>(steps, cap));
        const int64_t perBlock = (rows + grid - 1) / grid;
        source.param = std::max<int>(int(rowsPerStep), roundUp(perBlock, rowsPerStep));
    }
    return CIE_HIP_OK;
}

// Starting point before autotuning: block tiles of 2048 entries (handles
// rows of any length). The staged kernel (rows bit-identical to the CPU loop)
// and vector rows are selected with cie_hip_csr_set_launch or the autotuner.
inline CsrLaunch defaultLaunch() {
    CsrLaunch launch;
    launch.kernel = CIE_HIP_SPMV_TILED;
    launch.width = kSpmvBlock;
    launch.unroll = 8;
    launch.block = kSpmvBlock;
    return launch;
}

// Grid of the elementwise/reduction kernels: depends only on n.
inline int vectorGrid(int64_t n) {
    const int64_t blocks = (n + int64_t(kReduceBlock) * 4 - 1) / (int64_t(kReduceBlock) * 4);
    return int(std::max<int64_t>(8, std::min<int64_t>(2048, roundUp(std::max<int64_t>(blocks, 1), kXcds))));
}

}  // namespace ciehip

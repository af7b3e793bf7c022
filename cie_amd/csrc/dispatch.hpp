// Compile-time dispatch over CSR launch shapes and launch-size helpers.
#pragma once

#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace ciehip {

constexpr int kSpmvBlock = 256;

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

inline int nextPow2(int v) {
    int p = 1;
    while (p < v) p *= 2;
    return p;
}

inline int roundUp(int64_t value, int64_t multiple) { return int((value + multiple - 1) / multiple * multiple); }

// Starting point before autotuning: about one pass of L*U entries per row.
inline CsrLaunch defaultLaunch(double meanRowLength, int computeUnits, int64_t rows) {
    CsrLaunch launch;
    launch.block = kSpmvBlock;
    int unroll = 4;
    int lanes = nextPow2(int(std::ceil(std::max(1.0, meanRowLength) / unroll)));
    if (lanes < 4) {
        unroll = 2;
        lanes = nextPow2(int(std::ceil(std::max(1.0, meanRowLength) / unroll)));
    }
    launch.lanes = std::min(64, std::max(1, lanes));
    launch.unroll = unroll;
    launch.grid = 0;  // filled by spmvGrid
    (void)computeUnits;
    (void)rows;
    return launch;
}

// Grid of the row-chunked SpMV kernels: a multiple of 8 (one run per XCD),
// at most 8 blocks per compute unit, at least one row group per block.
inline int spmvGrid(int64_t rows, const CsrLaunch& launch, int computeUnits) {
    const int64_t rowsPerIteration = launch.block / launch.lanes;
    const int64_t groups = std::max<int64_t>(1, (rows + rowsPerIteration - 1) / rowsPerIteration);
    const int64_t cap = std::max(8, computeUnits * 8);
    int grid = int(std::min<int64_t>(groups, cap));
    grid = roundUp(grid, kXcds);
    return std::min(grid, kMaxPartials);
}

inline int rowsPerChunk(int64_t rows, const CsrLaunch& launch) {
    const int64_t rowsPerIteration = launch.block / launch.lanes;
    const int64_t perBlock = (rows + launch.grid - 1) / launch.grid;
    return std::max<int>(int(rowsPerIteration), roundUp(perBlock, rowsPerIteration));
}

// Grid of the elementwise/reduction kernels: depends only on n.
inline int vectorGrid(int64_t n) {
    const int64_t blocks = (n + int64_t(kReduceBlock) * 4 - 1) / (int64_t(kReduceBlock) * 4);
    return int(std::max<int64_t>(8, std::min<int64_t>(2048, roundUp(std::max<int64_t>(blocks, 1), kXcds))));
}

}  // namespace ciehip

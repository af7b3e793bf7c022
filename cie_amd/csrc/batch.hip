// Batched Jacobi-PCG: many independent SPD systems of uneven size solved
// together. No counterpart in the reference (it solves one system per call,
// ConjugateGradients_impl.hpp:49-201); each system follows that algorithm.
//
// Layout: the systems are the diagonal blocks of one CSR matrix (rows of
// system s are [offsets[s], offsets[s+1]) and its columns stay inside that
// range). Two work lists are cut once per batch, on the host:
//   * block tiles for the SpMV kernels (whole rows of one system, at most
//     kTileEntries entries and rowsPerTile[s] rows: the same block-tile
//     kernel as the single-system SpMV, one workgroup per tile);
//   * chunks of at most kChunkRows rows of one system for the vector kernels
//     (kChunkRowsPerThread rows per thread).
// Each tile or chunk writes its share of its system's dot products; one block
// per system then sums that system's partials in order and advances its
// scalars (rho, alpha, beta, residuals, flags), kept on the device.
//
// Convergence mask: after every trip a one-block kernel lists the systems
// still running with the prefix sums of their tile and chunk counts, and a
// second kernel expands that into the list of their tiles and chunks; the work
// kernels walk only those items, so finished systems cost nothing. The mixed
// form builds a third list, of the tiles of the systems that refine this trip,
// for its FP64 residual kernel. Grids are fixed (the chain of one trip is
// captured once in a HIP graph and replayed until every system is done); the
// item counts are read on the device.
//
// Rows per tile (32, 64, 128 or 256) are chosen per system by
// cie_hip_batch_autotune from timed SpMV runs of the system's own tiles.
#include "dispatch.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

using namespace ciehip;

namespace {

constexpr int kTileBlock = 256;
constexpr int kTileUnroll = 8;
constexpr int kTileEntries = kTileBlock * kTileUnroll;  // entries per tile at most
constexpr int kChunkThreads = 256;                       // threads per vector chunk
constexpr int kChunkRowsPerThread = 4;                   // rows per thread, kChunkThreads apart
constexpr int kChunkRows = kChunkThreads * kChunkRowsPerThread;  // rows per vector chunk
constexpr int kCompactThreads = 1024;
constexpr int kGraphIterations = 8;
constexpr int kDefaultTileRows = 256;

constexpr int kBatchBlockWidth = 3;  // column-block width of a batch that walks blocks
constexpr int kBatchPaired = 53;     // walker code of paired 1 x 5 / 1 x 3 tiles
// Blocks per lane and width of paired tiles: a tile holds at most 256 rows, and
// 2D rows hold few blocks (about 2 of each width on 2D Q2), so 512 per width
// fill the tiles the row cap allows.
constexpr int kBatchPairedUnroll = 2;

// The tile walker of a batch by its code: matrix entries (0), 1 x 3 column
// blocks (3), the same blocks in column-sorted tiles (-3), or paired 1 x 5 /
// 1 x 3 tiles (kBatchPaired).
template <class T, int W>
struct BatchWalker {
    using Type = BlockTiles<T, T, kTileUnroll, kTileBlock>;
};
template <class T>
struct BatchWalker<T, kBatchBlockWidth> {
    using Type = ColumnBlockTiles<T, T, kTileUnroll, kTileBlock, kBatchBlockWidth>;
};
template <class T>
struct BatchWalker<T, -kBatchBlockWidth> {
    using Type = SortedBlockTiles<T, T, kTileUnroll, kTileBlock, kBatchBlockWidth>;
};
template <class T>
struct BatchWalker<T, kBatchPaired> {
    using Type = PairedBlockTiles<T, T, kBatchPairedUnroll, kTileBlock, kBatchPaired / 10, kBatchPaired % 10>;
};
template <class T, int W>
using BatchTiles = typename BatchWalker<T, W>::Type;
static_assert(BatchTiles<double, kBatchPaired>::kCapacity <= kTileEntries, "paired tiles stage in the tile's LDS");

struct SystemState {
    double rho, alpha, beta, rr, initialNorm, absolute, relative, updateNorm;
    double recursiveRho;  // mixed: r^T z of the recursive residual on the trip that folds
    uint64_t iteration, refinements;
    int32_t done, refine;
};

struct BatchSettings {
    double absoluteTolerance, relativeTolerance, updateThreshold;
    uint64_t maxIterations;
    int32_t restart, mixed;
    int32_t rowSumCorrection;  // mixed: the FP32 product adds c_i p_i (the FP32 plane's row-sum correction)
    int32_t unused;
};

// A work list: the selected systems (in system order), the prefix sums of
// their tile and chunk counts (count + 1 entries each; the last is the total),
// and the expanded items {tile or chunk, system}.
struct ActiveWork {
    int32_t* systems;
    int32_t* tilePrefix;
    int32_t* chunkPrefix;
    int32_t* count;  // [0] systems, [1] tiles, [2] chunks
    int2* tileList;
    int2* chunkList;
};

enum Selection : int { kRunning = 0, kRefining = 1 };

// Where the work of a batch lives on the device.
struct BatchLayout {
    RowSource source;                 // block tiles of all systems (tiles / tileCount)
    const int32_t* systemTiles;       // systems + 1: first tile of each system
    const int32_t* systemChunks;      // systems + 1: first chunk of each system
    const int32_t* chunkRows;         // chunks + 1: first row of each chunk
};

enum Phase : int { kPhaseStart = 0, kPhaseAlpha = 1, kPhaseUpdate = 2, kPhaseUpdateMixed = 3, kPhaseRefine = 4 };

// Largest k < count with prefix[k] <= item (the prefix is increasing).
__device__ __forceinline__ int locate(int item, const int32_t* __restrict__ prefix, int count) {
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// Work kernels: item = blockIdx.x, blockIdx.x + gridDim.x, ... over the active
// tiles (tile kernels) or chunks (vector kernels); with one block per item
// the loop runs at most once.
// ---------------------------------------------------------------------------

// r = b - A x (FP64), z = D^-1 r; start: p = z. Tile partials r.r (FP64), r.z (TR).
template <class TR, bool kStart, int W>
__global__ __launch_bounds__(kTileBlock) void batchResidualKernel(
    BatchLayout layout, ActiveWork work, const double* __restrict__ entries, const double* __restrict__ x,
    const double* __restrict__ b, TR* __restrict__ r, const TR* __restrict__ inverse, TR* __restrict__ p,
    double* __restrict__ partials, const SystemState* __restrict__ state) {
    __shared__ double stage[kTileEntries];
    __shared__ int32_t extents[kTileBlock + 1];
    __shared__ double ldsRR[kTileBlock / kWave];
    __shared__ TR ldsRZ[kTileBlock / kWave];
    const int items = work.count[1];
    for (int item = blockIdx.x; item < items; item += gridDim.x) {
        const int t = work.tileList[item].x;
        double rr = 0.0;
        TR rz = TR(0);
        auto epilogue = [&](int row, double sum) {
            const double residual64 = b[row] - sum;
            const TR residual = TR(residual64);
            const TR z = residual * inverse[row];
            r[row] = residual;
            if (kStart) p[row] = z;
            rr += residual64 * residual64;
            rz += residual * z;
        };
        BatchTiles<double, W>::tile(t, layout.source, entries, x, epilogue, stage, extents);
        rr = blockSum<kTileBlock>(rr, ldsRR);
        rz = blockSum<kTileBlock>(rz, ldsRZ);
        if (threadIdx.x == 0) {
            partials[2 * t] = rr;
            partials[2 * t + 1] = double(rz);
        }
        __syncthreads();  // stage and extents are reused by the next item
    }
}

// q = A p; tile partial p.q. With `correction` (the mixed form), q_i gets
// c_i p_i added: the FP32 plane's row-sum correction. A tile holds at most
// kTileBlock rows, and thread `tid` owns row first + tid: it loads p (and the
// correction) of that row when the tile starts, so the loads are in flight
// while the tile streams; the walker leaves each row sum in LDS, and the
// owning thread writes q (coalesced) and adds its p.q term.
template <class T, int W>
__global__ __launch_bounds__(kTileBlock) void batchProductKernel(BatchLayout layout, ActiveWork work,
                                                                 const T* __restrict__ entries, const T* __restrict__ p,
                                                                 T* __restrict__ q, double* __restrict__ partials,
                                                                 const T* __restrict__ correction) {
    __shared__ T stage[kTileEntries];
    __shared__ int32_t extents[kTileBlock + 1];
    __shared__ T rowSums[kTileBlock];
    __shared__ T lds[kTileBlock / kWave];
    const int items = work.count[1];
    const int tid = threadIdx.x;
    for (int item = blockIdx.x; item < items; item += gridDim.x) {
        const int t = work.tileList[item].x;
        const int first = layout.source.tiles[t].x;
        const int rows = layout.source.tiles[t + 1].x - first;
        T pi = T(0), ci = T(0);
        if (tid < rows) {
            pi = p[first + tid];
            if constexpr (sizeof(T) == sizeof(float))  // only the FP32 plane has a correction
                if (correction) ci = correction[first + tid];
        }
        auto stash = [&](int row, T sum) { rowSums[row - first] = sum; };
        BatchTiles<T, W>::tile(t, layout.source, entries, p, stash, stage, extents);
        __syncthreads();
        T pq = T(0);
        if (tid < rows) {
            const T sum = rowSums[tid];
            T value = sum;
            if constexpr (sizeof(T) == sizeof(float))
                if (correction) value = sum + ci * pi;
            q[first + tid] = value;
            pq = pi * value;
        }
        pq = blockSum<kTileBlock>(pq, lds);
        if (tid == 0) partials[2 * t] = double(pq);
        __syncthreads();  // stage, extents and rowSums are reused by the next item
    }
}

// The same product over an explicit tile range (the autotuner's timing runs).
template <class T, int W>
__global__ __launch_bounds__(kTileBlock) void batchTileRangeKernel(RowSource source, int firstTile,
                                                                   const T* __restrict__ entries,
                                                                   const T* __restrict__ p, T* __restrict__ q) {
    __shared__ T stage[kTileEntries];
    __shared__ int32_t extents[kTileBlock + 1];
    auto epilogue = [&](int row, T sum) { q[row] = sum; };
    BatchTiles<T, W>::tile(firstTile + int(blockIdx.x), source, entries, p, epilogue, stage, extents);
}

// The vector kernels walk chunks of kChunkRows rows of one system: a block
// of kChunkThreads threads, each owning kChunkRowsPerThread rows
// kChunkThreads apart (coalesced). A thread issues all its rows' loads as
// soon as the chunk's row range is known, before it reads the system's state,
// so one block keeps several loads per thread in flight behind a single
// dependent chain (item -> chunk rows -> data).
struct ChunkRows {
    int first, end;
    __device__ __forceinline__ int row(int k) const { return first + int(threadIdx.x) + k * kChunkThreads; }
    __device__ __forceinline__ bool has(int k) const { return row(k) < end; }
};

__device__ __forceinline__ ChunkRows chunkRowsOf(const BatchLayout& layout, int c) {
    return {layout.chunkRows[c], layout.chunkRows[c + 1]};
}

// x += alpha p; r -= alpha q; z = D^-1 r; chunk partials r.r and r.z.
template <class T, class TX>
__global__ __launch_bounds__(kChunkThreads) void batchUpdateKernel(BatchLayout layout, ActiveWork work,
                                                                   TX* __restrict__ x, T* __restrict__ r,
                                                                   const T* __restrict__ p, const T* __restrict__ q,
                                                                   const T* __restrict__ inverse,
                                                                   double* __restrict__ partials,
                                                                   const SystemState* __restrict__ state) {
    constexpr int R = kChunkRowsPerThread;
    __shared__ T lds[kChunkThreads / kWave];
    __shared__ T lds2[kChunkThreads / kWave];
    const int items = work.count[2];
    for (int item = blockIdx.x; item < items; item += gridDim.x) {
        const int2 chunk = work.chunkList[item];
        const int c = chunk.x;
        const ChunkRows rows = chunkRowsOf(layout, c);
        TX xs[R];
        T ps[R], qs[R], rs[R], is[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (rows.has(k)) {
                const int i = rows.row(k);
                xs[k] = x[i];
                ps[k] = p[i];
                qs[k] = q[i];
                rs[k] = r[i];
                is[k] = inverse[i];
            }
        }
        const SystemState& s = state[chunk.y];
        T rr = T(0), rz = T(0);
        if (!s.done) {  // NOT_SPD may end a system between the product and this kernel
            const T alpha = T(s.alpha);
#pragma unroll
            for (int k = 0; k < R; ++k) {
                if (rows.has(k)) {
                    const int i = rows.row(k);
                    x[i] = xs[k] + TX(alpha) * TX(ps[k]);
                    const T residual = rs[k] + (-alpha) * qs[k];
                    r[i] = residual;
                    const T z = residual * is[k];
                    rr += residual * residual;
                    rz += residual * z;
                }
            }
        }
        rr = blockSum<kChunkThreads>(rr, lds);
        rz = blockSum<kChunkThreads>(rz, lds2);
        if (threadIdx.x == 0) {
            partials[2 * c] = double(rr);
            partials[2 * c + 1] = double(rz);
        }
    }
}

// Mixed form: the FP32 correction d += alpha p, folded into the FP64 x (and
// cleared) for a system that refines this trip; r -= alpha q; z = D^-1 r;
// chunk partials r.r and r.z.
__global__ __launch_bounds__(kChunkThreads) void batchMixedUpdateKernel(BatchLayout layout, ActiveWork work,
                                                                        double* __restrict__ x, float* __restrict__ d,
                                                                        float* __restrict__ r,
                                                                        const float* __restrict__ p,
                                                                        const float* __restrict__ q,
                                                                        const float* __restrict__ inverse,
                                                                        double* __restrict__ partials,
                                                                        const SystemState* __restrict__ state) {
    constexpr int R = kChunkRowsPerThread;
    __shared__ float lds[kChunkThreads / kWave];
    __shared__ float lds2[kChunkThreads / kWave];
    const int items = work.count[2];
    for (int item = blockIdx.x; item < items; item += gridDim.x) {
        const int2 chunk = work.chunkList[item];
        const int c = chunk.x;
        const ChunkRows rows = chunkRowsOf(layout, c);
        float ds[R], ps[R], qs[R], rs[R], is[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (rows.has(k)) {
                const int i = rows.row(k);
                ds[k] = d[i];
                ps[k] = p[i];
                qs[k] = q[i];
                rs[k] = r[i];
                is[k] = inverse[i];
            }
        }
        const SystemState& s = state[chunk.y];
        float rr = 0.0f, rz = 0.0f;
        if (!s.done) {  // NOT_SPD may end a system between the product and this kernel
            const float alpha = float(s.alpha);
            const bool refine = s.refine;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                if (rows.has(k)) {
                    const int i = rows.row(k);
                    const float correction = ds[k] + alpha * ps[k];
                    if (refine) {
                        x[i] = x[i] + double(correction);
                        d[i] = 0.0f;
                    } else {
                        d[i] = correction;
                    }
                    const float residual = rs[k] + (-alpha) * qs[k];
                    r[i] = residual;
                    const float z = residual * is[k];
                    rr += residual * residual;
                    rz += residual * z;
                }
            }
        }
        rr = blockSum<kChunkThreads>(rr, lds);
        rz = blockSum<kChunkThreads>(rz, lds2);
        if (threadIdx.x == 0) {
            partials[2 * c] = double(rr);
            partials[2 * c + 1] = double(rz);
        }
    }
}

// p = beta p + D^-1 r
template <class T>
__global__ __launch_bounds__(kChunkThreads) void batchDirectionKernel(BatchLayout layout, ActiveWork work,
                                                                      T* __restrict__ p, const T* __restrict__ r,
                                                                      const T* __restrict__ inverse,
                                                                      const SystemState* __restrict__ state) {
    constexpr int R = kChunkRowsPerThread;
    const int items = work.count[2];
    for (int item = blockIdx.x; item < items; item += gridDim.x) {
        const int2 chunk = work.chunkList[item];
        const ChunkRows rows = chunkRowsOf(layout, chunk.x);
        T ps[R], rs[R], is[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (rows.has(k)) {
                const int i = rows.row(k);
                ps[k] = p[i];
                rs[k] = r[i];
                is[k] = inverse[i];
            }
        }
        const SystemState& s = state[chunk.y];
        if (s.done) continue;
        const T beta = T(s.beta);
#pragma unroll
        for (int k = 0; k < R; ++k)
            if (rows.has(k)) p[rows.row(k)] = ps[k] * beta + rs[k] * is[k];
    }
}

// ---------------------------------------------------------------------------
// Per-system finish: one block per system sums its partials in order (tile
// partials after the SpMV kernels, chunk partials after the update).
// ---------------------------------------------------------------------------

template <class TR, int kPhase>
__global__ __launch_bounds__(kTileBlock) void batchFinishKernel(const int32_t* __restrict__ systemItems,
                                                                const double* __restrict__ partials,
                                                                SystemState* __restrict__ state,
                                                                BatchSettings settings, int* active) {
    __shared__ double lds64[kTileBlock / kWave];
    __shared__ TR ldsR[kTileBlock / kWave];
    SystemState& s = state[blockIdx.x];
    if (s.done) return;
    if (kPhase == kPhaseRefine && !s.refine) return;
    const int begin = systemItems[blockIdx.x], end = systemItems[blockIdx.x + 1];
    // first value: r.r (FP64 in the residual phases, TR otherwise) or p.q; second: r.z (TR)
    double first64 = 0.0;
    TR firstR = TR(0), second = TR(0);
    for (int t = begin + threadIdx.x; t < end; t += kTileBlock) {
        first64 += partials[2 * t];
        firstR += TR(partials[2 * t]);
        second += TR(partials[2 * t + 1]);
    }
    constexpr bool kFirst64 = kPhase == kPhaseStart || kPhase == kPhaseRefine;
    double A64 = 0.0;
    TR AR = TR(0), B = TR(0);
    if (kFirst64) A64 = blockSum<kTileBlock>(first64, lds64);
    else AR = blockSum<kTileBlock>(firstR, ldsR);
    if (kPhase != kPhaseAlpha) B = blockSum<kTileBlock>(second, ldsR);
    if (threadIdx.x != 0) return;
    bool finished = false;
    if (kPhase == kPhaseStart) {
        const double initial = sqrt(A64);
        s.rr = A64;
        s.initialNorm = initial;
        s.absolute = initial;
        s.relative = 1.0;
        s.updateNorm = A64;
        if (floatLess(initial, settings.absoluteTolerance) || floatLess(1.0, settings.relativeTolerance)) {
            s.done = kConverged;
        } else if (!(TR(0) < B)) {
            s.done = kBreakdown;
        } else {
            s.rho = double(B);
        }
        finished = s.done != kRunning;
    } else if (kPhase == kPhaseAlpha) {
        // Mixed: does this trip refine? The FP32 r.r the previous trip left fell by the refinement factor since
        // the last refinement, looks converged, or the iteration cap is near. The update then folds the
        // correction into x and the residual kernel recomputes r from it.
        bool refine = false;
        if (settings.mixed) {
            const double absolute = sqrt(s.rr);
            const double relative = absolute / s.initialNorm;
            refine = s.rr <= settings.updateThreshold * s.updateNorm ||
                     floatLess(absolute, settings.absoluteTolerance) ||
                     floatLess(relative, settings.relativeTolerance) || s.iteration + 1 >= settings.maxIterations;
        }
        // p.q = 0 on a refining trip: the FP32 recurrence solved the system exactly (r, then p, became
        // zero); the trip folds without a step and the true residual decides. Any other p.q <= 0 ends it.
        const bool vanished = AR == TR(0) && refine;
        if (!(TR(0) < AR) && !vanished) {
            s.done = kNotSpd;
            finished = true;
        } else {
            s.alpha = vanished ? 0.0 : double(TR(s.rho) / AR);
            if (settings.mixed) s.refine = refine;
        }
    } else if (kPhase == kPhaseUpdate) {
        const double absolute = sqrt(double(AR));
        const double relative = absolute / s.initialNorm;
        const double previous = s.rho;
        s.rho = double(B);
        s.beta = double(B) / previous;
        s.rr = double(AR);
        s.absolute = absolute;
        s.relative = relative;
        if (floatLess(absolute, settings.absoluteTolerance) || floatLess(relative, settings.relativeTolerance)) {
            s.done = kConverged;
        } else {
            s.iteration += 1;
            if (s.iteration >= settings.maxIterations) s.done = kMaxIterations;
        }
        finished = s.done != kRunning;
    } else if (kPhase == kPhaseUpdateMixed) {
        const double absolute = sqrt(double(AR));
        const double relative = absolute / s.initialNorm;
        if (s.refine) {  // a refining trip advances in the refinement phase, from the true residual
            s.recursiveRho = double(B);
        } else {
            const TR previous = TR(s.rho);
            s.rho = double(B);
            s.beta = previous == TR(0) ? 0.0 : double(B / previous);  // a vanished residual restarts p
            s.rr = double(AR);
            s.absolute = absolute;
            s.relative = relative;
            s.iteration += 1;
        }
    } else {  // kPhaseRefine
        s.refine = 0;
        s.refinements += 1;
        s.rr = A64;
        s.absolute = sqrt(A64);
        s.relative = s.absolute / s.initialNorm;
        if (floatLess(s.absolute, settings.absoluteTolerance) || floatLess(s.relative, settings.relativeTolerance)) {
            s.done = kConverged;
        } else {
            const TR previous = TR(s.rho);
            s.rho = double(B);
            // A replaced residual far above the recursive one restarts the direction (kRestartGap).
            const bool gap = double(B) > kRestartGap * s.recursiveRho;
            s.beta = settings.restart || previous == TR(0) || gap ? 0.0 : double(B / previous);
            s.updateNorm = A64;
            s.iteration += 1;
            if (s.iteration >= settings.maxIterations) s.done = kMaxIterations;
        }
        finished = s.done != kRunning;
    }
    if (finished) atomicSub(active, 1);
}

// Lists the selected systems (running, or running and refining this trip)
// and the prefix sums of their tile and chunk counts (one block; thread i
// handles a contiguous run of systems).
template <int kSelect>
__global__ __launch_bounds__(kCompactThreads) void batchCompactKernel(int systems, const SystemState* __restrict__ state,
                                                                      const int32_t* __restrict__ systemTiles,
                                                                      const int32_t* __restrict__ systemChunks,
                                                                      ActiveWork work) {
    __shared__ int32_t scanRun[kCompactThreads], scanTiles[kCompactThreads], scanChunks[kCompactThreads];
    const int tid = threadIdx.x;
    const int per = (systems + kCompactThreads - 1) / kCompactThreads;
    const int lo = min(systems, tid * per), hi = min(systems, lo + per);
    int run = 0, tiles = 0, chunks = 0;
    auto selected = [&](int s) { return !state[s].done && (kSelect == kRunning || state[s].refine); };
    for (int s = lo; s < hi; ++s) {
        if (!selected(s)) continue;
        run += 1;
        tiles += systemTiles[s + 1] - systemTiles[s];
        chunks += systemChunks[s + 1] - systemChunks[s];
    }
    scanRun[tid] = run;
    scanTiles[tid] = tiles;
    scanChunks[tid] = chunks;
    __syncthreads();
    // inclusive Hillis-Steele scan
    for (int offset = 1; offset < kCompactThreads; offset *= 2) {
        const int a = tid >= offset ? scanRun[tid - offset] : 0;
        const int b = tid >= offset ? scanTiles[tid - offset] : 0;
        const int c = tid >= offset ? scanChunks[tid - offset] : 0;
        __syncthreads();
        scanRun[tid] += a;
        scanTiles[tid] += b;
        scanChunks[tid] += c;
        __syncthreads();
    }
    int k = scanRun[tid] - run, tileAt = scanTiles[tid] - tiles, chunkAt = scanChunks[tid] - chunks;
    for (int s = lo; s < hi; ++s) {
        if (!selected(s)) continue;
        work.systems[k] = s;
        work.tilePrefix[k] = tileAt;
        work.chunkPrefix[k] = chunkAt;
        tileAt += systemTiles[s + 1] - systemTiles[s];
        chunkAt += systemChunks[s + 1] - systemChunks[s];
        k += 1;
    }
    if (tid == kCompactThreads - 1) {
        work.count[0] = scanRun[tid];
        work.count[1] = scanTiles[tid];
        work.count[2] = kSelect == kRunning ? scanChunks[tid] : 0;
        work.tilePrefix[scanRun[tid]] = scanTiles[tid];
        work.chunkPrefix[scanRun[tid]] = scanChunks[tid];
    }
}

// Expands the selected systems into their items: thread i writes tile item i
// and chunk item i, {index, system}, by binary search over the prefix sums.
__global__ __launch_bounds__(kTileBlock) void batchExpandKernel(const int32_t* __restrict__ systemTiles,
                                                                const int32_t* __restrict__ systemChunks,
                                                                ActiveWork work) {
    const int i = int(blockIdx.x) * kTileBlock + int(threadIdx.x);
    const int selected = work.count[0];
    if (i < work.count[1]) {
        const int k = locate(i, work.tilePrefix, selected);
        const int s = work.systems[k];
        work.tileList[i] = make_int2(systemTiles[s] + (i - work.tilePrefix[k]), s);
    }
    if (i < work.count[2]) {
        const int k = locate(i, work.chunkPrefix, selected);
        const int s = work.systems[k];
        work.chunkList[i] = make_int2(systemChunks[s] + (i - work.chunkPrefix[k]), s);
    }
}

// The two steps above in one kernel, for batches of at most kSelectSystems
// systems: every block ranks the selected systems itself (their prefix sums
// of tiles and chunks, in LDS; the order and sums are the compaction's) and
// writes its share of the items; block 0 writes the counts.
constexpr int kSelectSystems = 4 * kTileBlock;

template <int kSelect>
__global__ __launch_bounds__(kTileBlock) void batchSelectKernel(int systems, const SystemState* __restrict__ state,
                                                                const int32_t* __restrict__ systemTiles,
                                                                const int32_t* __restrict__ systemChunks,
                                                                ActiveWork work) {
    __shared__ int32_t selectedSystems[kSelectSystems];
    __shared__ int32_t tilePrefix[kSelectSystems + 1], chunkPrefix[kSelectSystems + 1];
    __shared__ int32_t scanRun[kTileBlock], scanTiles[kTileBlock], scanChunks[kTileBlock];
    const int tid = threadIdx.x;
    const int per = (systems + kTileBlock - 1) / kTileBlock;  // at most 4
    const int lo = min(systems, tid * per), hi = min(systems, lo + per);
    auto selected = [&](int s) { return !state[s].done && (kSelect == kRunning || state[s].refine); };
    int run = 0, tiles = 0, chunks = 0;
    for (int s = lo; s < hi; ++s) {
        if (!selected(s)) continue;
        run += 1;
        tiles += systemTiles[s + 1] - systemTiles[s];
        chunks += systemChunks[s + 1] - systemChunks[s];
    }
    scanRun[tid] = run;
    scanTiles[tid] = tiles;
    scanChunks[tid] = chunks;
    __syncthreads();
    for (int offset = 1; offset < kTileBlock; offset *= 2) {  // inclusive Hillis-Steele scan
        const int a = tid >= offset ? scanRun[tid - offset] : 0;
        const int b = tid >= offset ? scanTiles[tid - offset] : 0;
        const int c = tid >= offset ? scanChunks[tid - offset] : 0;
        __syncthreads();
        scanRun[tid] += a;
        scanTiles[tid] += b;
        scanChunks[tid] += c;
        __syncthreads();
    }
    int k = scanRun[tid] - run, tileAt = scanTiles[tid] - tiles, chunkAt = scanChunks[tid] - chunks;
    for (int s = lo; s < hi; ++s) {
        if (!selected(s)) continue;
        selectedSystems[k] = s;
        tilePrefix[k] = tileAt;
        chunkPrefix[k] = chunkAt;
        tileAt += systemTiles[s + 1] - systemTiles[s];
        chunkAt += systemChunks[s + 1] - systemChunks[s];
        k += 1;
    }
    const int count = scanRun[kTileBlock - 1];
    const int tileTotal = scanTiles[kTileBlock - 1];
    const int chunkTotal = kSelect == kRunning ? scanChunks[kTileBlock - 1] : 0;
    if (tid == 0) {
        tilePrefix[count] = tileTotal;
        chunkPrefix[count] = chunkTotal;
    }
    __syncthreads();
    if (blockIdx.x == 0 && tid == 0) {
        work.count[0] = count;
        work.count[1] = tileTotal;
        work.count[2] = chunkTotal;
    }
    const int i = int(blockIdx.x) * kTileBlock + tid;
    if (i < tileTotal) {
        const int j = locate(i, tilePrefix, count);
        const int s = selectedSystems[j];
        work.tileList[i] = make_int2(systemTiles[s] + (i - tilePrefix[j]), s);
    }
    if (i < chunkTotal) {
        const int j = locate(i, chunkPrefix, count);
        const int s = selectedSystems[j];
        work.chunkList[i] = make_int2(systemChunks[s] + (i - chunkPrefix[j]), s);
    }
}

__global__ void batchCheckKernel(int systems, const int64_t* __restrict__ offsets, const int32_t* __restrict__ rowExtents,
                                 const int32_t* __restrict__ columnIndices, int* __restrict__ bad) {
    const int s = blockIdx.x;
    if (s >= systems) return;
    const int64_t lo = offsets[s], hi = offsets[s + 1];
    for (int64_t row = lo + threadIdx.x; row < hi; row += blockDim.x)
        for (int k = rowExtents[row]; k < rowExtents[row + 1]; ++k)
            if (columnIndices[k] < lo || columnIndices[k] >= hi) atomicOr(bad, 1);
}

}  // namespace

// ---------------------------------------------------------------------------
// Batch handle
// ---------------------------------------------------------------------------

struct cie_hip_batch_s {
    cie_hip_context context = nullptr;
    cie_hip_csr matrix = nullptr;
    int systems = 0;
    std::vector<int64_t> offsets;       // host, systems + 1
    std::vector<int32_t> tileRows;      // rows per tile of each system
    int2* tiles = nullptr;              // device, tileCount + 1: {first row, first entry}
    int32_t* systemTiles = nullptr;     // device, systems + 1
    int tileCount = 0;
    int32_t* chunkRows = nullptr;       // device, chunkCount + 1
    int32_t* systemChunks = nullptr;    // device, systems + 1
    int chunkCount = 0;
    ActiveWork running{};               // device work list of the running systems
    ActiveWork refining{};              // device work list of the systems refining this trip (mixed)
    SystemState* state = nullptr;
    double* tilePartials = nullptr;     // 2 per tile
    double* chunkPartials = nullptr;    // 2 per chunk
    int* active = nullptr;
    int* pinnedActive = nullptr;
    SystemState* pinnedStates = nullptr;
    void* workspace = nullptr;
    size_t workspaceBytes = 0;
    hipGraphExec_t graph = nullptr;
    int blockWidth = 0;                 // 0: tiles of entries; kBatchBlockWidth: tiles of column blocks
    bool sortedTiles = false;           // tiles of column blocks stored in column order (blockWidth > 0)
    SortedCopy sorted;                  // the sorted copy of the block form over the batch's tiles
    bool pairedTiles = false;           // paired 1 x 5 / 1 x 3 tiles (blockWidth 0)
    PairedForm* paired = nullptr;       // the paired form cut into the batch's tiles
    int graphKey[5] = {0, 0, 0, 0, 0};  // what the captured graph was built for
    const void* graphPointers[3] = {nullptr, nullptr, nullptr};  // b, x, inverse diagonal
    BatchSettings graphSettings{};      // the settings its kernels were captured with (passed by value)
};

namespace {

void releaseGraph(cie_hip_batch batch) {
    if (batch->graph) (void)hipGraphExecDestroy(batch->graph);
    batch->graph = nullptr;
}

// The walker code of the batch's tile format (the W of BatchTiles).
int walkerCode(const cie_hip_batch_s* batch) {
    if (batch->pairedTiles) return kBatchPaired;
    return batch->blockWidth ? (batch->sortedTiles ? -batch->blockWidth : batch->blockWidth) : 0;
}

template <class T>
void freeDevice(T*& pointer) {
    if (pointer) (void)hipFree(pointer);
    pointer = nullptr;
}

// Cuts every system into block tiles (<= kTileEntries entries, <= tileRows[s]
// rows; a longer row is a tile of its own) and into chunks of kChunkRows rows.
// Paired tiles are cut by the paired form's builder with the same rules
// (<= tileRows[s] rows, <= 256 * kBatchPairedUnroll blocks of each width;
// a matrix with a longer row does not fit).
int buildWork(cie_hip_batch batch) {
    std::vector<int2> tiles;
    std::vector<int32_t> systemTiles(size_t(batch->systems) + 1, 0);
    std::vector<int32_t> chunkRows;
    std::vector<int32_t> systemChunks(size_t(batch->systems) + 1, 0);
    const int rows = int(batch->offsets[batch->systems]);
    if (batch->pairedTiles) {
        releasePairedForm(batch->context, batch->paired);
        batch->paired = new PairedForm;
        batch->paired->wide = kBatchPaired / 10;
        batch->paired->narrow = kBatchPaired % 10;
        batch->paired->unroll = kBatchPairedUnroll;
        const int status = buildPairedForm(batch->matrix, *batch->paired,
                                           PairedCut{batch->offsets.data(), batch->tileRows.data(), batch->systems,
                                                     &systemTiles});
        if (status != CIE_HIP_OK) return status;
        if (batch->paired->state < 0) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, batch->paired->error);
        batch->tileCount = batch->paired->tileCount;
    } else {
        // Tiles are cut over the rows' entries, or over their column blocks (as many per row as the block form holds).
        cie_hip_csr form = batch->matrix;
        if (batch->blockWidth) {
            const int status = ensureColumnBlocks(batch->matrix, batch->blockWidth, &form);
            if (status != CIE_HIP_OK) return status;
        }
        const int32_t* extents = form->hostExtents.data();
        for (int s = 0; s < batch->systems; ++s) {
            systemTiles[s] = int32_t(tiles.size());
            const int lo = int(batch->offsets[s]), hi = int(batch->offsets[s + 1]);
            const int cap = batch->tileRows[s];
            int row = lo;
            while (row < hi) {
                const int first = row;
                if (extents[row + 1] - extents[row] > kTileEntries) {
                    ++row;
                } else {
                    while (row < hi && row - first < cap && extents[row + 1] - extents[first] <= kTileEntries) ++row;
                }
                tiles.push_back(make_int2(first, extents[first]));
            }
        }
        systemTiles[batch->systems] = int32_t(tiles.size());
        batch->tileCount = int(tiles.size());
        tiles.push_back(make_int2(rows, extents[rows]));
    }
    for (int s = 0; s < batch->systems; ++s) {
        systemChunks[s] = int32_t(chunkRows.size());
        for (int first = int(batch->offsets[s]); first < int(batch->offsets[s + 1]); first += kChunkRows)
            chunkRows.push_back(first);
    }
    systemChunks[batch->systems] = int32_t(chunkRows.size());
    batch->chunkCount = int(chunkRows.size());
    chunkRows.push_back(rows);
    releaseSortedCopy(batch->context, batch->sorted);  // it follows the tiles
    if (!batch->pairedTiles) {
        releasePairedForm(batch->context, batch->paired);
        batch->paired = nullptr;
    }
    freeDevice(batch->tiles);
    freeDevice(batch->chunkRows);
    freeDevice(batch->tilePartials);
    freeDevice(batch->chunkPartials);
    for (ActiveWork* work : {&batch->running, &batch->refining}) {
        freeDevice(work->tileList);
        freeDevice(work->chunkList);
        CIE_HIP_CHECK(hipMalloc(&work->tileList, sizeof(int2) * std::max(1, batch->tileCount)));
        CIE_HIP_CHECK(hipMalloc(&work->chunkList, sizeof(int2) * std::max(1, batch->chunkCount)));
    }
    if (!tiles.empty()) {  // paired tiles live in the paired form
        CIE_HIP_CHECK(hipMalloc(&batch->tiles, sizeof(int2) * tiles.size()));
        CIE_HIP_CHECK(hipMemcpy(batch->tiles, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice));
    }
    CIE_HIP_CHECK(hipMalloc(&batch->chunkRows, sizeof(int32_t) * chunkRows.size()));
    CIE_HIP_CHECK(hipMalloc(&batch->tilePartials, 2 * sizeof(double) * std::max(1, batch->tileCount)));
    CIE_HIP_CHECK(hipMalloc(&batch->chunkPartials, 2 * sizeof(double) * std::max(1, batch->chunkCount)));
    CIE_HIP_CHECK(hipMemcpy(batch->chunkRows, chunkRows.data(), sizeof(int32_t) * chunkRows.size(),
                            hipMemcpyHostToDevice));
    CIE_HIP_CHECK(hipMemcpy(batch->systemTiles, systemTiles.data(), sizeof(int32_t) * systemTiles.size(),
                            hipMemcpyHostToDevice));
    CIE_HIP_CHECK(hipMemcpy(batch->systemChunks, systemChunks.data(), sizeof(int32_t) * systemChunks.size(),
                            hipMemcpyHostToDevice));
    releaseGraph(batch);
    return CIE_HIP_OK;
}

int batchLayout(cie_hip_batch batch, BatchLayout* layout) {
    cie_hip_csr m = batch->matrix;
    if (batch->pairedTiles) {  // the paired form, its values current
        PairedForm& form = *batch->paired;
        const int status = refreshPairedValues(m, form);
        if (status != CIE_HIP_OK) return status;
        RowSource& source = layout->source;
        source = RowSource{};
        source.rows = int(m->rows);
        source.param = 1;
        source.rowExtents = form.extents;
        source.tiles = form.tiles;
        source.tileCount = form.tileCount;
        source.sortedKeys = form.keys;
        source.sortedValues64 = form.values64;
        source.sortedValues32 = form.values32;
        source.sortedBlocks = form.blocks[0] + form.blocks[1];
        source.pairedBegins = form.begins;
        source.pairedBases = form.bases;
        source.pairedWide = form.blocks[0];
        layout->systemTiles = batch->systemTiles;
        layout->systemChunks = batch->systemChunks;
        layout->chunkRows = batch->chunkRows;
        return CIE_HIP_OK;
    }
    if (batch->blockWidth) {  // the block form, its values current
        const int status = ensureColumnBlocks(batch->matrix, batch->blockWidth, &m);
        if (status != CIE_HIP_OK) return status;
    }
    int status = ensureNarrowColumns(m);
    if (status != CIE_HIP_OK) return status;
    RowSource& source = layout->source;
    source.tileBases = nullptr;
    source.sortedKeys = nullptr;
    source.sortedValues64 = nullptr;
    source.sortedValues32 = nullptr;
    source.sortedBlocks = 0;
    source.pairedBegins = nullptr;
    source.pairedBases = nullptr;
    source.pairedWide = 0;
    if (batch->blockWidth && batch->sortedTiles) {
        status = ensureSortedCopy(m, batch->tiles, batch->tileCount, kTileEntries, batch->sorted);
        if (status != CIE_HIP_OK) return status;
        source.tileBases = batch->sorted.bases;
        source.sortedKeys = batch->sorted.keys;
        source.sortedValues64 = batch->sorted.values64;
        source.sortedValues32 = batch->sorted.values32;
        source.sortedBlocks = m->nnz;
    }
    source.blockValues64 = batch->blockWidth ? m->entries64 : nullptr;
    source.blockValues32 = batch->blockWidth ? m->entries32 : nullptr;
    source.rows = int(m->rows);
    source.param = 1;
    source.rowExtents = m->rowExtents;
    source.columnIndices = m->columnIndices;
    source.tiles = batch->tiles;
    source.tileCount = batch->tileCount;
    source.narrowRuns = m->narrowState > 0 ? m->narrowRuns : 0;
    source.narrowColumns = m->narrowState > 0 ? m->narrowColumns : nullptr;
    source.narrowBases = m->narrowState > 0 ? m->narrowBases : nullptr;
    layout->systemTiles = batch->systemTiles;
    layout->systemChunks = batch->systemChunks;
    layout->chunkRows = batch->chunkRows;
    return CIE_HIP_OK;
}


// Blocks of a work kernel. The main kernels get one block per item (blocks
// past the active count exit at once; the hardware balances the rest as
// blocks retire). The mixed form's residual kernel, which walks the usually
// empty list of refining systems, gets one resident round walking its items
// grid-stride, so an empty list costs a few hundred empty blocks.
template <class Kernel>
dim3 workGrid(Kernel kernel, cie_hip_context ctx, int block, int items, bool residentRound = false) {
    if (!residentRound) return dim3(std::max(1, items));
    const int resident = std::max(1, ctx->computeUnits) * residentBlocks(reinterpret_cast<const void*>(kernel), block);
    return dim3(std::max(1, std::min(items, resident)));
}

void releaseBatch(cie_hip_batch batch) {
    if (!batch) return;
    if (batch->blockWidth && batch->matrix->columnBlocks[batch->blockWidth])
        --batch->matrix->columnBlocks[batch->blockWidth]->blockUsers;
    if (batch->context) {
        (void)hipSetDevice(batch->context->device);
        (void)hipStreamSynchronize(batch->context->stream);
    }
    releaseGraph(batch);
    releaseSortedCopy(batch->context, batch->sorted);
    releasePairedForm(batch->context, batch->paired);
    freeDevice(batch->tiles);
    freeDevice(batch->systemTiles);
    freeDevice(batch->chunkRows);
    freeDevice(batch->systemChunks);
    for (ActiveWork* work : {&batch->running, &batch->refining}) {
        freeDevice(work->systems);
        freeDevice(work->tilePrefix);
        freeDevice(work->chunkPrefix);
        freeDevice(work->count);
        freeDevice(work->tileList);
        freeDevice(work->chunkList);
    }
    freeDevice(batch->state);
    freeDevice(batch->tilePartials);
    freeDevice(batch->chunkPartials);
    freeDevice(batch->active);
    if (batch->pinnedActive) (void)hipHostFree(batch->pinnedActive);
    if (batch->pinnedStates) (void)hipHostFree(batch->pinnedStates);
    freeDevice(batch->workspace);
    cie_hip_context context = batch->context;
    cie_hip_csr matrix = batch->matrix;
    delete batch;
    if (matrix) dropMatrix(matrix);
    if (context) dropContext(context);
}

// Switches the batch between tiles of entries (0) and of column blocks,
// sorted or not; the block form in use is pinned so the matrix's own
// autotuner keeps it.
int setBlockWidth(cie_hip_batch batch, int width, bool sorted = false) {
    sorted = sorted && width;
    bool rebuild = false;
    if (batch->pairedTiles) {  // leaving paired tiles: re-cut
        batch->pairedTiles = false;
        releaseGraph(batch);
        rebuild = true;
    }
    if (sorted != batch->sortedTiles) {
        releaseSortedCopy(batch->context, batch->sorted);
        batch->sortedTiles = sorted;
        releaseGraph(batch);
    }
    if (width == batch->blockWidth) return rebuild ? buildWork(batch) : CIE_HIP_OK;
    cie_hip_csr m = batch->matrix;
    if (width) {
        cie_hip_csr form = nullptr;
        const int status = ensureColumnBlocks(m, width, &form);
        if (status != CIE_HIP_OK) return status;
        ++form->blockUsers;
    }
    if (batch->blockWidth && m->columnBlocks[batch->blockWidth]) --m->columnBlocks[batch->blockWidth]->blockUsers;
    batch->blockWidth = width;
    return buildWork(batch);  // the tiles always match the walker
}

// Switches the batch to paired 1 x 5 / 1 x 3 tiles (built over the batch's
// own tile cut); fails with CIE_HIP_ERROR_INVALID_ARGUMENT where the matrix
// does not fit the paired form (the batch then keeps entry tiles).
int setPairedTiles(cie_hip_batch batch) {
    if (batch->pairedTiles) return CIE_HIP_OK;
    int status = setBlockWidth(batch, 0, false);
    if (status != CIE_HIP_OK) return status;
    batch->pairedTiles = true;
    releaseGraph(batch);
    status = buildWork(batch);
    if (status != CIE_HIP_OK) {
        const std::string message = cie_hip_last_error();
        batch->pairedTiles = false;
        if (buildWork(batch) != CIE_HIP_OK) return CIE_HIP_ERROR_HIP;
        return setError(status, message);
    }
    return CIE_HIP_OK;
}

// The batch in the tile format of walker code `code`.
int setTileFormat(cie_hip_batch batch, int code) {
    if (code == kBatchPaired) return setPairedTiles(batch);
    return setBlockWidth(batch, std::abs(code), code < 0);
}

// Calls f(integral_constant<W>) for the batch's walker code.
template <class F>
void withWalkerCode(int code, F&& f) {
    if (code == kBatchBlockWidth) f(std::integral_constant<int, kBatchBlockWidth>{});
    else if (code == -kBatchBlockWidth) f(std::integral_constant<int, -kBatchBlockWidth>{});
    else if (code == kBatchPaired) f(std::integral_constant<int, kBatchPaired>{});
    else f(std::integral_constant<int, 0>{});
}

// The batched product over tiles [t0, t0 + tiles) (the autotuner's timing runs).
void launchTileRange(int code, int precision, int tiles, hipStream_t stream, const RowSource& source, int t0,
                     cie_hip_csr m, void* buffer) {
    withWalkerCode(code, [&](auto w) {
        constexpr int W = decltype(w)::value;
        if (precision == 32) {
            float* in = static_cast<float*>(buffer);
            hipLaunchKernelGGL((batchTileRangeKernel<float, W>), dim3(tiles), dim3(kTileBlock), 0, stream, source, t0,
                               m->entries32, in, in + m->rows);
        } else {
            double* in = static_cast<double*>(buffer);
            hipLaunchKernelGGL((batchTileRangeKernel<double, W>), dim3(tiles), dim3(kTileBlock), 0, stream, source, t0,
                               m->entries64, in, in + m->rows);
        }
    });
}

// Rebuilds a work list from the system states.
template <int kSelect>
void launchSelect(cie_hip_batch batch, const ActiveWork& work) {
    const hipStream_t stream = batch->context->stream;
    const int items = std::max(batch->tileCount, kSelect == kRunning ? batch->chunkCount : 0);
    const dim3 grid(std::max(1, (items + kTileBlock - 1) / kTileBlock));
    if (batch->systems <= kSelectSystems) {
        hipLaunchKernelGGL((batchSelectKernel<kSelect>), grid, dim3(kTileBlock), 0, stream, batch->systems,
                           batch->state, batch->systemTiles, batch->systemChunks, work);
        return;
    }
    hipLaunchKernelGGL((batchCompactKernel<kSelect>), dim3(1), dim3(kCompactThreads), 0, stream, batch->systems,
                       batch->state, batch->systemTiles, batch->systemChunks, work);
    hipLaunchKernelGGL(batchExpandKernel, grid, dim3(kTileBlock), 0, stream, batch->systemTiles, batch->systemChunks,
                       work);
}

// Enqueues the kernels of one iteration on the context stream.
constexpr int kProfiledTrips = 8;  // product launches averaged by the profiling summary

template <class TR, int W>
void enqueueIteration(cie_hip_batch batch, const BatchLayout& layout, const double* b, double* x, TR* r, TR* p, TR* q,
                      float* correction, const TR* inverse, const BatchSettings& settings, bool mixed,
                      ProductProfile* profile) {
    cie_hip_context ctx = batch->context;
    cie_hip_csr m = batch->matrix;
    const hipStream_t stream = ctx->stream;
    const ActiveWork& work = batch->running;
    const dim3 tileBlock(kTileBlock), chunkBlock(kChunkThreads), systems(batch->systems);
    const TR* entries;
    if constexpr (std::is_same_v<TR, float>) entries = m->entries32;
    else entries = m->entries64;
    const dim3 productGrid = workGrid(batchProductKernel<TR, W>, ctx, kTileBlock, batch->tileCount);
    const TR* rowCorrection = nullptr;
    if constexpr (std::is_same_v<TR, float>)
        if (settings.rowSumCorrection) rowCorrection = m->rowCorrection32;
    if (profile)  // eager launches only (a captured graph is never profiled)
        profile->launch(ctx, batchProductKernel<TR, W>, productGrid, tileBlock, layout, work, entries,
                        static_cast<const TR*>(p), q, batch->tilePartials, rowCorrection);
    else
        hipLaunchKernelGGL((batchProductKernel<TR, W>), productGrid, tileBlock, 0, stream, layout, work, entries, p, q,
                           batch->tilePartials, rowCorrection);
    hipLaunchKernelGGL((batchFinishKernel<TR, kPhaseAlpha>), systems, tileBlock, 0, stream, batch->systemTiles,
                       batch->tilePartials, batch->state, settings, batch->active);
    if constexpr (std::is_same_v<TR, float>) {  // the mixed form: FP32 correction, folded when refining
        hipLaunchKernelGGL(batchMixedUpdateKernel,
                           workGrid(batchMixedUpdateKernel, ctx, kChunkThreads, batch->chunkCount), chunkBlock, 0,
                           stream, layout, work, x, correction, r, p, q, inverse, batch->chunkPartials, batch->state);
    } else {
        hipLaunchKernelGGL((batchUpdateKernel<TR, double>),
                           workGrid(batchUpdateKernel<TR, double>, ctx, kChunkThreads, batch->chunkCount), chunkBlock, 0,
                           stream, layout, work, x, r, p, q, inverse, batch->chunkPartials, batch->state);
    }
    if (mixed) {
        hipLaunchKernelGGL((batchFinishKernel<TR, kPhaseUpdateMixed>), systems, tileBlock, 0, stream,
                           batch->systemChunks, batch->chunkPartials, batch->state, settings, batch->active);
        launchSelect<kRefining>(batch, batch->refining);
        hipLaunchKernelGGL((batchResidualKernel<TR, false, W>),
                           workGrid(batchResidualKernel<TR, false, W>, ctx, kTileBlock, batch->tileCount, true), tileBlock, 0,
                           stream, layout, batch->refining, m->entries64, x, b, r, inverse, p, batch->tilePartials,
                           batch->state);
        hipLaunchKernelGGL((batchFinishKernel<TR, kPhaseRefine>), systems, tileBlock, 0, stream, batch->systemTiles,
                           batch->tilePartials, batch->state, settings, batch->active);
    } else {
        hipLaunchKernelGGL((batchFinishKernel<TR, kPhaseUpdate>), systems, tileBlock, 0, stream, batch->systemChunks,
                           batch->chunkPartials, batch->state, settings, batch->active);
    }
    hipLaunchKernelGGL((batchDirectionKernel<TR>), workGrid(batchDirectionKernel<TR>, ctx, kChunkThreads, batch->chunkCount),
                       chunkBlock, 0, stream, layout, work, p, r, inverse, batch->state);
    launchSelect<kRunning>(batch, work);
}

// One trip with the batch's walker.
template <class TR>
void enqueueTrip(cie_hip_batch batch, const BatchLayout& layout, const double* b, double* x, TR* r, TR* p, TR* q,
                 float* correction, const TR* inverse, const BatchSettings& settings, bool mixed,
                 ProductProfile* profile) {
    withWalkerCode(walkerCode(batch), [&](auto w) {
        enqueueIteration<TR, decltype(w)::value>(batch, layout, b, x, r, p, q, correction, inverse, settings, mixed,
                                                 profile);
    });
}

template <class TR>
int runBatch(cie_hip_batch batch, const double* b, double* x, const BatchSettings& settings, bool mixed,
             bool useGraph, double* seconds, int* launches, ProductProfile* profile) {
    cie_hip_context ctx = batch->context;
    cie_hip_csr m = batch->matrix;
    const int64_t n = m->rows;
    int status = ensureInverseDiagonal(m, mixed ? 32 : 64);
    if (status) return status;
    BatchLayout layout{};
    status = batchLayout(batch, &layout);
    if (status) return status;
    const TR* inverse;
    if constexpr (std::is_same_v<TR, float>) inverse = m->inverseDiagonal32;
    else inverse = m->inverseDiagonal64;
    const size_t vector = (sizeof(TR) * size_t(std::max<int64_t>(1, n)) + 255) / 256 * 256;
    const size_t vector32 = (sizeof(float) * size_t(std::max<int64_t>(1, n)) + 255) / 256 * 256;
    const size_t bytes = 3 * vector + (mixed ? vector32 : 0);
    if (batch->workspaceBytes < bytes) {
        freeDevice(batch->workspace);
        CIE_HIP_CHECK(hipMalloc(&batch->workspace, bytes));
        batch->workspaceBytes = bytes;
        releaseGraph(batch);
    }
    TR* r = reinterpret_cast<TR*>(batch->workspace);
    TR* p = reinterpret_cast<TR*>(static_cast<char*>(batch->workspace) + vector);
    TR* q = reinterpret_cast<TR*>(static_cast<char*>(batch->workspace) + 2 * vector);
    float* correction = mixed ? reinterpret_cast<float*>(static_cast<char*>(batch->workspace) + 3 * vector) : nullptr;

    // Initial states and the active-system counter (uploaded, not memset: the
    // graph below never records a memset node).
    std::vector<SystemState> init(size_t(batch->systems));
    CIE_HIP_CHECK(hipMemcpyAsync(batch->state, init.data(), sizeof(SystemState) * init.size(), hipMemcpyHostToDevice,
                                 ctx->stream));
    *batch->pinnedActive = batch->systems;
    CIE_HIP_CHECK(hipMemcpyAsync(batch->active, batch->pinnedActive, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));

    ScopedEvents<3> events;  // the start, and one mark per polling slot
    CIE_HIP_CHECK(events.create());
    const hipEvent_t start = events.event[0];
    hipEvent_t* const marks = events.event + 1;
    CIE_HIP_CHECK(hipEventRecord(start, ctx->stream));
    const ActiveWork& work = batch->running;
    if (correction)  // a fill kernel, as everywhere in the batch's stream (no memset nodes near its graph)
        hipLaunchKernelGGL((vectorKernel<float>), dim3(vectorGrid(n)), dim3(kReduceBlock), 0, ctx->stream, int(kFill),
                           n, correction, static_cast<const float*>(nullptr), 0.0f);
    launchSelect<kRunning>(batch, work);  // every system runs
    withWalkerCode(walkerCode(batch), [&](auto w) {
        constexpr int W = decltype(w)::value;
        hipLaunchKernelGGL((batchResidualKernel<TR, true, W>),
                           workGrid(batchResidualKernel<TR, true, W>, ctx, kTileBlock, batch->tileCount),
                           dim3(kTileBlock), 0, ctx->stream, layout, work, m->entries64, x, b, r, inverse, p,
                           batch->tilePartials, batch->state);
    });
    hipLaunchKernelGGL((batchFinishKernel<TR, kPhaseStart>), dim3(batch->systems), dim3(kTileBlock), 0, ctx->stream,
                       batch->systemTiles, batch->tilePartials, batch->state, settings, batch->active);
    launchSelect<kRunning>(batch, work);
    CIE_HIP_CHECK(hipGetLastError());

    if (useGraph) {
        const int key[5] = {mixed ? 1 : 0, settings.restart, int(sizeof(TR)), batch->tileCount, walkerCode(batch)};
        // The finish kernels take the settings by value, so a graph captured
        // with other tolerances, caps or refinement threshold is not reused.
        const BatchSettings& g = batch->graphSettings;
        const bool sameSettings = g.absoluteTolerance == settings.absoluteTolerance &&
                                  g.relativeTolerance == settings.relativeTolerance &&
                                  g.updateThreshold == settings.updateThreshold &&
                                  g.maxIterations == settings.maxIterations && g.restart == settings.restart &&
                                  g.mixed == settings.mixed && g.rowSumCorrection == settings.rowSumCorrection;
        const bool same = batch->graph && sameSettings && std::equal(key, key + 5, batch->graphKey) &&
                          batch->graphPointers[0] == b && batch->graphPointers[1] == x &&
                          batch->graphPointers[2] == inverse;
        if (!same) {
            releaseGraph(batch);
            hipGraph_t graph = nullptr;
            CIE_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < kGraphIterations; ++i)
                enqueueTrip<TR>(batch, layout, b, x, r, p, q, correction, inverse, settings, mixed, nullptr);
            const hipError_t captured = hipStreamEndCapture(ctx->stream, &graph);
            if (captured != hipSuccess) {
                return setError(CIE_HIP_ERROR_HIP, std::string("batch graph capture: ") + hipGetErrorString(captured));
            }
            const hipError_t made = hipGraphInstantiate(&batch->graph, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (made != hipSuccess) {
                batch->graph = nullptr;
                return setError(CIE_HIP_ERROR_HIP, std::string("batch graph instantiate: ") + hipGetErrorString(made));
            }
            std::copy(key, key + 5, batch->graphKey);
            batch->graphSettings = settings;
            batch->graphPointers[0] = b;
            batch->graphPointers[1] = x;
            batch->graphPointers[2] = inverse;
        }
    }

    int slot = 0;
    bool pending = false;
    int count = 0;
    for (;;) {
        if (useGraph) {
            CIE_HIP_CHECK(hipGraphLaunch(batch->graph, ctx->stream));
        } else {
            for (int i = 0; i < kGraphIterations; ++i)
                enqueueTrip<TR>(batch, layout, b, x, r, p, q, correction, inverse, settings, mixed, profile);
            CIE_HIP_CHECK(hipGetLastError());
        }
        ++count;
        CIE_HIP_CHECK(hipMemcpyAsync(batch->pinnedActive + 1 + slot, batch->active, sizeof(int), hipMemcpyDeviceToHost,
                                     ctx->stream));
        CIE_HIP_CHECK(hipEventRecord(marks[slot], ctx->stream));
        if (pending) {
            const int previous = slot ^ 1;
            CIE_HIP_CHECK(hipEventSynchronize(marks[previous]));
            if (batch->pinnedActive[1 + previous] <= 0) {
                float ms = 0.0f;
                CIE_HIP_CHECK(hipEventElapsedTime(&ms, start, marks[previous]));
                *seconds = ms * 1e-3;
                break;
            }
        }
        pending = true;
        slot ^= 1;
    }
    CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    *launches = count;
    return CIE_HIP_OK;
}

}  // namespace

extern "C" {

int cie_hip_batch_create(cie_hip_context context, cie_hip_csr matrix, int32_t systems, const int64_t* offsets,
                         cie_hip_batch* batch) {
    if (!batch) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_batch_create: null output");
    *batch = nullptr;
    if (!context || !matrix || systems < 1 || !offsets)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_batch_create: invalid arguments");
    CIE_HIP_REQUIRE_CONTEXT(context, matrix->context, "cie_hip_batch_create");
    if (matrix->rows != matrix->columns || offsets[0] != 0 || offsets[systems] != matrix->rows)
        return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_batch_create: offsets must cover the square matrix's rows");
    for (int s = 0; s < systems; ++s)
        if (offsets[s + 1] <= offsets[s])
            return setError(CIE_HIP_ERROR_SHAPE, "cie_hip_batch_create: system " + std::to_string(s) + " is empty");
    if (!matrix->entries64 || !matrix->entries32)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_batch_create: the matrix needs both value planes");
    CIE_HIP_CHECK(hipSetDevice(context->device));
    auto* b = new cie_hip_batch_s;
    b->context = context;
    b->matrix = matrix;
    holdContext(context);
    holdMatrix(matrix);
    b->systems = systems;
    b->offsets.assign(offsets, offsets + systems + 1);
    b->tileRows.assign(size_t(systems), kDefaultTileRows);
    auto bail = [&](int status, const std::string& message) {
        releaseBatch(b);
        return setError(status, message);
    };
    int64_t* deviceOffsets = nullptr;
    const size_t perSystem = sizeof(int32_t) * (size_t(systems) + 1);
    bool lists = true;
    for (ActiveWork* work : {&b->running, &b->refining})
        lists = lists && hipMalloc(&work->systems, perSystem) == hipSuccess &&
                hipMalloc(&work->tilePrefix, perSystem) == hipSuccess &&
                hipMalloc(&work->chunkPrefix, perSystem) == hipSuccess &&
                hipMalloc(&work->count, sizeof(int32_t) * 4) == hipSuccess;
    if (!lists || hipMalloc(&b->systemTiles, perSystem) != hipSuccess ||
        hipMalloc(&b->systemChunks, perSystem) != hipSuccess ||
        hipMalloc(&b->state, sizeof(SystemState) * systems) != hipSuccess ||
        hipMalloc(&b->active, sizeof(int) * 2) != hipSuccess ||
        hipHostMalloc(&b->pinnedActive, sizeof(int) * 4) != hipSuccess ||
        hipHostMalloc(&b->pinnedStates, sizeof(SystemState) * systems) != hipSuccess ||
        hipMalloc(&deviceOffsets, sizeof(int64_t) * (systems + 1)) != hipSuccess)
        return bail(CIE_HIP_ERROR_OUT_OF_MEMORY, "cie_hip_batch_create: allocation failed");
    // Block-diagonal check: every column of system s lies in its row range.
    b->pinnedActive[0] = 0;
    bool ok = hipMemcpy(deviceOffsets, offsets, sizeof(int64_t) * (systems + 1), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(b->active, b->pinnedActive, sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(batchCheckKernel, dim3(systems), dim3(256), 0, context->stream, systems, deviceOffsets,
                           matrix->rowExtents, matrix->columnIndices, b->active);
        ok = hipStreamSynchronize(context->stream) == hipSuccess &&
             hipMemcpy(b->pinnedActive, b->active, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess;
    }
    (void)hipFree(deviceOffsets);
    if (!ok) return bail(CIE_HIP_ERROR_HIP, "cie_hip_batch_create: block check failed to run");
    if (b->pinnedActive[0]) return bail(CIE_HIP_ERROR_SHAPE, "cie_hip_batch_create: the matrix is not block diagonal");
    const int status = buildWork(b);
    if (status) return bail(status, cie_hip_last_error());
    *batch = b;
    return CIE_HIP_OK;
}

int cie_hip_batch_destroy(cie_hip_batch batch) {
    releaseBatch(batch);
    return CIE_HIP_OK;
}

int cie_hip_batch_set_tile_rows(cie_hip_batch batch, const int32_t* rowsPerTile) {
    if (!batch || !rowsPerTile) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_batch_set_tile_rows: null argument");
    for (int s = 0; s < batch->systems; ++s)
        if (rowsPerTile[s] != 32 && rowsPerTile[s] != 64 && rowsPerTile[s] != 128 && rowsPerTile[s] != 256)
            return setError(CIE_HIP_ERROR_INVALID_ARGUMENT,
                            "cie_hip_batch_set_tile_rows: rows per tile must be 32, 64, 128 or 256");
    batch->tileRows.assign(rowsPerTile, rowsPerTile + batch->systems);
    return buildWork(batch);
}

int cie_hip_batch_get_tile_rows(cie_hip_batch batch, int32_t* rowsPerTile) {
    if (!batch || !rowsPerTile) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_batch_get_tile_rows: null argument");
    std::copy(batch->tileRows.begin(), batch->tileRows.end(), rowsPerTile);
    return CIE_HIP_OK;
}

int cie_hip_batch_autotune(cie_hip_batch batch, int precision, int repetitions) {
    if (!batch || (precision != 32 && precision != 64))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_batch_autotune: invalid arguments");
    cie_hip_context ctx = batch->context;
    cie_hip_csr m = batch->matrix;
    CIE_HIP_CHECK(hipSetDevice(ctx->device));
    repetitions = std::max(1, repetitions);
    const size_t valueSize = precision == 32 ? sizeof(float) : sizeof(double);
    void* buffer = nullptr;
    CIE_HIP_CHECK(hipMalloc(&buffer, 2 * valueSize * size_t(std::max<int64_t>(1, m->rows))));
    ScopedEvents<2> events;
    if (const hipError_t error = events.create(); error != hipSuccess ||
        hipMemsetAsync(buffer, 0, 2 * valueSize * size_t(std::max<int64_t>(1, m->rows)), ctx->stream) != hipSuccess) {
        (void)hipFree(buffer);
        return setError(CIE_HIP_ERROR_HIP, "cie_hip_batch_autotune: timing set-up failed");
    }
    const hipEvent_t start = events.event[0], stop = events.event[1];
    // Formats (walker codes): tiles of entries, and tiles of 1 x 3 column blocks, plain or column-sorted, where
    // their padding stays small (fill <= 5/3).
    std::vector<int> codes{0};
    if (m->columns >= kBatchBlockWidth) {
        cie_hip_csr form = nullptr;
        const int status = ensureColumnBlocks(m, kBatchBlockWidth, &form);
        if (status != CIE_HIP_OK) {
            (void)hipFree(buffer);
            return status;
        }
        if (3 * form->nnz * kBatchBlockWidth <= 5 * std::max<int64_t>(1, m->nnz)) {
            codes.push_back(kBatchBlockWidth);
            codes.push_back(-kBatchBlockWidth);
        }
    }
    if (m->columns >= kBatchPaired / 10) codes.push_back(kBatchPaired);
    // The time of `tiles` tiles from `t0` on (one launch, the best of `repetitions`), or a negative status.
    const auto timeTiles = [&](int code, const BatchLayout& layout, int t0, int tiles, float* best) -> int {
        *best = std::numeric_limits<float>::infinity();
        for (int pass = 0; pass < 2; ++pass) {
            (void)hipEventRecord(start, ctx->stream);
            for (int r = 0; r < (pass ? repetitions : 1); ++r)
                launchTileRange(code, precision, tiles, ctx->stream, layout.source, t0, m, buffer);
            if (hipEventRecord(stop, ctx->stream) != hipSuccess || hipEventSynchronize(stop) != hipSuccess)
                return setError(CIE_HIP_ERROR_HIP, "cie_hip_batch_autotune: timing failed");
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, start, stop);
            if (pass) *best = ms / float(repetitions);
        }
        return CIE_HIP_OK;
    };
    // Candidates per format: one cap for every system (32, 64, 128 or 256 rows), and each system's own fastest
    // cap timed alone. A small system alone runs fastest on small tiles (more blocks for the GPU), but the batch
    // runs every system at once, where fewer, fuller tiles win; so every candidate is judged by the time of the
    // whole batch's product in one launch.
    double bestTotal = std::numeric_limits<double>::infinity();
    int bestCode = 0;
    std::vector<int32_t> bestChoice(size_t(batch->systems), kDefaultTileRows);
    int status = CIE_HIP_OK;
    for (int code : codes) {
        status = setTileFormat(batch, code);
        if (status == CIE_HIP_ERROR_INVALID_ARGUMENT && code == kBatchPaired) {  // does not fit the paired form
            status = CIE_HIP_OK;
            continue;
        }
        if (status) break;
        std::vector<float> best(size_t(batch->systems), std::numeric_limits<float>::infinity());
        std::vector<int32_t> choice(size_t(batch->systems), kDefaultTileRows);
        std::vector<std::vector<int32_t>> candidates;
        bool skipped = false;
        const auto judge = [&](const std::vector<int32_t>& rows, bool perSystem) -> int {
            batch->tileRows = rows;
            int result = buildWork(batch);
            if (result && code == kBatchPaired && batch->paired && batch->paired->state < 0) {  // does not fit
                skipped = true;
                return CIE_HIP_OK;
            }
            if (result) return result;
            BatchLayout layout{};
            result = batchLayout(batch, &layout);
            if (result && code < 0 && batch->sorted.state < 0) {  // tiles too wide for the sorted keys
                skipped = true;
                return CIE_HIP_OK;
            }
            if (result) return result;
            float whole = 0.0f;
            result = timeTiles(code, layout, 0, batch->tileCount, &whole);
            if (result) return result;
            if (whole < bestTotal) {
                bestTotal = whole;
                bestCode = code;
                bestChoice = rows;
            }
            if (!perSystem) return CIE_HIP_OK;
            std::vector<int32_t> systemTiles(size_t(batch->systems) + 1);
            if (hipMemcpy(systemTiles.data(), batch->systemTiles, sizeof(int32_t) * systemTiles.size(),
                          hipMemcpyDeviceToHost) != hipSuccess)
                return setError(CIE_HIP_ERROR_HIP, "cie_hip_batch_autotune: reading the tile layout failed");
            for (int sys = 0; sys < batch->systems; ++sys) {
                float ms = 0.0f;
                result = timeTiles(code, layout, systemTiles[sys], systemTiles[sys + 1] - systemTiles[sys], &ms);
                if (result) return result;
                if (ms < best[sys]) {
                    best[sys] = ms;
                    choice[sys] = rows[0];
                }
            }
            return CIE_HIP_OK;
        };
        for (int rows : {32, 64, 128, 256}) {
            status = judge(std::vector<int32_t>(size_t(batch->systems), rows), true);
            if (status || skipped) break;
        }
        if (status) break;
        if (skipped) continue;
        status = judge(choice, false);
        if (status) break;
    }
    (void)hipFree(buffer);
    if (status) return status;
    batch->tileRows = bestChoice;
    status = setTileFormat(batch, bestCode);
    if (status) return status;
    return buildWork(batch);
}

int cie_hip_batch_get_tile_kernel(cie_hip_batch batch, int* kernel) {
    if (!batch || !kernel) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_batch_get_tile_kernel: null argument");
    *kernel = batch->pairedTiles ? CIE_HIP_SPMV_PAIRED
              : !batch->blockWidth ? CIE_HIP_SPMV_TILED
              : batch->sortedTiles ? CIE_HIP_SPMV_SORTED
                                   : CIE_HIP_SPMV_BLOCKS;
    return CIE_HIP_OK;
}

int cie_hip_batch_set_tile_kernel(cie_hip_batch batch, int kernel) {
    if (!batch || (kernel != CIE_HIP_SPMV_TILED && kernel != CIE_HIP_SPMV_BLOCKS && kernel != CIE_HIP_SPMV_SORTED &&
                   kernel != CIE_HIP_SPMV_PAIRED))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT,
                        "cie_hip_batch_set_tile_kernel: kernel must be CIE_HIP_SPMV_TILED, _BLOCKS, _SORTED or _PAIRED");
    const int needed = kernel == CIE_HIP_SPMV_PAIRED ? kBatchPaired / 10 : kBatchBlockWidth;
    if (kernel != CIE_HIP_SPMV_TILED && batch->matrix->columns < needed)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_batch_set_tile_kernel: fewer columns than a block");
    CIE_HIP_CHECK(hipSetDevice(batch->context->device));
    if (kernel == CIE_HIP_SPMV_PAIRED) return setPairedTiles(batch);
    return setBlockWidth(batch, kernel == CIE_HIP_SPMV_TILED ? 0 : kBatchBlockWidth, kernel == CIE_HIP_SPMV_SORTED);
}

int cie_hip_batch_get_column_blocks(cie_hip_batch batch, int* width) {
    if (!batch || !width) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_batch_get_column_blocks: null argument");
    *width = batch->blockWidth;
    return CIE_HIP_OK;
}

int cie_hip_batch_set_column_blocks(cie_hip_batch batch, int width) {
    if (!batch || (width != 0 && width != kBatchBlockWidth))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_batch_set_column_blocks: width must be 0 or " +
                                                            std::to_string(kBatchBlockWidth));
    CIE_HIP_CHECK(hipSetDevice(batch->context->device));
    return setBlockWidth(batch, width, false);
}

int cie_hip_pcg_batched(cie_hip_context context, cie_hip_batch batch, const double* b, double* x,
                        const cie_hip_mixed_settings* settings, int mixed, int useGraph, cie_hip_statistics* stats,
                        int32_t* status, uint64_t* refinements, cie_hip_solve_info* summary) {
    if (!context || !batch || !settings || (!b || !x))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_pcg_batched: invalid arguments");
    if (context != batch->context)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_pcg_batched: batch belongs to another context");
    CIE_HIP_CHECK(hipSetDevice(context->device));
    BatchSettings s{};
    s.absoluteTolerance = settings->settings.absoluteResidual;
    s.relativeTolerance = settings->settings.relativeResidual;
    double factor = settings->refinementFactor;
    if (!(factor > 0.0 && factor < 1.0)) factor = CIE_HIP_DEFAULT_REFINEMENT_FACTOR;
    s.updateThreshold = factor * factor;
    s.maxIterations = settings->settings.iterationCount;
    s.restart = settings->restart ? 1 : 0;
    s.mixed = mixed ? 1 : 0;
    s.rowSumCorrection = mixed && (settings->flags & CIE_HIP_MIXED_PLAIN_FP32) == 0 ? 1 : 0;
    if (s.maxIterations == 0)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_pcg_batched: iterationCount must be positive");
    double seconds = 0.0;
    int launches = 0;
    ProductProfile profile;  // product kernels bracketed with events (profiling on, no graph)
    const int run = mixed ? runBatch<float>(batch, b, x, s, true, useGraph != 0, &seconds, &launches, &profile)
                          : runBatch<double>(batch, b, x, s, false, useGraph != 0, &seconds, &launches, &profile);
    if (run) return run;
    CIE_HIP_CHECK(hipMemcpy(batch->pinnedStates, batch->state, sizeof(SystemState) * batch->systems,
                            hipMemcpyDeviceToHost));
    int converged = 0, worst = CIE_HIP_OK;
    uint64_t maxIterations = 0, totalRefinements = 0;
    double maxRelative = 0.0;
    for (int i = 0; i < batch->systems; ++i) {
        const SystemState& st = batch->pinnedStates[i];
        if (stats) stats[i] = {st.iteration, st.absolute, st.relative};
        const int code = st.done == kNotSpd ? CIE_HIP_ERROR_NOT_SPD
                       : st.done == kBreakdown ? CIE_HIP_ERROR_BREAKDOWN : CIE_HIP_OK;
        if (status) status[i] = st.done;
        if (refinements) refinements[i] = st.refinements;
        converged += st.done == kConverged;
        if (code && !worst) worst = code;
        maxIterations = std::max<uint64_t>(maxIterations, st.iteration);
        totalRefinements += st.refinements;
        maxRelative = std::max(maxRelative, st.relative);
    }
    if (summary) {
        *summary = cie_hip_solve_info{};
        summary->stats = {maxIterations, 0.0, maxRelative};
        summary->converged = converged == batch->systems;
        summary->status = worst;
        summary->seconds = seconds;
        summary->refinements = totalRefinements;
        summary->spmvCount = uint64_t(launches) * kGraphIterations;
        if (context->profiling && profile.used) {
            // The first trips, where every system still runs (the bench checks the fewest iterations).
            const int samples = std::min(profile.used, kProfiledTrips);
            summary->productSamples = uint64_t(samples);
            summary->productSeconds = profile.mean(context, samples);
        }
    }
    if (worst) return setError(worst, "cie_hip_pcg_batched: a system broke down (see the per-system status)");
    return CIE_HIP_OK;
}

}  // extern "C"

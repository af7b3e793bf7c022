// Batched Jacobi-PCG: many independent SPD systems of uneven size solved
// together. No counterpart in the reference (it solves one system per call,
// ConjugateGradients_impl.hpp:49-201); each system follows that algorithm.
//
// Layout: the systems are the diagonal blocks of one CSR matrix (rows of
// system s are [offsets[s], offsets[s+1]) and its columns stay inside that
// range). Rows are cut into tiles of up to 64 rows that never straddle two
// systems; each tile is processed by one wavefront with the staged row sum
// (rows bit-identical to a sequential CPU loop), and writes its share of the
// system's dot products. Every kernel covers all systems in one launch,
// walking the tile list grid-stride; tiles of converged systems are skipped
// (per-system convergence mask). Per-system scalars (rho, alpha, beta,
// residuals, flags) live on the device and are finished by one block per
// system that sums that system's tile partials in tile order. The chain of
// kernels of one iteration is captured once in a HIP graph and replayed until
// every system is done.
//
// Tile height (rows per wavefront: 16, 32 or 64) is chosen per system by
// cie_hip_batch_autotune from timed runs of the system's own tiles.
#include "dispatch.hpp"
#include "kernels.hpp"

#include <cmath>
#include <limits>
#include <vector>

using namespace ciehip;

namespace {

constexpr int kBatchBlock = 256;
constexpr int kBatchWaves = kBatchBlock / kWave;
constexpr int kBatchStage = 1024;
constexpr int kBatchUnroll = 8;
constexpr int kGraphIterations = 8;

struct Tile {
    int32_t first;
    int32_t count;
    int32_t system;
    int32_t pad;
};

struct SystemState {
    double rho, alpha, beta, rr, initialNorm, absolute, relative, updateNorm;
    uint64_t iteration, refinements;
    int32_t done, refine;
};

struct BatchSettings {
    double absoluteTolerance, relativeTolerance, updateThreshold;
    uint64_t maxIterations;
    int32_t restart, mixed;
};

enum Phase : int { kPhaseStart = 0, kPhaseAlpha = 1, kPhaseUpdate = 2, kPhaseUpdateMixed = 3, kPhaseRefine = 4 };

// ---------------------------------------------------------------------------
// Tile kernels (one wavefront per tile, grid-stride over the tile list)
// ---------------------------------------------------------------------------

#define CIEHIP_FOR_EACH_TILE(tileBegin, tileEnd)                                                        \
    const int wave = threadIdx.x / kWave;                                                              \
    const int lane = threadIdx.x % kWave;                                                              \
    for (int t = (tileBegin) + int(blockIdx.x) * kBatchWaves + wave; t < (tileEnd); t += int(gridDim.x) * kBatchWaves)

// r = b - A x (FP64), z = D^-1 r; start: p = z. Tile partials: r.r (FP64), r.z (TR).
template <class TR, bool kStart>
__global__ __launch_bounds__(kBatchBlock) void batchResidualKernel(
    int tileBegin, int tileEnd, const Tile* __restrict__ tiles, const int32_t* __restrict__ rowExtents,
    const int32_t* __restrict__ columnIndices, const double* __restrict__ entries, const double* __restrict__ x,
    const double* __restrict__ b, TR* __restrict__ r, const TR* __restrict__ inverse, TR* __restrict__ p,
    double* __restrict__ partials, const SystemState* __restrict__ state) {
    __shared__ double stage[kBatchWaves][kBatchStage];
    CIEHIP_FOR_EACH_TILE(tileBegin, tileEnd) {
        const Tile tile = tiles[t];
        const SystemState& s = state[tile.system];
        if (s.done || (!kStart && !s.refine)) continue;
        const double sum = stagedTileSum<double, double, kBatchStage, kBatchUnroll>(
            tile.first, tile.count, lane, rowExtents, columnIndices, entries, x, stage[wave]);
        double rr = 0.0;
        TR rz = TR(0);
        if (lane < tile.count) {
            const int row = tile.first + lane;
            const double residual64 = b[row] - sum;
            const TR residual = TR(residual64);
            const TR z = residual * inverse[row];
            r[row] = residual;
            if (kStart) p[row] = z;
            rr = residual64 * residual64;
            rz = residual * z;
        }
        rr = groupSum<kWave>(rr);
        rz = groupSum<kWave>(rz);
        if (lane == 0) {
            partials[2 * t] = rr;
            partials[2 * t + 1] = double(rz);
        }
    }
}

// q = A p; tile partial p.q.
template <class T>
__global__ __launch_bounds__(kBatchBlock) void batchProductKernel(
    int tileBegin, int tileEnd, const Tile* __restrict__ tiles, const int32_t* __restrict__ rowExtents,
    const int32_t* __restrict__ columnIndices, const T* __restrict__ entries, const T* __restrict__ p,
    T* __restrict__ q, double* __restrict__ partials, const SystemState* __restrict__ state) {
    __shared__ T stage[kBatchWaves][kBatchStage];
    CIEHIP_FOR_EACH_TILE(tileBegin, tileEnd) {
        const Tile tile = tiles[t];
        if (state[tile.system].done) continue;
        const T sum = stagedTileSum<T, T, kBatchStage, kBatchUnroll>(tile.first, tile.count, lane, rowExtents,
                                                                    columnIndices, entries, p, stage[wave]);
        T pq = T(0);
        if (lane < tile.count) {
            const int row = tile.first + lane;
            q[row] = sum;
            pq = p[row] * sum;
        }
        pq = groupSum<kWave>(pq);
        if (lane == 0) partials[2 * t] = double(pq);
    }
}

// x += alpha p; r -= alpha q; z = D^-1 r; tile partials r.r and r.z.
template <class T, class TX>
__global__ __launch_bounds__(kBatchBlock) void batchUpdateKernel(
    int tileBegin, int tileEnd, const Tile* __restrict__ tiles, TX* __restrict__ x, T* __restrict__ r,
    const T* __restrict__ p, const T* __restrict__ q, const T* __restrict__ inverse, double* __restrict__ partials,
    const SystemState* __restrict__ state) {
    CIEHIP_FOR_EACH_TILE(tileBegin, tileEnd) {
        const Tile tile = tiles[t];
        const SystemState& s = state[tile.system];
        if (s.done) continue;
        const T alpha = T(s.alpha);
        T rr = T(0), rz = T(0);
        if (lane < tile.count) {
            const int row = tile.first + lane;
            x[row] = x[row] + TX(alpha) * TX(p[row]);
            const T residual = r[row] + (-alpha) * q[row];
            r[row] = residual;
            const T z = residual * inverse[row];
            rr = residual * residual;
            rz = residual * z;
        }
        rr = groupSum<kWave>(rr);
        rz = groupSum<kWave>(rz);
        if (lane == 0) {
            partials[2 * t] = double(rr);
            partials[2 * t + 1] = double(rz);
        }
    }
}

// p = beta p + D^-1 r
template <class T>
__global__ __launch_bounds__(kBatchBlock) void batchDirectionKernel(int tileBegin, int tileEnd,
                                                                    const Tile* __restrict__ tiles, T* __restrict__ p,
                                                                    const T* __restrict__ r,
                                                                    const T* __restrict__ inverse,
                                                                    const SystemState* __restrict__ state) {
    CIEHIP_FOR_EACH_TILE(tileBegin, tileEnd) {
        const Tile tile = tiles[t];
        const SystemState& s = state[tile.system];
        if (s.done) continue;
        const T beta = T(s.beta);
        if (lane < tile.count) {
            const int row = tile.first + lane;
            p[row] = p[row] * beta + r[row] * inverse[row];
        }
    }
}

// ---------------------------------------------------------------------------
// Per-system finish: one block per system sums its tile partials in order.
// ---------------------------------------------------------------------------

template <class TR, int kPhase>
__global__ __launch_bounds__(kBatchBlock) void batchFinishKernel(const int32_t* __restrict__ systemTiles,
                                                                 const double* __restrict__ partials,
                                                                 SystemState* __restrict__ state,
                                                                 BatchSettings settings, int* active) {
    __shared__ double lds64[kBatchBlock / kWave];
    __shared__ TR ldsR[kBatchBlock / kWave];
    SystemState& s = state[blockIdx.x];
    if (s.done) return;
    if (kPhase == kPhaseRefine && !s.refine) return;
    const int begin = systemTiles[blockIdx.x], end = systemTiles[blockIdx.x + 1];
    // first value: r.r (FP64 in the residual phases, TR otherwise) or p.q; second: r.z (TR)
    double first64 = 0.0;
    TR firstR = TR(0), second = TR(0);
    for (int t = begin + threadIdx.x; t < end; t += kBatchBlock) {
        first64 += partials[2 * t];
        firstR += TR(partials[2 * t]);
        second += TR(partials[2 * t + 1]);
    }
    constexpr bool kFirst64 = kPhase == kPhaseStart || kPhase == kPhaseRefine;
    double A64 = 0.0;
    TR AR = TR(0), B = TR(0);
    if (kFirst64) A64 = blockSum<kBatchBlock>(first64, lds64);
    else AR = blockSum<kBatchBlock>(firstR, ldsR);
    if (kPhase != kPhaseAlpha) B = blockSum<kBatchBlock>(second, ldsR);
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
        if (!(TR(0) < AR)) {
            s.done = kNotSpd;
            finished = true;
        } else {
            s.alpha = double(TR(s.rho) / AR);
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
        const bool refine = double(AR) <= settings.updateThreshold * s.updateNorm ||
                            floatLess(absolute, settings.absoluteTolerance) ||
                            floatLess(relative, settings.relativeTolerance) || s.iteration + 1 >= settings.maxIterations;
        if (refine) {
            s.refine = 1;
        } else {
            const TR previous = TR(s.rho);
            s.rho = double(B);
            s.beta = double(B / previous);
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
            s.beta = settings.restart ? 0.0 : double(B / previous);
            s.updateNorm = A64;
            s.iteration += 1;
            if (s.iteration >= settings.maxIterations) s.done = kMaxIterations;
        }
        finished = s.done != kRunning;
    }
    if (finished) atomicSub(active, 1);
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
    Tile* tiles = nullptr;
    int32_t* systemTiles = nullptr;     // device, systems + 1
    int tileCount = 0;
    SystemState* state = nullptr;
    double* partials = nullptr;
    int* active = nullptr;
    int* pinnedActive = nullptr;
    SystemState* pinnedStates = nullptr;
    void* workspace = nullptr;
    size_t workspaceBytes = 0;
    hipGraphExec_t graph = nullptr;
    int graphKey[4] = {0, 0, 0, 0};     // what the captured graph was built for
    const void* graphPointers[3] = {nullptr, nullptr, nullptr};  // b, x, inverse diagonal
};

namespace {

void releaseGraph(cie_hip_batch batch) {
    if (batch->graph) (void)hipGraphExecDestroy(batch->graph);
    batch->graph = nullptr;
}

int buildTiles(cie_hip_batch batch) {
    std::vector<Tile> tiles;
    std::vector<int32_t> systemTiles(size_t(batch->systems) + 1, 0);
    for (int s = 0; s < batch->systems; ++s) {
        systemTiles[s] = int32_t(tiles.size());
        const int rows = batch->tileRows[s];
        for (int64_t first = batch->offsets[s]; first < batch->offsets[s + 1]; first += rows)
            tiles.push_back({int32_t(first), int32_t(std::min<int64_t>(rows, batch->offsets[s + 1] - first)), s, 0});
    }
    systemTiles[batch->systems] = int32_t(tiles.size());
    if (batch->tiles) CIE_HIP_CHECK(hipFree(batch->tiles));
    if (batch->partials) CIE_HIP_CHECK(hipFree(batch->partials));
    batch->tiles = nullptr;
    batch->partials = nullptr;
    batch->tileCount = int(tiles.size());
    CIE_HIP_CHECK(hipMalloc(&batch->tiles, sizeof(Tile) * std::max<size_t>(1, tiles.size())));
    CIE_HIP_CHECK(hipMalloc(&batch->partials, 2 * sizeof(double) * std::max<size_t>(1, tiles.size())));
    if (!tiles.empty())
        CIE_HIP_CHECK(hipMemcpy(batch->tiles, tiles.data(), sizeof(Tile) * tiles.size(), hipMemcpyHostToDevice));
    CIE_HIP_CHECK(hipMemcpy(batch->systemTiles, systemTiles.data(), sizeof(int32_t) * systemTiles.size(),
                            hipMemcpyHostToDevice));
    releaseGraph(batch);
    return CIE_HIP_OK;
}

// One resident round of blocks of `kernel` (grid-stride covers the rest).
template <class Kernel>
dim3 tileGrid(Kernel kernel, cie_hip_context ctx, int tiles) {
    const int wanted = (tiles + kBatchWaves - 1) / kBatchWaves;
    const int resident = std::max(1, ctx->computeUnits) *
                         residentBlocks(reinterpret_cast<const void*>(kernel), kBatchBlock);
    return dim3(std::max(1, std::min(wanted, resident)));
}

void releaseBatch(cie_hip_batch batch) {
    if (!batch) return;
    if (batch->context) {
        (void)hipSetDevice(batch->context->device);
        (void)hipStreamSynchronize(batch->context->stream);
    }
    releaseGraph(batch);
    if (batch->tiles) (void)hipFree(batch->tiles);
    if (batch->systemTiles) (void)hipFree(batch->systemTiles);
    if (batch->state) (void)hipFree(batch->state);
    if (batch->partials) (void)hipFree(batch->partials);
    if (batch->active) (void)hipFree(batch->active);
    if (batch->pinnedActive) (void)hipHostFree(batch->pinnedActive);
    if (batch->pinnedStates) (void)hipHostFree(batch->pinnedStates);
    if (batch->workspace) (void)hipFree(batch->workspace);
    delete batch;
}

// Enqueues the kernels of one iteration on the context stream.
template <class TR>
void enqueueIteration(cie_hip_batch batch, const double* b, double* x, TR* r, TR* p, TR* q, const TR* inverse,
                      const BatchSettings& settings, bool mixed) {
    cie_hip_context ctx = batch->context;
    cie_hip_csr m = batch->matrix;
    const hipStream_t stream = ctx->stream;
    const int tiles = batch->tileCount;
    const dim3 block(kBatchBlock), systems(batch->systems);
    const TR* entries;
    if constexpr (std::is_same_v<TR, float>) entries = m->entries32;
    else entries = m->entries64;
    hipLaunchKernelGGL((batchProductKernel<TR>), tileGrid(batchProductKernel<TR>, ctx, tiles), block, 0, stream, 0,
                       tiles, batch->tiles, m->rowExtents, m->columnIndices, entries, p, q, batch->partials,
                       batch->state);
    hipLaunchKernelGGL((batchFinishKernel<TR, kPhaseAlpha>), systems, block, 0, stream, batch->systemTiles,
                       batch->partials, batch->state, settings, batch->active);
    hipLaunchKernelGGL((batchUpdateKernel<TR, double>), tileGrid(batchUpdateKernel<TR, double>, ctx, tiles), block, 0,
                       stream, 0, tiles, batch->tiles, x, r, p, q, inverse, batch->partials, batch->state);
    if (mixed) {
        hipLaunchKernelGGL((batchFinishKernel<TR, kPhaseUpdateMixed>), systems, block, 0, stream, batch->systemTiles,
                           batch->partials, batch->state, settings, batch->active);
        hipLaunchKernelGGL((batchResidualKernel<TR, false>), tileGrid(batchResidualKernel<TR, false>, ctx, tiles),
                           block, 0, stream, 0, batch->tileCount, batch->tiles,
                           m->rowExtents, m->columnIndices, m->entries64, x, b, r, inverse, p, batch->partials,
                           batch->state);
        hipLaunchKernelGGL((batchFinishKernel<TR, kPhaseRefine>), systems, block, 0, stream, batch->systemTiles,
                           batch->partials, batch->state, settings, batch->active);
    } else {
        hipLaunchKernelGGL((batchFinishKernel<TR, kPhaseUpdate>), systems, block, 0, stream, batch->systemTiles,
                           batch->partials, batch->state, settings, batch->active);
    }
    hipLaunchKernelGGL((batchDirectionKernel<TR>), tileGrid(batchDirectionKernel<TR>, ctx, tiles), block, 0, stream, 0, batch->tileCount, batch->tiles, p, r,
                       inverse, batch->state);
}

template <class TR>
int runBatch(cie_hip_batch batch, const double* b, double* x, const BatchSettings& settings, bool mixed,
             bool useGraph, double* seconds, int* launches) {
    cie_hip_context ctx = batch->context;
    cie_hip_csr m = batch->matrix;
    const int64_t n = m->rows;
    int status = ensureInverseDiagonal(m, mixed ? 32 : 64);
    if (status) return status;
    const TR* inverse;
    if constexpr (std::is_same_v<TR, float>) inverse = m->inverseDiagonal32;
    else inverse = m->inverseDiagonal64;
    const size_t vector = (sizeof(TR) * size_t(std::max<int64_t>(1, n)) + 255) / 256 * 256;
    if (batch->workspaceBytes < 3 * vector) {
        if (batch->workspace) CIE_HIP_CHECK(hipFree(batch->workspace));
        batch->workspace = nullptr;
        CIE_HIP_CHECK(hipMalloc(&batch->workspace, 3 * vector));
        batch->workspaceBytes = 3 * vector;
        releaseGraph(batch);
    }
    TR* r = reinterpret_cast<TR*>(batch->workspace);
    TR* p = reinterpret_cast<TR*>(static_cast<char*>(batch->workspace) + vector);
    TR* q = reinterpret_cast<TR*>(static_cast<char*>(batch->workspace) + 2 * vector);

    // Initial states and the active-system counter (uploaded, not memset: the
    // graph below never records a memset node).
    std::vector<SystemState> init(size_t(batch->systems));
    CIE_HIP_CHECK(hipMemcpyAsync(batch->state, init.data(), sizeof(SystemState) * init.size(), hipMemcpyHostToDevice,
                                 ctx->stream));
    *batch->pinnedActive = batch->systems;
    CIE_HIP_CHECK(hipMemcpyAsync(batch->active, batch->pinnedActive, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));

    hipEvent_t start, marks[2];
    CIE_HIP_CHECK(hipEventCreate(&start));
    CIE_HIP_CHECK(hipEventCreate(&marks[0]));
    CIE_HIP_CHECK(hipEventCreate(&marks[1]));
    auto cleanup = [&]() {
        (void)hipEventDestroy(start);
        (void)hipEventDestroy(marks[0]);
        (void)hipEventDestroy(marks[1]);
    };
    CIE_HIP_CHECK(hipEventRecord(start, ctx->stream));
    const dim3 grid(tileGrid(batchResidualKernel<TR, true>, ctx, batch->tileCount)), block(kBatchBlock), systems(batch->systems);
    hipLaunchKernelGGL((batchResidualKernel<TR, true>), grid, block, 0, ctx->stream, 0, batch->tileCount, batch->tiles,
                       m->rowExtents, m->columnIndices, m->entries64, x, b, r, inverse, p, batch->partials,
                       batch->state);
    hipLaunchKernelGGL((batchFinishKernel<TR, kPhaseStart>), systems, block, 0, ctx->stream, batch->systemTiles,
                       batch->partials, batch->state, settings, batch->active);
    CIE_HIP_CHECK(hipGetLastError());

    if (useGraph) {
        const int key[4] = {mixed ? 1 : 0, settings.restart, int(sizeof(TR)), batch->tileCount};
        const bool same = batch->graph && std::equal(key, key + 4, batch->graphKey) && batch->graphPointers[0] == b &&
                          batch->graphPointers[1] == x && batch->graphPointers[2] == inverse;
        if (!same) {
            releaseGraph(batch);
            hipGraph_t graph = nullptr;
            CIE_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < kGraphIterations; ++i) enqueueIteration<TR>(batch, b, x, r, p, q, inverse, settings, mixed);
            const hipError_t captured = hipStreamEndCapture(ctx->stream, &graph);
            if (captured != hipSuccess) {
                cleanup();
                return setError(CIE_HIP_ERROR_HIP, std::string("batch graph capture: ") + hipGetErrorString(captured));
            }
            const hipError_t made = hipGraphInstantiate(&batch->graph, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (made != hipSuccess) {
                batch->graph = nullptr;
                cleanup();
                return setError(CIE_HIP_ERROR_HIP, std::string("batch graph instantiate: ") + hipGetErrorString(made));
            }
            std::copy(key, key + 4, batch->graphKey);
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
            for (int i = 0; i < kGraphIterations; ++i) enqueueIteration<TR>(batch, b, x, r, p, q, inverse, settings, mixed);
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
    cleanup();
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
    b->systems = systems;
    b->offsets.assign(offsets, offsets + systems + 1);
    b->tileRows.assign(size_t(systems), kWave);
    auto bail = [&](int status, const std::string& message) {
        releaseBatch(b);
        return setError(status, message);
    };
    int64_t* deviceOffsets = nullptr;
    if (hipMalloc(&b->systemTiles, sizeof(int32_t) * (systems + 1)) != hipSuccess ||
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
    const int status = buildTiles(b);
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
        if (rowsPerTile[s] != 16 && rowsPerTile[s] != 32 && rowsPerTile[s] != 64)
            return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_batch_set_tile_rows: rows per tile must be 16, 32 or 64");
    batch->tileRows.assign(rowsPerTile, rowsPerTile + batch->systems);
    return buildTiles(batch);
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
    std::vector<SystemState> running(size_t(batch->systems));
    CIE_HIP_CHECK(hipMemcpy(batch->state, running.data(), sizeof(SystemState) * running.size(), hipMemcpyHostToDevice));
    CIE_HIP_CHECK(hipMemsetAsync(buffer, 0, 2 * valueSize * size_t(std::max<int64_t>(1, m->rows)), ctx->stream));
    hipEvent_t start, stop;
    CIE_HIP_CHECK(hipEventCreate(&start));
    CIE_HIP_CHECK(hipEventCreate(&stop));
    std::vector<double> best(size_t(batch->systems), std::numeric_limits<double>::infinity());
    std::vector<int32_t> choice(size_t(batch->systems), kWave);
    int status = CIE_HIP_OK;
    for (int rows : {16, 32, 64}) {
        std::vector<int32_t> uniform(size_t(batch->systems), rows);
        batch->tileRows = uniform;
        status = buildTiles(batch);
        if (status) break;
        std::vector<int32_t> systemTiles(size_t(batch->systems) + 1);
        CIE_HIP_CHECK(hipMemcpy(systemTiles.data(), batch->systemTiles, sizeof(int32_t) * systemTiles.size(),
                                hipMemcpyDeviceToHost));
        for (int s = 0; s < batch->systems && status == CIE_HIP_OK; ++s) {
            const int t0 = systemTiles[s], t1 = systemTiles[s + 1];
            const dim3 grid(precision == 32 ? tileGrid(batchProductKernel<float>, ctx, t1 - t0) : tileGrid(batchProductKernel<double>, ctx, t1 - t0)), block(kBatchBlock);
            for (int pass = 0; pass < 2; ++pass) {
                (void)hipEventRecord(start, ctx->stream);
                for (int r = 0; r < (pass ? repetitions : 1); ++r) {
                    if (precision == 32) {
                        float* in = static_cast<float*>(buffer);
                        hipLaunchKernelGGL((batchProductKernel<float>), grid, block, 0, ctx->stream, t0, t1,
                                           batch->tiles, m->rowExtents, m->columnIndices, m->entries32, in,
                                           in + m->rows, batch->partials, batch->state);
                    } else {
                        double* in = static_cast<double*>(buffer);
                        hipLaunchKernelGGL((batchProductKernel<double>), grid, block, 0, ctx->stream, t0, t1,
                                           batch->tiles, m->rowExtents, m->columnIndices, m->entries64, in,
                                           in + m->rows, batch->partials, batch->state);
                    }
                }
                if (hipEventRecord(stop, ctx->stream) != hipSuccess || hipEventSynchronize(stop) != hipSuccess) {
                    status = setError(CIE_HIP_ERROR_HIP, "cie_hip_batch_autotune: timing failed");
                    break;
                }
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, start, stop);
                if (pass && ms < best[s]) {
                    best[s] = ms;
                    choice[s] = rows;
                }
            }
        }
    }
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    (void)hipFree(buffer);
    if (status) return status;
    batch->tileRows = choice;
    return buildTiles(batch);
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
    if (!(factor > 0.0 && factor < 1.0)) factor = 0.1;
    s.updateThreshold = factor * factor;
    s.maxIterations = settings->settings.iterationCount;
    s.restart = settings->restart ? 1 : 0;
    s.mixed = mixed ? 1 : 0;
    if (s.maxIterations == 0)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_pcg_batched: iterationCount must be positive");
    double seconds = 0.0;
    int launches = 0;
    const int run = mixed ? runBatch<float>(batch, b, x, s, true, useGraph != 0, &seconds, &launches)
                          : runBatch<double>(batch, b, x, s, false, useGraph != 0, &seconds, &launches);
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
    }
    if (worst) return setError(worst, "cie_hip_pcg_batched: a system broke down (see the per-system status)");
    return CIE_HIP_OK;
}

}  // extern "C"

// Internal definitions shared by the HIP translation units of libcie_hip.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/cie_hip.h"

namespace ciehip {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kReduceBlock = 256;  // threads per block of the vector/reduction kernels
constexpr int kMaxPartials = 4096; // partial slots allocated with a context (grown on demand)
constexpr int kLevelBlocks = 256;  // a reduction over more partials than this gets a first-level sum
constexpr int kVectorPartials = 512;  // partial slots of the solver's update kernel
constexpr int kXcds = 8;           // MI355X accelerator complex dies (blocks are dealt round-robin)

// Error state: the C ABI never throws; each failure sets a thread-local message.
int setError(int status, const std::string& message);

// A matrix, mask or batch belongs to the context it was created on: its
// derived forms, workspace and solver scratch live there, and its work is
// ordered on that context's stream. A call that names another context is
// refused (CIE_HIP_ERROR_INVALID_ARGUMENT).
#define CIE_HIP_REQUIRE_CONTEXT(context, owner, name)                                   \
    do {                                                                                  \
        if ((context) != (owner))                                                         \
            return ::ciehip::setError(CIE_HIP_ERROR_INVALID_ARGUMENT,                     \
                                      std::string(name) + ": an argument belongs to another context"); \
    } while (0)

#define CIE_HIP_CHECK(call)                                                               \
    do {                                                                                  \
        hipError_t cieHipError_ = (call);                                                 \
        if (cieHipError_ != hipSuccess)                                                   \
            return ::ciehip::setError(cieHipError_ == hipErrorOutOfMemory                 \
                                          ? CIE_HIP_ERROR_OUT_OF_MEMORY                   \
                                          : CIE_HIP_ERROR_HIP,                            \
                                      std::string(#call) + ": " + hipGetErrorString(cieHipError_)); \
    } while (0)

// Device-side scalar state of one solve. Lives in device memory; kernels read
// and update it so that a whole PCG iteration needs no host round trip.
struct SolverState {
    double rhos[2];        // r^T z of the current iterate, in the slot of the iteration's parity
    double alpha;          // search scale p^T r / p^T A p
    double beta;           // direction update scale
    double rr;             // r^T r
    double initialNorm;    // ||r_0||
    double absolute;       // ||r||
    double relative;       // ||r|| / ||r_0||
    double absoluteTolerance;
    double relativeTolerance;
    double updateNorm;     // mixed: r^T r at the last FP64 refinement step
    double updateThreshold;// mixed: refine when r^T r <= threshold * updateNorm
    uint64_t iteration;    // Statistics::iterationCount
    uint64_t maxIterations;
    uint64_t refinements;  // mixed: FP64 refinement steps taken
    int32_t done;          // 0 running, else a SolveEnd value
    int32_t refined[2];    // mixed: 1 if the trip of that parity recomputed the true residual
    int32_t restart;       // mixed: 1 resets the search direction at each refinement
    int32_t startExit;     // 1 if the initial residual met the tolerance (no loop trip ran)
    double* history;       // caller's residual history (device), or null: ||r|| of loop trip k in [k]
    uint64_t historyCapacity;
};

// Mixed forms: after a refinement, a replaced (true) residual whose r^T z
// exceeds the recursive one's by more than this factor restarts the search
// direction. The recursion then ran below the FP32 floor of the true
// residual (a system that one FP32 step solves, such as a single row), and
// combining the stale direction with beta = rho_true / rho_recursive can
// grow the residual without bound. At an ordinary refinement the two agree
// closely, so the factor never triggers there.
constexpr double kRestartGap = 4.0;

enum SolveEnd : int32_t {
    kRunning = 0,
    kConverged = 1,
    kMaxIterations = 2,
    kNotSpd = 3,      // p^T A p <= 0
    kBreakdown = 4,   // r^T z <= 0 before the loop
};

// Per-context scratch for grid reductions and scalar results.
struct Scratch {
    double* partials = nullptr;   // partialSlots * 4 doubles (grown on demand)
    int partialSlots = 0;
    double* level = nullptr;      // kLevelBlocks * 2 doubles: first level of a large partial reduction
    double* vectorPartials = nullptr;  // kVectorPartials * 2 doubles: partials of the solver's vector kernels
    double* refinePartials = nullptr;  // kMaxPartials * 2 doubles: FP64 r.r partials of the mixed refinement
    double* scalars = nullptr;    // 64 doubles of device scalar results
    double* pinned = nullptr;     // 64 doubles of pinned host memory for readback
    SolverState* state = nullptr; // one solver state
    SolverState* pinnedState = nullptr; // two pinned host slots (double-buffered polling)
};

}  // namespace ciehip

// Lifetime of the library's objects: a context is held by every object created
// on it, a matrix by the batches, assembly and multigrid plans over it, a mask
// by the multigrid plans over it. Destroying an object that is still held only
// marks it; it is released with its last holder, so objects may be destroyed
// in any order.
struct Holds {
    int holders = 0;
    bool destroyRequested = false;
};

struct cie_hip_context_s {
    Holds holds;
    int device = 0;
    hipStream_t stream = nullptr;
    bool ownsStream = false;
    int computeUnits = 0;
    ciehip::Scratch scratch;
    bool profiling = false;
    static constexpr int kProfileEvents = 64;
    hipEvent_t profileEvents[2 * kProfileEvents] = {};
    double kernelSeconds = 0.0;  // profiling on: the last call's main kernel (cie_hip_context_kernel_seconds)
    double* history = nullptr;   // residual history of the fused solvers (cie_hip_context_set_residual_history)
    int64_t historyCapacity = 0;
};

namespace ciehip {
// With the context's profiling on, times up to kProfileEvents launches of one
// kernel (the product of a solver loop) with a pair of HIP events each,
// recorded by the launch itself at the kernel's start and end (so the launch
// latency before the kernel is not counted); mean() is their average duration
// over the first `samples` of them. Further launches, and all launches with
// profiling off, are plain.
struct ProductProfile {
    int used = 0;
    template <class Kernel, class... Args>
    void launch(cie_hip_context ctx, Kernel kernel, dim3 grid, dim3 block, Args... args) {
        if (ctx->profiling && used < cie_hip_context_s::kProfileEvents) {
            hipExtLaunchKernelGGL(kernel, grid, block, 0, ctx->stream, ctx->profileEvents[2 * used],
                                  ctx->profileEvents[2 * used + 1], 0, args...);
            ++used;
        } else {
            hipLaunchKernelGGL(kernel, grid, block, 0, ctx->stream, args...);
        }
    }
    double mean(cie_hip_context ctx, int samples) const {
        double total = 0.0;
        int count = 0;
        for (int i = 0; i < std::min(used, samples); ++i) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, ctx->profileEvents[2 * i], ctx->profileEvents[2 * i + 1]) == hipSuccess) {
                total += ms * 1e-3;
                ++count;
            }
        }
        return count ? total / count : 0.0;
    }
};

// N events, destroyed on every return path of the scope that holds them.
template <int N>
struct ScopedEvents {
    hipEvent_t event[N] = {};
    ScopedEvents() = default;
    ScopedEvents(const ScopedEvents&) = delete;
    ScopedEvents& operator=(const ScopedEvents&) = delete;
    ~ScopedEvents() {
        for (hipEvent_t e : event)
            if (e) (void)hipEventDestroy(e);
    }
    hipError_t create() {
        for (hipEvent_t& e : event)
            if (const hipError_t error = hipEventCreate(&e); error != hipSuccess) return error;
        return hipSuccess;
    }
};

// With the context's profiling on, brackets one kernel launch on its stream
// with HIP events; finish() waits for the end event and keeps the duration.
struct KernelTimer {
    cie_hip_context context;
    explicit KernelTimer(cie_hip_context c) : context(c) {
        if (context->profiling) (void)hipEventRecord(context->profileEvents[0], context->stream);
    }
    void stop() const {
        if (context->profiling) (void)hipEventRecord(context->profileEvents[1], context->stream);
    }
    void finish() const {
        if (!context->profiling) return;
        float ms = 0.0f;
        if (hipEventSynchronize(context->profileEvents[1]) == hipSuccess &&
            hipEventElapsedTime(&ms, context->profileEvents[0], context->profileEvents[1]) == hipSuccess)
            context->kernelSeconds = double(ms) * 1e-3;
    }
};
}  // namespace ciehip

// Launch configuration of a CSR SpMV for one value precision. The grid is
// sized per launch from the kernel's occupancy (one resident round).
struct CsrLaunch {
    int kernel = 0;   // CIE_HIP_SPMV_VECTOR, CIE_HIP_SPMV_STAGED, CIE_HIP_SPMV_TILED or CIE_HIP_SPMV_BLOCKS
    int width = 0;    // vector: lanes per row; staged: LDS entries per segment; tiled: threads per block;
                      // blocks: columns per block
    int unroll = 0;   // entries per lane loaded before they are used
    int block = 0;    // threads per block
};

// A column-block matrix's blocks in column order within each tile of a tile
// list (see SortedBlockTiles), built on the device on first use.
struct SortedCopy {
    int state = 0;               // 0 not built, 1 built, -1 a tile spans too many columns
    int32_t* bases = nullptr;    // per tile: smallest block column
    uint32_t* keys = nullptr;    // per block: column offset and tile position
    double* values64 = nullptr;  // per block: w values, sorted order
    float* values32 = nullptr;
    uint64_t version = 0;        // the block matrix's filledVersion its values hold
};

// Block tiles of a matrix for one tile capacity (entries per tile).
struct TileSet {
    int capacity = 0;
    int count = 0;          // tiles; the device array holds count + 1 entries
    int2* tiles = nullptr;  // {first row, first entry}
    SortedCopy sorted;      // of a column-block matrix, built on first use by a sorted launch
};

// A matrix covered by 1 x wide and 1 x narrow column blocks, in tiles of
// whole rows with each width's blocks column-sorted within the tile (see
// PairedBlockTiles), for one tile capacity (unroll blocks per lane and
// width). Built on the device on first use; the values are refilled whenever
// the matrix's entries change.
struct PairedForm {
    int wide = 0, narrow = 0, unroll = 0;
    int state = 0;                   // 1 built, -1 the matrix does not fit the form (error says why)
    std::string error;
    int64_t blocks[2] = {0, 0};      // wide, narrow
    int tileCount = 0;
    int32_t* extents = nullptr;      // per row: its first block, both widths in row order (rows + 1)
    int2* tiles = nullptr;           // tileCount + 1: {first row, first block}
    int2* begins = nullptr;          // tileCount + 1: {first wide block, first narrow block}
    int2* bases = nullptr;           // per tile: smallest wide and smallest narrow block column
    uint32_t* keys = nullptr;        // the wide blocks' keys, then the narrow blocks', sorted within tiles
    int32_t* sortedStart = nullptr;  // per sorted block: first column, first CSR entry and entry count
    int32_t* sortedEntry = nullptr;  //   (the refill reads the CSR through them)
    uint8_t* sortedCount = nullptr;
    double* values64 = nullptr;      // wide planes, then narrow planes (PlaneRun layout)
    float* values32 = nullptr;
    uint64_t version = 0;            // the matrix's valuesVersion the values hold
};

struct cie_hip_csr_s {
    Holds holds;
    cie_hip_context context = nullptr;
    int64_t rows = 0, columns = 0, nnz = 0;
    int32_t* rowExtents = nullptr;
    int32_t* columnIndices = nullptr;
    double* entries64 = nullptr;
    float* entries32 = nullptr;
    bool ownsIndices = false, ownsEntries64 = false;
    double* inverseDiagonal64 = nullptr;  // built lazily for the fused solvers
    float* inverseDiagonal32 = nullptr;
    float* rowCorrection32 = nullptr;     // with inverseDiagonal32: sum_j (a_ij - fl32(a_ij)) per row (mixed solvers)
    double meanRowLength = 0.0;
    int maxRowLength = 0;
    CsrLaunch launch64, launch32;
    void* workspace = nullptr;      // solver vectors, grown on demand, freed with the matrix
    size_t workspaceBytes = 0;
    std::vector<int32_t> hostExtents;  // host copy of rowExtents (tile cuts are made on the host)
    std::vector<TileSet> tileSets;
    int narrowState = 0;               // narrow column plane: 0 not built yet, 1 built, -1 does not fit
    int narrowRuns = 0;
    uint16_t* narrowColumns = nullptr;
    int32_t* narrowBases = nullptr;
    uint64_t valuesVersion = 0;        // bumped whenever the entries change
    // The matrix as 1 x w blocks of adjacent columns, per block width w (built
    // on first use): itself a matrix whose "entries" are the blocks (extents,
    // first columns, tiles and narrow plane as above) and whose value planes
    // hold w values per block, zero where the row has no entry.
    static constexpr int kMaxBlockWidth = 4;
    cie_hip_csr_s* columnBlocks[kMaxBlockWidth + 1] = {};
    int blockWidth = 0;                // of a block matrix; 0 for an ordinary one
    int32_t* blockEntryBegin = nullptr;// of a block matrix: each block's first entry of the parent (blocks + 1)
    int blockUsers = 0;                // of a block matrix: holders (batches) that keep it from being freed
    uint64_t filledVersion = 0;        // of a block matrix: the parent's valuesVersion its values hold
    std::vector<PairedForm*> pairedForms;  // built on first use by paired launches
};

namespace ciehip {
// Holds (see Holds): hold on creation of a holder, drop on its release; a
// drop that ends the holds of an object whose destruction was asked for
// releases it.
void holdContext(cie_hip_context context);
void dropContext(cie_hip_context context);
void holdMatrix(cie_hip_csr matrix);
void dropMatrix(cie_hip_csr matrix);
}  // namespace ciehip

namespace ciehip {
// Returns workspace of at least `bytes` bytes owned by the matrix handle.
int matrixWorkspace(cie_hip_csr matrix, size_t bytes, void** pointer);
// Builds the FP64/FP32 inverse diagonal of the matrix if it is missing.
int ensureInverseDiagonal(cie_hip_csr matrix, int precision);
// The block tiles of `capacity` entries (built and uploaded on first use).
int ensureTiles(cie_hip_csr matrix, int capacity, const TileSet** tiles);
// The sorted copy of a column-block matrix over `tiles` (tileCount tiles of at
// most `capacity` blocks; a longer tile holds one row and is walked unsorted):
// keys built on first use, values refreshed whenever the block values changed.
// Fails with CIE_HIP_ERROR_INVALID_ARGUMENT when a tile's block columns span
// more than the keys hold. The caller keeps `copy` with its tile list.
int ensureSortedCopy(cie_hip_csr blocks, const int2* tiles, int tileCount, int capacity, SortedCopy& copy);
// Frees a sorted copy (it is rebuilt on its next use); waits for the stream.
void releaseSortedCopy(cie_hip_context context, SortedCopy& copy);
// The tiles of `capacity` blocks of a column-block matrix with their sorted
// copy (see ensureSortedCopy).
int ensureSortedTiles(cie_hip_csr blocks, int capacity, const TileSet** tiles);
// The paired block form of `matrix` for widths wide > narrow and `unroll`
// blocks per lane and width, values current (built on first use, refilled
// after the entries changed). Fails with CIE_HIP_ERROR_INVALID_ARGUMENT when
// a row holds more blocks of one width than a tile, or a tile's blocks of one
// width span more columns than the keys hold.
int ensurePairedForm(cie_hip_csr matrix, int wide, int narrow, int unroll, const PairedForm** form);
// Frees a paired form's device arrays and the form; waits for the stream.
void releasePairedForm(cie_hip_context context, PairedForm* form);
// How a paired form's tiles are cut: never across the row offsets of
// `systems` systems (null: one system of all rows), at most rowCaps[s] rows
// per tile (null: 256); `systemTiles` (if given) receives the first tile of
// each system and the tile count.
struct PairedCut {
    const int64_t* offsets = nullptr;
    const int32_t* rowCaps = nullptr;
    int systems = 1;
    std::vector<int32_t>* systemTiles = nullptr;
};
// Builds a paired form (form.wide, narrow and unroll set) with its tiles cut
// as asked; a matrix that does not fit leaves form.state -1 and form.error.
int buildPairedForm(cie_hip_csr matrix, PairedForm& form, const PairedCut& cut);
// Refills a built paired form's values if the matrix's entries changed.
int refreshPairedValues(cie_hip_csr matrix, PairedForm& form);
// Builds the narrow (16-bit) column plane on first use if every run of
// entries fits; matrix->narrowState tells whether it did.
int ensureNarrowColumns(cie_hip_csr matrix);
// The 1 x width column-block form of `matrix` with values current (built on
// first use, refilled after the entries changed).
int ensureColumnBlocks(cie_hip_csr matrix, int width, cie_hip_csr* blocks);
// After the entries changed: derived value planes (the Jacobi inverse
// diagonals, the column-block values) are rebuilt before their next use.
int entriesChanged(cie_hip_csr matrix);
// A library-owned matrix over device arrays allocated by the caller of this
// function (ownership passes to the matrix; entries32 may be null).
int adoptMatrix(cie_hip_context context, int64_t rows, int64_t columns, int32_t* rowExtents, int32_t* columnIndices,
                double* entries64, float* entries32, std::vector<int32_t> hostExtents, cie_hip_csr* matrix);
// Grows the context's partials to at least `slots` slots of 4 doubles. A
// growth waits for the stream first (queued kernels may use the old buffer).
int ensurePartials(cie_hip_context context, int slots);
// This device's address of the static scratch of the no-allocation dot
// product (looked up on the first call per device; the device must be current).
int staticScratchAddress(int device, double** address);
}  // namespace ciehip

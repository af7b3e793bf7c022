// Fused Jacobi-preconditioned conjugate gradients on one device.
//
// Algorithm and step order follow ConjugateGradients<TS>::product
// (linalg/packages/solvers/impl/ConjugateGradients_impl.hpp:49-201) with the
// Jacobi DiagonalOperator as preconditioner:
//   r = b - A x; ||r0|| = sqrt(r.r); early exit;  z = D^-1 r; rho = r.z; p = z
//   loop: q = A p; alpha = rho / p.q; x += alpha p; r -= alpha q; rr = r.r;
//         z = D^-1 r; rho' = r.z; beta = rho'/rho; p = beta p + z;
//         abs = sqrt(rr), rel = abs / ||r0||; stop on Comparison::less.
// Each loop trip is three kernels that keep all scalars on the device:
//   product:   q = A p and p.q (last block: alpha, p.q <= 0 check)
//   update:    x, r updates, z, r.r and r.z (last block: beta, convergence)
//   direction: p = beta p + D^-1 r
// The host only polls the device state once per chunk of iterations.
//
// The mixed form keeps A, r, p, q, z and their dot products in FP32 and the
// solution x in FP64. Whenever the FP32 residual has fallen by the refinement
// factor since the last refinement (or looks converged, or the iteration cap
// is near), a refinement step recomputes r = b - A x with the FP64 matrix,
// tests convergence on that FP64 residual, and restarts the FP32 recurrence
// from it; the search direction is kept (restart = 0) or reset (restart = 1).
#include "dispatch.hpp"
#include "kernels.hpp"

#include <cmath>
#include <limits>

using namespace ciehip;

namespace {

enum Counter { kStartCounter = 8, kProductCounter = 9, kUpdateCounter = 10, kRefineCounter = 11 };

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// r = b - A x (FP64), z = D^-1 r, start: p = z. Partials: r.r (FP64), r.z (TR).
template <class TR, int L, int U, int BS, bool kStart>
__global__ __launch_bounds__(BS) void residualKernel(int rows, int rowsPerChunk, const int32_t* __restrict__ rowExtents,
                                                     const int32_t* __restrict__ columnIndices,
                                                     const double* __restrict__ entries, const double* __restrict__ x,
                                                     const double* __restrict__ b, TR* __restrict__ r,
                                                     const TR* __restrict__ inverse, TR* __restrict__ p,
                                                     double* partials, unsigned* counter, SolverState* state) {
    __shared__ double ldsRR[BS / kWave];
    __shared__ TR ldsRZ[BS / kWave];
    if (state->done || (!kStart && !state->refine)) return;
    const int chunk = xcdChunk();
    const RowChunks range = rowChunk(rows, rowsPerChunk, chunk);
    const int group = threadIdx.x / L;
    const int lane = threadIdx.x % L;
    double rr = 0.0;
    TR rz = TR(0);
    for (int row0 = range.rowBegin; row0 < range.rowEnd; row0 += BS / L) {
        const int row = row0 + group;
        if (row < range.rowEnd) {
            const double sum = rowProduct<double, double, L, U>(row, lane, rowExtents, columnIndices, entries, x);
            if (lane == 0) {
                const double t = b[row] - sum;
                const TR residual = TR(t);
                const TR z = residual * inverse[row];
                r[row] = residual;
                if (kStart) p[row] = z;
                rr += t * t;
                rz += residual * z;
            }
        }
    }
    rr = blockSum<BS>(rr, ldsRR);
    rz = blockSum<BS>(rz, ldsRZ);
    if (threadIdx.x == 0) {
        partials[2 * chunk] = rr;
        partials[2 * chunk + 1] = double(rz);
    }
    if (!arriveLast(counter)) return;
    const double RR = sumPartials<BS>(partials, gridDim.x, 2, ldsRR);
    TR rzSum = TR(0);
    for (int i = threadIdx.x; i < int(gridDim.x); i += BS) rzSum += TR(partials[2 * i + 1]);
    const TR RZ = blockSum<BS>(rzSum, ldsRZ);
    if (threadIdx.x != 0) return;
    resetCounter(counter);
    SolverState& s = *state;
    s.rr = RR;
    if (kStart) {
        // ConjugateGradients_impl.hpp:93-132
        const double initial = sqrt(RR);
        s.initialNorm = initial;
        s.absolute = initial;
        s.relative = 1.0;
        s.updateNorm = RR;
        if (floatLess(initial, s.absoluteTolerance) || floatLess(1.0, s.relativeTolerance)) {
            s.done = kConverged;
        } else if (!(TR(0) < RZ)) {
            s.done = kBreakdown;
        } else {
            s.rho = double(RZ);
        }
    } else {
        // FP64 refinement: true residual, FP64 convergence test, restart the FP32 recurrence.
        s.refine = 0;
        s.refinements += 1;
        s.absolute = sqrt(RR);
        s.relative = s.absolute / s.initialNorm;
        if (floatLess(s.absolute, s.absoluteTolerance) || floatLess(s.relative, s.relativeTolerance)) {
            s.done = kConverged;
        } else {
            const TR previous = TR(s.rho);
            s.rho = double(RZ);
            s.beta = s.restart ? 0.0 : double(RZ / previous);
            s.updateNorm = RR;
            s.iteration += 1;
            if (s.iteration >= s.maxIterations) s.done = kMaxIterations;
        }
    }
}

// q = A p, partial p.q; last block: alpha = rho / p.q (ConjugateGradients_impl.hpp:136-151).
template <class T, int L, int U, int BS>
__global__ __launch_bounds__(BS) void productKernel(int rows, int rowsPerChunk, const int32_t* __restrict__ rowExtents,
                                                    const int32_t* __restrict__ columnIndices,
                                                    const T* __restrict__ entries, const T* __restrict__ p,
                                                    T* __restrict__ q, double* partials, unsigned* counter,
                                                    SolverState* state) {
    __shared__ T lds[BS / kWave];
    if (state->done) return;
    const int chunk = xcdChunk();
    const RowChunks range = rowChunk(rows, rowsPerChunk, chunk);
    const int group = threadIdx.x / L;
    const int lane = threadIdx.x % L;
    T pq = T(0);
    for (int row0 = range.rowBegin; row0 < range.rowEnd; row0 += BS / L) {
        const int row = row0 + group;
        if (row < range.rowEnd) {
            const T sum = rowProduct<T, T, L, U>(row, lane, rowExtents, columnIndices, entries, p);
            if (lane == 0) {
                q[row] = sum;
                pq += p[row] * sum;
            }
        }
    }
    pq = blockSum<BS>(pq, lds);
    if (threadIdx.x == 0) partials[chunk] = double(pq);
    if (!arriveLast(counter)) return;
    T total = T(0);
    for (int i = threadIdx.x; i < int(gridDim.x); i += BS) total += T(partials[i]);
    total = blockSum<BS>(total, lds);
    if (threadIdx.x != 0) return;
    resetCounter(counter);
    if (!(T(0) < total)) {
        state->done = kNotSpd;
    } else {
        state->alpha = double(T(state->rho) / total);
    }
}

// x += alpha p; r -= alpha q; z = D^-1 r; partials r.r and r.z.
// Plain: last block computes beta and the convergence test.
// Mixed: last block either continues the FP32 recurrence or requests a refinement.
template <class T, class TX, bool kMixed>
__global__ __launch_bounds__(kReduceBlock) void updateKernel(int64_t n, TX* __restrict__ x, T* __restrict__ r,
                                                             const T* __restrict__ p, const T* __restrict__ q,
                                                             const T* __restrict__ inverse, double* partials,
                                                             unsigned* counter, SolverState* state) {
    __shared__ T lds[kReduceBlock / kWave];
    __shared__ T lds2[kReduceBlock / kWave];
    if (state->done) return;
    const T alpha = T(state->alpha);
    const TX alphaX = TX(alpha);
    T rr = T(0), rz = T(0);
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        x[i] = x[i] + alphaX * TX(p[i]);
        const T residual = r[i] + (-alpha) * q[i];
        r[i] = residual;
        const T z = residual * inverse[i];
        rr += residual * residual;
        rz += residual * z;
    }
    rr = blockSum<kReduceBlock>(rr, lds);
    rz = blockSum<kReduceBlock>(rz, lds2);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = double(rr);
        partials[2 * blockIdx.x + 1] = double(rz);
    }
    if (!arriveLast(counter)) return;
    T rrSum = T(0), rzSum = T(0);
    for (int i = threadIdx.x; i < int(gridDim.x); i += kReduceBlock) {
        rrSum += T(partials[2 * i]);
        rzSum += T(partials[2 * i + 1]);
    }
    const T RR = blockSum<kReduceBlock>(rrSum, lds);
    const T RZ = blockSum<kReduceBlock>(rzSum, lds2);
    if (threadIdx.x != 0) return;
    resetCounter(counter);
    SolverState& s = *state;
    const double absolute = sqrt(double(RR));
    const double relative = absolute / s.initialNorm;
    if (kMixed) {
        const bool refine = double(RR) <= s.updateThreshold * s.updateNorm || floatLess(absolute, s.absoluteTolerance)
                         || floatLess(relative, s.relativeTolerance) || s.iteration + 1 >= s.maxIterations;
        if (refine) {
            s.refine = 1;
            return;
        }
        const T previous = T(s.rho);
        s.rho = double(RZ);
        s.beta = double(RZ / previous);
        s.rr = double(RR);
        s.absolute = absolute;
        s.relative = relative;
        s.iteration += 1;
        return;
    }
    // ConjugateGradients_impl.hpp:158-185
    const double previous = s.rho;
    s.rho = double(RZ);
    s.beta = double(RZ) / previous;
    s.rr = double(RR);
    s.absolute = absolute;
    s.relative = relative;
    if (floatLess(absolute, s.absoluteTolerance) || floatLess(relative, s.relativeTolerance)) {
        s.done = kConverged;
    } else {
        s.iteration += 1;
        if (s.iteration >= s.maxIterations) s.done = kMaxIterations;
    }
}

// p = beta p + D^-1 r (ConjugateGradients_impl.hpp:168-170).
template <class T>
__global__ __launch_bounds__(kReduceBlock) void directionKernel(int64_t n, T* __restrict__ p, const T* __restrict__ r,
                                                                const T* __restrict__ inverse,
                                                                const SolverState* state) {
    if (state->done) return;
    const T beta = T(state->beta);
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        p[i] = p[i] * beta + r[i] * inverse[i];
}

// ---------------------------------------------------------------------------
// Host driver
// ---------------------------------------------------------------------------

size_t alignUp(size_t bytes) { return (bytes + 255) / 256 * 256; }

template <class TR>
int launchResidual(cie_hip_csr m, bool start, const double* x, const double* b, TR* r, const TR* inverse, TR* p) {
    cie_hip_context ctx = m->context;
    const CsrLaunch& launch = m->launch64;
    const int perChunk = rowsPerChunk(m->rows, launch);
    const bool ok = withShape(launch.lanes, launch.unroll, [&](auto lanes, auto unroll) {
        constexpr int L = decltype(lanes)::value;
        constexpr int U = decltype(unroll)::value;
        if (start)
            hipLaunchKernelGGL((residualKernel<TR, L, U, kSpmvBlock, true>), dim3(launch.grid), dim3(kSpmvBlock), 0,
                               ctx->stream, int(m->rows), perChunk, m->rowExtents, m->columnIndices, m->entries64, x, b,
                               r, inverse, p, ctx->scratch.partials, ctx->scratch.counters + kStartCounter,
                               ctx->scratch.state);
        else
            hipLaunchKernelGGL((residualKernel<TR, L, U, kSpmvBlock, false>), dim3(launch.grid), dim3(kSpmvBlock), 0,
                               ctx->stream, int(m->rows), perChunk, m->rowExtents, m->columnIndices, m->entries64, x, b,
                               r, inverse, p, ctx->scratch.partials, ctx->scratch.counters + kRefineCounter,
                               ctx->scratch.state);
    });
    if (!ok) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "unsupported SpMV launch shape");
    return CIE_HIP_OK;
}

template <class T>
int launchProductStep(cie_hip_csr m, const T* entries, const CsrLaunch& launch, const T* p, T* q) {
    cie_hip_context ctx = m->context;
    const int perChunk = rowsPerChunk(m->rows, launch);
    const bool ok = withShape(launch.lanes, launch.unroll, [&](auto lanes, auto unroll) {
        constexpr int L = decltype(lanes)::value;
        constexpr int U = decltype(unroll)::value;
        hipLaunchKernelGGL((productKernel<T, L, U, kSpmvBlock>), dim3(launch.grid), dim3(kSpmvBlock), 0, ctx->stream,
                           int(m->rows), perChunk, m->rowExtents, m->columnIndices, entries, p, q,
                           ctx->scratch.partials, ctx->scratch.counters + kProductCounter, ctx->scratch.state);
    });
    if (!ok) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "unsupported SpMV launch shape");
    return CIE_HIP_OK;
}

// out = inScale * out + outScale * solution (ConjugateGradients_impl.hpp:190-198).
int writeOutput(cie_hip_context ctx, double inScale, double outScale, double* out, const double* solution, int64_t n) {
    const dim3 grid(vectorGrid(n)), block(kReduceBlock);
    if (inScale == 1.0) {
        hipLaunchKernelGGL((vectorKernel<double>), grid, block, 0, ctx->stream, int(kAdd), n, out, solution, outScale);
    } else if (inScale == 0.0) {
        hipLaunchKernelGGL((vectorKernel<double>), grid, block, 0, ctx->stream, int(kAssign), n, out, solution, 0.0);
        hipLaunchKernelGGL((vectorKernel<double>), grid, block, 0, ctx->stream, int(kScaleUniform), n, out,
                           static_cast<const double*>(nullptr), outScale);
    } else {
        hipLaunchKernelGGL((vectorKernel<double>), grid, block, 0, ctx->stream, int(kScaleUniform), n, out,
                           static_cast<const double*>(nullptr), inScale);
        hipLaunchKernelGGL((vectorKernel<double>), grid, block, 0, ctx->stream, int(kAdd), n, out, solution, outScale);
    }
    CIE_HIP_CHECK(hipGetLastError());
    return CIE_HIP_OK;
}

// ||b - A x|| with the FP64 plane, using `temp` (n doubles).
int trueResidualNorm(cie_hip_csr m, const double* b, const double* x, double* temp, double* norm) {
    cie_hip_context ctx = m->context;
    int status = cie_hip_assign_f64(ctx, temp, b, m->rows);
    if (status) return status;
    status = cie_hip_csr_product_f64(ctx, m, 1.0, x, m->columns, -1.0, temp, m->rows);
    if (status) return status;
    double rr = 0.0;
    status = cie_hip_inner_product_f64(ctx, temp, temp, m->rows, &rr);
    if (status) return status;
    *norm = std::sqrt(rr);
    return CIE_HIP_OK;
}

int statusOf(int32_t done) {
    switch (done) {
    case kNotSpd: return setError(CIE_HIP_ERROR_NOT_SPD, "p^T @ A @ p vanished");
    case kBreakdown: return setError(CIE_HIP_ERROR_BREAKDOWN, "preconditioned norm vanished in iteration 0 (r^T @ z = 0)");
    default: return CIE_HIP_OK;
    }
}

struct Timer {
    hipEvent_t start = nullptr, marks[2] = {nullptr, nullptr};
    ~Timer() {
        if (start) (void)hipEventDestroy(start);
        for (auto& e : marks) if (e) (void)hipEventDestroy(e);
    }
};

constexpr int kChunk = 16;

// Brackets the first kProfileEvents loop SpMV launches with events (profiling on).
struct ProductProfile {
    int used = 0;
    int begin(cie_hip_context ctx) {
        if (!ctx->profiling || used >= cie_hip_context_s::kProfileEvents) return -1;
        (void)hipEventRecord(ctx->profileEvents[2 * used], ctx->stream);
        return used++;
    }
    void end(cie_hip_context ctx, int mark) {
        if (mark >= 0) (void)hipEventRecord(ctx->profileEvents[2 * mark + 1], ctx->stream);
    }
    // Mean duration of the bracketed kernels whose state was still running is
    // not separable from early exits; the caller passes how many did work.
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

// Runs the loop until the device state says done; returns the solve time.
template <class Enqueue>
int runLoop(cie_hip_context ctx, Timer& timer, Enqueue&& enqueueIteration, double* seconds) {
    SolverState* pinned = ctx->scratch.pinnedState;
    int slot = 0;
    bool pending = false;
    for (;;) {
        for (int i = 0; i < kChunk; ++i) {
            const int status = enqueueIteration();
            if (status) return status;
        }
        CIE_HIP_CHECK(hipGetLastError());
        CIE_HIP_CHECK(hipMemcpyAsync(pinned + slot, ctx->scratch.state, sizeof(SolverState), hipMemcpyDeviceToHost,
                                     ctx->stream));
        CIE_HIP_CHECK(hipEventRecord(timer.marks[slot], ctx->stream));
        if (pending) {
            const int previous = slot ^ 1;
            CIE_HIP_CHECK(hipEventSynchronize(timer.marks[previous]));
            if (pinned[previous].done) {
                float ms = 0.0f;
                CIE_HIP_CHECK(hipEventElapsedTime(&ms, timer.start, timer.marks[previous]));
                *seconds = ms * 1e-3;
                CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
                return CIE_HIP_OK;
            }
        }
        pending = true;
        slot ^= 1;
    }
}

struct Request {
    double inScale, outScale;
    const double* b;
    double* x;
    int64_t size;
    cie_hip_statistics settings;
    bool mixed;
    double refinementFactor;
    int restart;
    bool trueResidual;
};

int solve(cie_hip_context ctx, cie_hip_csr m, const Request& q, cie_hip_solve_info* info) {
    if (!ctx || !m) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "pcg: null handle");
    if (m->rows != m->columns || q.size != m->rows)
        return setError(CIE_HIP_ERROR_SHAPE, "incompatible vector sizes " + std::to_string(q.size) + " != "
                                                 + std::to_string(m->rows));
    if (q.size && (!q.b || !q.x)) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "pcg: null vector");
    if (!m->entries64) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "pcg: matrix has no FP64 plane");
    if (q.mixed && !m->entries32) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "pcg_mixed: matrix has no FP32 plane");
    CIE_HIP_CHECK(hipSetDevice(ctx->device));
    cie_hip_solve_info result{};
    result.stats.iterationCount = 0;
    result.stats.absoluteResidual = std::numeric_limits<double>::max();
    result.stats.relativeResidual = std::numeric_limits<double>::max();
    const int64_t n = q.size;
    if (n == 0 || q.settings.iterationCount == 0) {
        // ConjugateGradients_impl.hpp:77: no iterations requested -> solution = x.
        if (n) {
            int status = writeOutput(ctx, q.inScale, q.outScale, q.x, q.x, n);
            if (status) return status;
            CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        }
        if (info) *info = result;
        return CIE_HIP_OK;
    }
    int status = ensureInverseDiagonal(m, q.mixed ? 32 : 64);
    if (status) return status;

    const size_t valueSize = q.mixed ? sizeof(float) : sizeof(double);
    const size_t vec64 = alignUp(sizeof(double) * n), vecR = alignUp(valueSize * n);
    void* workspace = nullptr;
    status = matrixWorkspace(m, 2 * vec64 + 3 * vecR, &workspace);
    if (status) return status;
    char* base = static_cast<char*>(workspace);
    double* solution = reinterpret_cast<double*>(base);
    double* temp64 = reinterpret_cast<double*>(base + vec64);
    void* r = base + 2 * vec64;
    void* p = base + 2 * vec64 + vecR;
    void* qv = base + 2 * vec64 + 2 * vecR;

    SolverState init{};
    init.absoluteTolerance = q.settings.absoluteResidual;
    init.relativeTolerance = q.settings.relativeResidual;
    init.maxIterations = q.settings.iterationCount;
    init.updateThreshold = q.refinementFactor * q.refinementFactor;
    init.restart = q.restart;
    *ctx->scratch.pinnedState = init;
    CIE_HIP_CHECK(hipMemcpyAsync(ctx->scratch.state, ctx->scratch.pinnedState, sizeof(SolverState),
                                 hipMemcpyHostToDevice, ctx->stream));
    CIE_HIP_CHECK(hipMemcpyAsync(solution, q.x, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));

    Timer timer;
    CIE_HIP_CHECK(hipEventCreate(&timer.start));
    CIE_HIP_CHECK(hipEventCreate(&timer.marks[0]));
    CIE_HIP_CHECK(hipEventCreate(&timer.marks[1]));
    CIE_HIP_CHECK(hipEventRecord(timer.start, ctx->stream));

    const dim3 vgrid(vectorGrid(n)), vblock(kReduceBlock);
    double seconds = 0.0;
    ProductProfile profile;
    if (!q.mixed) {
        double* r64 = static_cast<double*>(r);
        double* p64 = static_cast<double*>(p);
        double* q64 = static_cast<double*>(qv);
        status = launchResidual<double>(m, true, solution, q.b, r64, m->inverseDiagonal64, p64);
        if (status) return status;
        status = runLoop(ctx, timer, [&]() -> int {
            const int mark = profile.begin(ctx);
            int s = launchProductStep<double>(m, m->entries64, m->launch64, p64, q64);
            if (s) return s;
            profile.end(ctx, mark);
            hipLaunchKernelGGL((updateKernel<double, double, false>), vgrid, vblock, 0, ctx->stream, n, solution, r64,
                               p64, q64, m->inverseDiagonal64, ctx->scratch.partials,
                               ctx->scratch.counters + kUpdateCounter, ctx->scratch.state);
            hipLaunchKernelGGL((directionKernel<double>), vgrid, vblock, 0, ctx->stream, n, p64, r64,
                               m->inverseDiagonal64, ctx->scratch.state);
            return CIE_HIP_OK;
        }, &seconds);
    } else {
        float* r32 = static_cast<float*>(r);
        float* p32 = static_cast<float*>(p);
        float* q32 = static_cast<float*>(qv);
        status = launchResidual<float>(m, true, solution, q.b, r32, m->inverseDiagonal32, p32);
        if (status) return status;
        status = runLoop(ctx, timer, [&]() -> int {
            const int mark = profile.begin(ctx);
            int s = launchProductStep<float>(m, m->entries32, m->launch32, p32, q32);
            if (s) return s;
            profile.end(ctx, mark);
            hipLaunchKernelGGL((updateKernel<float, double, true>), vgrid, vblock, 0, ctx->stream, n, solution, r32,
                               p32, q32, m->inverseDiagonal32, ctx->scratch.partials,
                               ctx->scratch.counters + kUpdateCounter, ctx->scratch.state);
            s = launchResidual<float>(m, false, solution, q.b, r32, m->inverseDiagonal32, p32);
            if (s) return s;
            hipLaunchKernelGGL((directionKernel<float>), vgrid, vblock, 0, ctx->stream, n, p32, r32,
                               m->inverseDiagonal32, ctx->scratch.state);
            return CIE_HIP_OK;
        }, &seconds);
    }
    if (status) return status;
    CIE_HIP_CHECK(hipMemcpyAsync(ctx->scratch.pinnedState, ctx->scratch.state, sizeof(SolverState),
                                 hipMemcpyDeviceToHost, ctx->stream));
    CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    const SolverState final = *ctx->scratch.pinnedState;
    result.stats.iterationCount = final.iteration;
    result.stats.absoluteResidual = final.absolute;
    result.stats.relativeResidual = final.relative;
    result.converged = final.done == kConverged;
    result.seconds = seconds;
    result.refinements = final.refinements;
    result.spmvCount = final.iteration + (final.done == kConverged ? 1 : 0) + 1 + final.refinements;
    if (ctx->profiling) {
        // Loop trips that ran the product kernel: iteration (+1 for the converged trip).
        const int worked = int(std::min<uint64_t>(final.iteration + (final.done == kConverged ? 1 : 0),
                                                  uint64_t(cie_hip_context_s::kProfileEvents)));
        result.productSamples = uint64_t(std::min(worked, profile.used));
        result.productSeconds = profile.mean(ctx, worked);
    }
    status = statusOf(final.done);
    result.status = status;
    if (status) {
        if (info) *info = result;
        return status;
    }
    if (q.trueResidual) {
        double norm = 0.0;
        status = trueResidualNorm(m, q.b, solution, temp64, &norm);
        if (status) return status;
        result.trueRelativeResidual = final.initialNorm > 0 ? norm / final.initialNorm : norm;
    }
    status = writeOutput(ctx, q.inScale, q.outScale, q.x, solution, n);
    if (status) return status;
    CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (info) *info = result;
    return CIE_HIP_OK;
}

}  // namespace

extern "C" {

int cie_hip_pcg_f64(cie_hip_context context, cie_hip_csr matrix, double inScale, const double* b, double outScale,
                    double* x, int64_t size, const cie_hip_statistics* settings, int trueResidual,
                    cie_hip_solve_info* info) {
    if (!settings) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_pcg_f64: null settings");
    Request q{inScale, outScale, b, x, size, *settings, false, 0.0, 0, trueResidual != 0};
    return solve(context, matrix, q, info);
}

int cie_hip_pcg_mixed(cie_hip_context context, cie_hip_csr matrix, double inScale, const double* b, double outScale,
                      double* x, int64_t size, const cie_hip_mixed_settings* settings, cie_hip_solve_info* info) {
    if (!settings) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_pcg_mixed: null settings");
    double factor = settings->refinementFactor;
    if (!(factor > 0.0 && factor < 1.0)) factor = 0.1;
    Request q{inScale, outScale, b, x, size, settings->settings, true, factor, settings->restart ? 1 : 0, true};
    return solve(context, matrix, q, info);
}

}  // extern "C"

// Fused Jacobi-preconditioned conjugate gradients on one device.
//
// Algorithm and step order follow ConjugateGradients<TS>::product
// (linalg/packages/solvers/impl/ConjugateGradients_impl.hpp:49-201) with the
// Jacobi DiagonalOperator as preconditioner:
//   r = b - A x; ||r0|| = sqrt(r.r); early exit;  z = D^-1 r; rho = r.z; p = z
//   loop: q = A p; alpha = rho / p.q; x += alpha p; r -= alpha q; rr = r.r;
//         z = D^-1 r; rho' = r.z; beta = rho'/rho; p = beta p + z;
//         abs = sqrt(rr), rel = abs / ||r0||; stop on Comparison::less.
// One loop trip is three streaming kernels (four in the mixed form), and the
// scalars never leave the device:
//   product:   q = A p, one p.q partial per block (then, for more than 4096
//              tiles, a first-level sum of those partials);
//   update:    every block sums the p.q partials itself, in one fixed order,
//              so all blocks hold the same alpha; x += alpha p, r -= alpha q,
//              one partial of r.r and r.z per block;
//   direction: every block sums the r.r and r.z partials the same way, forms
//              beta and the convergence test, and updates p = beta p + D^-1 r.
// Block 0 of a kernel records its scalars in the solver state; no block reads
// a state field that the same kernel writes (rho lives in two slots, picked
// by the host's iteration parity, so rho is read from one slot and written to
// the other), so no cross-block synchronisation is needed. The host only polls
// the device state once per chunk of iterations.
//
// The mixed form keeps A, r, p, q, z, the correction d and their dot products
// in FP32 and the solution x in FP64. Whenever the FP32 residual of the
// previous trip has fallen by the refinement factor since the last refinement
// (or looks converged, or the iteration cap is near), the update folds the
// correction into x (x += d in FP64, d = 0) and a refinement step recomputes
// r = b - A x with the FP64 matrix, tests convergence on that FP64 residual,
// and restarts the FP32 recurrence from it; the search direction is kept
// (restart = 0) or reset (restart = 1). The refinement kernel runs every
// trip and returns at once when the update did not decide to refine.
#include "dispatch.hpp"
#include "kernels.hpp"

#include <cmath>
#include <limits>

using namespace ciehip;

namespace {

constexpr int kVectorThreads = 1024;  // update and direction blocks
constexpr int kVectorBlocksMax = kVectorPartials;  // at most this many update partials
constexpr int kFinishBlock = 1024;
// Rows per thread of the vector kernels loaded before their partial sums: as
// many as keep two 1024-thread blocks resident per compute unit (<= 64 VGPRs).
template <class T>
constexpr int aheadRows() { return sizeof(T) == sizeof(double) ? 2 : 3; }

// Blocks of the update and direction kernels for n rows: about 4 rows per
// thread for large systems; a smaller system still spreads over up to 256
// blocks (one per CU) with fewer rows per thread.
int vectorBlocks(int64_t n) {
    const int64_t four = (n + int64_t(kVectorThreads) * 4 - 1) / (int64_t(kVectorThreads) * 4);
    const int64_t one = std::min<int64_t>(256, (n + kVectorThreads - 1) / kVectorThreads);
    return int(std::max<int64_t>(1, std::min<int64_t>(kVectorBlocksMax, std::max(four, one))));
}

// Sum over the block of `count` partials (value `which` of each slot of two),
// each converted to TR, in one fixed order; every thread returns the same bits.
template <int BS, class TR>
__device__ __forceinline__ TR sumPartialsAll(const double* __restrict__ partials, int count, int which, TR* lds) {
    TR value = TR(0);
    for (int i = threadIdx.x; i < count; i += BS) value += TR(partials[2 * i + which]);
    value = groupSum<kWave>(value);
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) lds[threadIdx.x / kWave] = value;
    __syncthreads();
    TR total = TR(0);
#pragma unroll
    for (int w = 0; w < BS / kWave; ++w) total += lds[w];
    __syncthreads();  // lds is reused by the caller
    return total;
}

__device__ __forceinline__ bool firstThread() { return blockIdx.x == 0 && threadIdx.x == 0; }

// The residual history's entry of loop trip `trip` (one thread, when recording).
__device__ __forceinline__ void recordResidual(const SolverState& s, uint64_t trip, double absolute) {
    if (s.history && trip < s.historyCapacity) s.history[trip] = absolute;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// Start: r = b - A x (FP64 matrix and x), z = D^-1 r, p = z; partials r.r
// (FP64) and r.z (recurrence precision TR) per block.
template <class Walker, class TR>
__global__ __launch_bounds__(Walker::kBlock) void startKernel(
    RowSource source, const double* __restrict__ entries, const double* __restrict__ x, const double* __restrict__ b,
    TR* __restrict__ r, const TR* __restrict__ inverse, TR* __restrict__ p, double* __restrict__ partials) {
    constexpr int BS = Walker::kBlock;
    __shared__ double ldsRR[BS / kWave];
    __shared__ TR ldsRZ[BS / kWave];
    double rr = 0.0;
    TR rz = TR(0);
    Walker::run(source, entries, x, [&](int row, double sum) {
        const double t = b[row] - sum;
        const TR residual = TR(t);
        const TR z = residual * inverse[row];
        r[row] = residual;
        p[row] = z;
        rr += t * t;
        rz += residual * z;
    });
    const int slot = Walker::slot();
    rr = blockSum<BS>(rr, ldsRR);
    rz = blockSum<BS>(rz, ldsRZ);
    if (threadIdx.x == 0) {
        partials[2 * slot] = rr;
        partials[2 * slot + 1] = double(rz);
    }
}

// q = A p, partial p.q (ConjugateGradients_impl.hpp:136-143). With
// `correction` (the mixed form's FP32 product), q_i = (A32 p)_i + c_i p_i:
// the row-sum correction restores the FP64 row sums that rounding the values
// to FP32 perturbs (all in FP32 arithmetic).
template <class Walker>
__global__ __launch_bounds__(Walker::kBlock) void productKernel(
    RowSource source, const typename Walker::Matrix* __restrict__ entries, const typename Walker::Value* __restrict__ p,
    typename Walker::Value* __restrict__ q, double* __restrict__ partials, const SolverState* __restrict__ state,
    const typename Walker::Value* __restrict__ correction) {
    using T = typename Walker::Value;
    constexpr int BS = Walker::kBlock;
    __shared__ T lds[BS / kWave];
    if (state->done) return;
    T pq = T(0);
    Walker::run(source, entries, p, [&](int row, T sum) {
        const T pi = p[row];
        const T value = correction ? sum + correction[row] * pi : sum;
        q[row] = value;
        pq += pi * value;
    });
    pq = blockSum<BS>(pq, lds);
    if (threadIdx.x == 0) partials[2 * Walker::slot()] = double(pq);
}

// The same product for walkers over tiles of whole rows. Thread `tid` owns
// row first + tid of the tile: it loads p and the correction of that row when
// the tile starts (the loads are in flight while the tile streams), the
// walker leaves each row sum in LDS, and thread `tid` then writes q and adds
// its p.q term: coalesced stores, and no dependent load after the row sums.
template <class Walker>
__global__ __launch_bounds__(Walker::kBlock) void tileProductKernel(
    RowSource source, const typename Walker::Matrix* __restrict__ entries, const typename Walker::Value* __restrict__ p,
    typename Walker::Value* __restrict__ q, double* __restrict__ partials, const SolverState* __restrict__ state,
    const typename Walker::Value* __restrict__ correction) {
    using T = typename Walker::Value;
    constexpr int BS = Walker::kBlock;
    __shared__ T stage[Walker::kCapacity];
    __shared__ int32_t extents[BS + 1];
    __shared__ T rowSums[BS];
    __shared__ T lds[BS / kWave];
    if (state->done) return;
    const int tid = threadIdx.x;
    T pq = T(0);
    for (int t = blockIdx.x; t < source.tileCount; t += gridDim.x) {
        const int first = source.tiles[t].x;
        const int rows = source.tiles[t + 1].x - first;  // at most BS rows per tile
        T pi = T(0), ci = T(0);
        if (tid < rows) {
            pi = p[first + tid];
            if (correction) ci = correction[first + tid];
        }
        auto stash = [&](int row, T sum) { rowSums[row - first] = sum; };
        Walker::tile(t, source, entries, p, stash, stage, extents);
        __syncthreads();
        if (tid < rows) {
            const T sum = rowSums[tid];
            const T value = correction ? sum + ci * pi : sum;
            q[first + tid] = value;
            pq += pi * value;
        }
        __syncthreads();  // stage, extents and rowSums are reused by the next tile
    }
    pq = blockSum<BS>(pq, lds);
    if (threadIdx.x == 0) partials[2 * blockIdx.x] = double(pq);
}

// First level of a sum over more than kLevelBlocks partials (one per tile of
// a tiled SpMV): block b sums slots [b * per, min(count, b * per + per)) in a
// fixed order, both values of a slot, in FP64. One compute unit cannot read
// tens of thousands of partials quickly; kLevelBlocks of them can.
__global__ __launch_bounds__(kReduceBlock) void partialLevelKernel(int count, int per,
                                                                   const double* __restrict__ partials,
                                                                   double* __restrict__ level,
                                                                   const SolverState* __restrict__ state) {
    __shared__ double ldsA[kReduceBlock / kWave];
    __shared__ double ldsB[kReduceBlock / kWave];
    if (state->done) return;
    const int begin = int(blockIdx.x) * per;
    const int end = min(count, begin + per);
    double a = 0.0, b = 0.0;
    for (int i = begin + int(threadIdx.x); i < end; i += kReduceBlock) {
        a += partials[2 * i];
        b += partials[2 * i + 1];
    }
    a = blockSum<kReduceBlock>(a, ldsA);
    b = blockSum<kReduceBlock>(b, ldsB);
    if (threadIdx.x == 0) {
        level[2 * blockIdx.x] = a;
        level[2 * blockIdx.x + 1] = b;
    }
}

// alpha = rho / p.q (:139-151; p.q <= 0 ends the solve), then x += alpha p;
// r -= alpha q; z = D^-1 r; partials r.r and r.z (:154-163).
template <class T, class TX>
__global__ __launch_bounds__(kVectorThreads) void updateKernel(int slot, int64_t n, TX* __restrict__ x,
                                                               T* __restrict__ r, const T* __restrict__ p,
                                                               const T* __restrict__ q, const T* __restrict__ inverse,
                                                               int productCount, const double* __restrict__ productPartials,
                                                               double* __restrict__ partials,
                                                               SolverState* __restrict__ state) {
    __shared__ T lds[kVectorThreads / kWave];
    __shared__ T lds2[kVectorThreads / kWave];
    if (state->done) return;
    constexpr int kAhead = aheadRows<TX>();
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    const int64_t first = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    TX xs[kAhead];
    T ps[kAhead], qs[kAhead], rs[kAhead], is[kAhead];
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {  // the first rows' loads are in flight during the partial sum
        const int64_t i = first + k * stride;
        if (i < n) {
            xs[k] = x[i];
            ps[k] = p[i];
            qs[k] = q[i];
            rs[k] = r[i];
            is[k] = inverse[i];
        }
    }
    const T pq = sumPartialsAll<kVectorThreads, T>(productPartials, productCount, 0, lds);
    if (!(T(0) < pq)) {
        if (firstThread()) state->done = kNotSpd;
        return;
    }
    const T alpha = T(T(state->rhos[slot]) / pq);
    if (firstThread()) state->alpha = double(alpha);
    const TX alphaX = TX(alpha);
    T rr = T(0), rz = T(0);
    const auto row = [&](int64_t i, TX xi, T pi, T qi, T ri, T ii) {
        x[i] = xi + alphaX * TX(pi);
        const T residual = ri + (-alpha) * qi;
        r[i] = residual;
        const T z = residual * ii;
        rr += residual * residual;
        rz += residual * z;
    };
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {
        const int64_t i = first + k * stride;
        if (i < n) row(i, xs[k], ps[k], qs[k], rs[k], is[k]);
    }
    for (int64_t i = first + kAhead * stride; i < n; i += stride) row(i, x[i], p[i], q[i], r[i], inverse[i]);
    rr = blockSum<kVectorThreads>(rr, lds);
    rz = blockSum<kVectorThreads>(rz, lds2);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = double(rr);
        partials[2 * blockIdx.x + 1] = double(rz);
    }
}

// Mixed form: does this trip refine? The FP32 r.r of the previous trip (as
// the direction kernel recorded it) fell by the refinement factor since the
// last refinement, looks converged, or the iteration cap is near.
__device__ __forceinline__ bool refineDue(const SolverState& s) {
    const double absolute = sqrt(s.rr);
    const double relative = absolute / s.initialNorm;
    return s.rr <= s.updateThreshold * s.updateNorm || floatLess(absolute, s.absoluteTolerance) ||
           floatLess(relative, s.relativeTolerance) || s.iteration + 1 >= s.maxIterations;
}

// Mixed form: alpha = rho / p.q; the FP32 correction d += alpha p, folded
// into the FP64 x (and cleared) on a trip that refines; r -= alpha q;
// partials r.r and r.z. Every block takes the refinement decision from the
// state the previous trip left (none of which this kernel writes) and block 0
// records it in the slot of this trip's parity. If the previous trip folded,
// this trip's level kernel has replaced r by the rounded true residual
// b - A x of the folded x, with its FP64 r.r in `refinePartials`: the
// convergence test runs on that FP64 norm first (x is final if it passes),
// and the refinement threshold restarts from it.
__global__ __launch_bounds__(kVectorThreads) void mixedUpdateKernel(int slot, int64_t n, double* __restrict__ x,
                                                                    float* __restrict__ d, float* __restrict__ r,
                                                                    const float* __restrict__ p,
                                                                    const float* __restrict__ q,
                                                                    const float* __restrict__ inverse,
                                                                    int productCount,
                                                                    const double* __restrict__ productPartials,
                                                                    int refineCount,
                                                                    const double* __restrict__ refinePartials,
                                                                    double* __restrict__ partials,
                                                                    SolverState* __restrict__ state) {
    __shared__ float lds[kVectorThreads / kWave];
    __shared__ float lds2[kVectorThreads / kWave];
    __shared__ double lds64[kVectorThreads / kWave];
    SolverState& s = *state;
    if (s.done) return;
    const bool replaced = s.refined[slot ^ 1] != 0;
    constexpr int kAhead = aheadRows<float>();
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    const int64_t first = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    float ds[kAhead], ps[kAhead], qs[kAhead], rs[kAhead], is[kAhead];
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {  // the first rows' loads are in flight during the partial sums
        const int64_t i = first + k * stride;
        if (i < n) {
            ds[k] = d[i];
            ps[k] = p[i];
            qs[k] = q[i];
            rs[k] = r[i];
            is[k] = inverse[i];
        }
    }
    bool fold;
    if (replaced) {
        const double rr = sumPartialsAll<kVectorThreads, double>(refinePartials, refineCount, 0, lds64);
        const double absolute = sqrt(rr);
        const double relative = absolute / s.initialNorm;
        const bool converged = floatLess(absolute, s.absoluteTolerance) || floatLess(relative, s.relativeTolerance);
        if (firstThread()) {
            s.refinements += 1;
            s.rr = rr;
            s.absolute = absolute;
            s.relative = relative;
            s.updateNorm = rr;
            // The trip that folded (its direction step counted it) reports the FP64 residual.
            recordResidual(s, s.iteration - 1, absolute);
            // The trip that folded counts as the last one: its direction step already counted it.
            if (converged) {
                s.iteration -= 1;
                s.done = kConverged;
            }
        }
        if (converged) return;
        fold = s.iteration + 1 >= s.maxIterations;  // just refined: only the iteration cap folds again
    } else {
        fold = refineDue(s);
    }
    const float pq = sumPartialsAll<kVectorThreads, float>(productPartials, productCount, 0, lds);
    // p.q = 0 on a trip that folds, or right after one: the FP32 recurrence solved its system exactly (r, then
    // p, became zero). The trip takes no step (folding, the true residual then decides; after a fold, the
    // direction restarts from the true residual); any other p.q <= 0 ends the solve.
    const bool vanished = pq == 0.0f && (fold || replaced);
    if (!(0.0f < pq) && !vanished) {
        if (firstThread()) s.done = kNotSpd;
        return;
    }
    const float alpha = vanished ? 0.0f : float(s.rhos[slot]) / pq;
    if (firstThread()) {
        s.alpha = double(alpha);
        s.refined[slot] = fold ? 1 : 0;
    }
    float rr = 0.0f, rz = 0.0f;
    const auto row = [&](int64_t i, float di, float pi, float qi, float ri, float ii) {
        const float correction = di + alpha * pi;
        if (fold) {
            x[i] = x[i] + double(correction);
            d[i] = 0.0f;
        } else {
            d[i] = correction;
        }
        const float residual = ri + (-alpha) * qi;
        r[i] = residual;
        const float z = residual * ii;
        rr += residual * residual;
        rz += residual * z;
    };
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {
        const int64_t i = first + k * stride;
        if (i < n) row(i, ds[k], ps[k], qs[k], rs[k], is[k]);
    }
    for (int64_t i = first + kAhead * stride; i < n; i += stride) row(i, d[i], p[i], q[i], r[i], inverse[i]);
    rr = blockSum<kVectorThreads>(rr, lds);
    rz = blockSum<kVectorThreads>(rz, lds2);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = double(rr);
        partials[2 * blockIdx.x + 1] = double(rz);
    }
}

// Mixed form, the kernel after the product, every trip. Blocks below
// `levelBlocks` first sum the product's FP32 p.q partials into `level`, in
// FP32 (the first level of the sum; none when the product wrote few enough
// partials).
// Then, if the previous trip folded the correction into x (refined[previous]),
// every block recomputes its rows of r = b - A x with the FP64 plane and x
// (rounded into the FP32 r) and writes a partial of the FP64 r.r to
// `refinePartials`; else it returns. The product ahead of it used only p, so
// the true residual replaces the recursive one before this trip's update,
// which tests convergence on its FP64 norm.
template <class Walker>
__global__ __launch_bounds__(Walker::kBlock) void levelRefineKernel(
    int levelBlocks, int levelCount, int levelPer, const double* __restrict__ productPartials,
    double* __restrict__ level, int previous, RowSource source, const double* __restrict__ entries,
    const double* __restrict__ x, const double* __restrict__ b, float* __restrict__ r,
    double* __restrict__ refinePartials, const SolverState* __restrict__ state) {
    constexpr int BS = Walker::kBlock;
    static_assert(BS == kReduceBlock, "the level sums use the reduction block size");
    __shared__ double ldsA[BS / kWave];
    __shared__ float ldsF[BS / kWave];
    if (state->done) return;
    if (int(blockIdx.x) < levelBlocks) {  // the FP32 p.q partials, summed in FP32
        const int begin = int(blockIdx.x) * levelPer;
        const int end = min(levelCount, begin + levelPer);
        float a = 0.0f;
        for (int i = begin + int(threadIdx.x); i < end; i += BS) a += float(productPartials[2 * i]);
        a = blockSum<BS>(a, ldsF);
        if (threadIdx.x == 0) level[2 * blockIdx.x] = double(a);
    }
    if (!state->refined[previous]) return;
    double rr = 0.0;
    Walker::run(source, entries, x, [&](int row, double sum) {
        const double t = b[row] - sum;
        r[row] = float(t);
        rr += t * t;
    });
    rr = blockSum<BS>(rr, ldsA);
    if (threadIdx.x == 0) refinePartials[2 * Walker::slot()] = rr;
}

// beta, the convergence test and the iteration count (:158-185), then
// p = beta p + D^-1 r (:168-170: scale, then add). kMixed: the FP32 recurrence
// step only (the mixed form tests convergence on FP64 true residuals, in the
// update kernel).
template <class T, bool kMixed>
__global__ __launch_bounds__(kVectorThreads) void directionKernel(int slot, int64_t n, T* __restrict__ p,
                                                                  const T* __restrict__ r, const T* __restrict__ inverse,
                                                                  int updateCount, const double* __restrict__ updatePartials,
                                                                  SolverState* __restrict__ state) {
    __shared__ T ldsT[kVectorThreads / kWave];
    SolverState& s = *state;
    if (s.done) return;
    constexpr int kAhead = aheadRows<T>();
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    const int64_t first = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    T ps[kAhead], rs[kAhead], is[kAhead];
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {  // the first rows' loads are in flight during the partial sums
        const int64_t i = first + k * stride;
        if (i < n) {
            ps[k] = p[i];
            rs[k] = r[i];
            is[k] = inverse[i];
        }
    }
    const double rhoPrevious = s.rhos[slot];
    bool converged = false;
    double beta = 0.0;
    if constexpr (!kMixed) {
        const T rr = sumPartialsAll<kVectorThreads, T>(updatePartials, updateCount, 0, ldsT);
        const T rz = sumPartialsAll<kVectorThreads, T>(updatePartials, updateCount, 1, ldsT);
        const double absolute = sqrt(double(rr));
        const double relative = absolute / s.initialNorm;
        converged = floatLess(absolute, s.absoluteTolerance) || floatLess(relative, s.relativeTolerance);
        beta = double(rz) / rhoPrevious;
        if (firstThread()) {
            s.rhos[slot ^ 1] = double(rz);
            s.beta = beta;
            s.rr = double(rr);
            s.absolute = absolute;
            s.relative = relative;
            recordResidual(s, s.iteration, absolute);
            if (converged) {
                s.done = kConverged;
            } else {
                s.iteration += 1;
                if (s.iteration >= s.maxIterations) s.done = kMaxIterations;
            }
        }
    } else {
        // The FP32 recurrence step. A trip that folded the correction into x
        // resets the search direction here when `restart` is set; its true
        // residual replaces r on the next trip (levelRefineKernel).
        const float rr = sumPartialsAll<kVectorThreads, float>(updatePartials, updateCount, 0, ldsT);
        const float rz = sumPartialsAll<kVectorThreads, float>(updatePartials, updateCount, 1, ldsT);
        // A vanished FP32 residual (rho 0) restarts the direction from the next residual, and so does a
        // replaced residual (this trip's r started from the true one) that exceeds the recursive one by more
        // than kRestartGap.
        const bool gap = s.refined[slot ^ 1] && double(rz) > kRestartGap * rhoPrevious;
        beta = (s.restart && s.refined[slot]) || float(rhoPrevious) == 0.0f || gap
                   ? 0.0
                   : double(rz / float(rhoPrevious));
        if (firstThread()) {
            const double absolute = sqrt(double(rr));
            s.rhos[slot ^ 1] = double(rz);
            s.beta = beta;
            s.rr = double(rr);
            s.absolute = absolute;
            s.relative = absolute / s.initialNorm;
            recordResidual(s, s.iteration, absolute);
            s.iteration += 1;
            if (s.iteration >= s.maxIterations) s.done = kMaxIterations;
        }
    }
    if (converged) return;
    const T scale = T(beta);
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {
        const int64_t i = first + k * stride;
        if (i < n) p[i] = ps[k] * scale + rs[k] * is[k];
    }
    for (int64_t i = first + kAhead * stride; i < n; i += stride) p[i] = p[i] * scale + r[i] * inverse[i];
}

// Single block, once per solve: the start sums (r.r in FP64, r.z in TR) and
// the early exit of ConjugateGradients_impl.hpp:93-132.
template <class TR>
__global__ __launch_bounds__(kFinishBlock) void startFinishKernel(int count, const double* __restrict__ partials,
                                                                  SolverState* __restrict__ state) {
    __shared__ double lds64[kFinishBlock / kWave];
    __shared__ TR ldsR[kFinishBlock / kWave];
    const double rr = sumPartialsAll<kFinishBlock, double>(partials, count, 0, lds64);
    const TR rz = sumPartialsAll<kFinishBlock, TR>(partials, count, 1, ldsR);
    if (threadIdx.x != 0) return;
    SolverState& s = *state;
    const double initial = sqrt(rr);
    s.rr = rr;
    s.initialNorm = initial;
    s.absolute = initial;
    s.relative = 1.0;
    s.updateNorm = rr;
    if (floatLess(initial, s.absoluteTolerance) || floatLess(1.0, s.relativeTolerance)) {
        s.done = kConverged;
        s.startExit = 1;
    } else if (!(TR(0) < rz)) {
        s.done = kBreakdown;
    } else {
        s.rhos[0] = double(rz);
    }
}

// ---------------------------------------------------------------------------
// Host driver
// ---------------------------------------------------------------------------

size_t alignUp(size_t bytes) { return (bytes + 255) / 256 * 256; }

// Partials of a product or start launch as their consumers (blocks of 1024
// threads) read them: up to kDirectPartials slots directly, four loads per
// thread at most; a first-level sum into kLevelBlocks slots above that.
struct PartialSource {
    const double* partials;
    int count;
};

constexpr int kDirectPartials = 4 * kVectorThreads;
static_assert(kFinishBlock == kVectorThreads, "the start sums read the partials as the update kernels do");

PartialSource levelPartials(cie_hip_context ctx, int count) {
    if (count <= kDirectPartials) return {ctx->scratch.partials, count};
    const int per = (count + kLevelBlocks - 1) / kLevelBlocks;
    const int blocks = (count + per - 1) / per;
    hipLaunchKernelGGL(partialLevelKernel, dim3(blocks), dim3(kReduceBlock), 0, ctx->stream, count, per,
                       ctx->scratch.partials, ctx->scratch.level, ctx->scratch.state);
    return {ctx->scratch.level, blocks};
}

// r = b - A x, z, p = z and the start sums.
template <class TR>
int launchStart(cie_hip_csr m, const double* x, const double* b, TR* r, const TR* inverse, TR* p) {
    cie_hip_context ctx = m->context;
    const CsrLaunch& launch = m->launch64;
    int grid = 0;
    int status = CIE_HIP_OK;
    const bool ok = withWalker<double, double>(launch, [&](auto walker) {
        using Walker = decltype(walker);
        RowSource source;
        status = prepareLaunch(m, launch, startKernel<Walker, TR>, source, grid);
        if (status == CIE_HIP_OK) status = ensurePartials(ctx, grid);
        if (status != CIE_HIP_OK) return;
        hipLaunchKernelGGL((startKernel<Walker, TR>), dim3(grid), dim3(Walker::kBlock), 0, ctx->stream, source,
                           m->entries64, x, b, r, inverse, p, ctx->scratch.partials);
    });
    if (!ok) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "unsupported SpMV launch shape");
    if (status != CIE_HIP_OK) return status;
    const PartialSource sums = levelPartials(ctx, grid);
    hipLaunchKernelGGL((startFinishKernel<TR>), dim3(1), dim3(kFinishBlock), 0, ctx->stream, sums.count, sums.partials,
                       ctx->scratch.state);
    return CIE_HIP_OK;
}

// Mixed form: the kernel after the product (levelRefineKernel) with the
// FP64 walker on a resident-round grid (it usually only sums the product's
// partials). `products` are the product's partials; returns where the update
// kernel reads them in `sums` and the refinement partials' count in `count`.
int launchLevelRefine(cie_hip_csr m, int previous, int products, const double* x, const double* b, float* r,
                      PartialSource* sums, int* count) {
    cie_hip_context ctx = m->context;
    const CsrLaunch& launch = m->launch64;
    int levelBlocks = 0, per = 0;
    if (products > kDirectPartials) {
        per = (products + kLevelBlocks - 1) / kLevelBlocks;
        levelBlocks = (products + per - 1) / per;
        *sums = {ctx->scratch.level, levelBlocks};
    } else {
        *sums = {ctx->scratch.partials, products};
    }
    int grid = 0;
    int status = CIE_HIP_OK;
    const bool ok = withWalker<double, double>(launch, [&](auto walker) {
        using Walker = decltype(walker);
        RowSource source;
        status = prepareLaunch(m, launch, levelRefineKernel<Walker>, source, grid, true);
        if (status != CIE_HIP_OK) return;
        // A larger grid only adds blocks without rows: the tile walkers stride
        // over tiles, the row-chunk walkers' extra chunks lie past the last row.
        grid = std::max(grid, levelBlocks);
        if (grid > kMaxPartials) {
            status = setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "refinement grid exceeds its partial slots");
            return;
        }
        hipLaunchKernelGGL((levelRefineKernel<Walker>), dim3(grid), dim3(Walker::kBlock), 0, ctx->stream, levelBlocks,
                           products, per, ctx->scratch.partials, ctx->scratch.level, previous, source, m->entries64, x,
                           b, r, ctx->scratch.refinePartials, static_cast<const SolverState*>(ctx->scratch.state));
    });
    if (!ok) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "unsupported SpMV launch shape");
    *count = grid;
    return status;
}

// Brackets the first kProfileEvents loop SpMV kernels with events on the
// solver stream (profiling on).
// q = A p with its p.q partials (first-level sum included); returns where the
// update kernel reads them.
// `sums` null: the caller sums the partials (the mixed form's level kernel);
// `*partialCount` receives their number.
template <class T>
int launchProduct(cie_hip_csr m, const T* entries, const CsrLaunch& launch, const T* p, T* q,
                  ProductProfile& profile, PartialSource* sums, const T* correction = nullptr,
                  int* partialCount = nullptr) {
    cie_hip_context ctx = m->context;
    int grid = 0;
    int status = CIE_HIP_OK;
    const bool ok = withWalker<T, T>(launch, [&](auto walker) {
        using Walker = decltype(walker);
        RowSource source;
        if constexpr (IsTileWalker<Walker>::value) {
            status = prepareLaunch(m, launch, tileProductKernel<Walker>, source, grid);
            if (status == CIE_HIP_OK) status = ensurePartials(ctx, grid);
            if (status != CIE_HIP_OK) return;
            profile.launch(ctx, tileProductKernel<Walker>, dim3(grid), dim3(Walker::kBlock), source, entries, p, q,
                           ctx->scratch.partials, static_cast<const SolverState*>(ctx->scratch.state), correction);
        } else {
            status = prepareLaunch(m, launch, productKernel<Walker>, source, grid);
            if (status == CIE_HIP_OK) status = ensurePartials(ctx, grid);
            if (status != CIE_HIP_OK) return;
            profile.launch(ctx, productKernel<Walker>, dim3(grid), dim3(Walker::kBlock), source, entries, p, q,
                           ctx->scratch.partials, static_cast<const SolverState*>(ctx->scratch.state), correction);
        }
    });
    if (!ok) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "unsupported SpMV launch shape");
    if (status != CIE_HIP_OK) return status;
    if (partialCount) *partialCount = grid;
    if (sums) *sums = levelPartials(ctx, grid);
    return CIE_HIP_OK;
}

// out = inScale * out + outScale * solution (ConjugateGradients_impl.hpp:190-198).
int writeOutput(cie_hip_context ctx, double inScale, double outScale, double* out, const double* solution, int64_t n) {
    const dim3 grid(vectorGrid(n)), block(kReduceBlock);
    if (inScale == 1.0) {
        hipLaunchKernelGGL((vectorKernel<double>), grid, block, 0, ctx->stream, int(kAdd), n, out, solution, outScale);
    } else if (inScale == 0.0) {
        hipLaunchKernelGGL((vectorKernel<double>), grid, block, 0, ctx->stream, int(kAssign), n, out, solution, 0.0);
        if (outScale != 1.0)  // x * 1 is x, bit for bit
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
        for (auto& e : marks)
            if (e) (void)hipEventDestroy(e);
    }
};

constexpr int kChunk = 16;
constexpr int kShortChunk = 4;            // trips per chunk once the residual is near the tolerance
constexpr double kNearTolerance = 100.0;  // "near": within this factor of the relative or absolute tolerance

// Enqueues chunks of iterations until the device state says done. The state
// of chunk k is read back while chunk k + 1 is already queued, so the device
// never idles on a poll; kernels of a finished solve return immediately.
template <class Enqueue>
int runLoop(cie_hip_context ctx, Timer& timer, Enqueue&& enqueueIteration, double* seconds) {
    SolverState* pinned = ctx->scratch.pinnedState;
    int slot = 0;
    bool pending = false;
    int chunk = kChunk;
    for (;;) {
        for (int i = 0; i < chunk; ++i) {
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
            const SolverState& state = pinned[previous];
            if (state.done) {
                float ms = 0.0f;
                CIE_HIP_CHECK(hipEventElapsedTime(&ms, timer.start, timer.marks[previous]));
                *seconds = ms * 1e-3;
                CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
                return CIE_HIP_OK;
            }
            // Near the tolerance, shorter chunks: fewer trips are queued past convergence (they only exit,
            // but each still launches its grids). The state is device-decided, so results do not change.
            if (state.relative < kNearTolerance * state.relativeTolerance ||
                state.absolute < kNearTolerance * state.absoluteTolerance)
                chunk = kShortChunk;
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
    bool rowSumCorrection;  // mixed: apply the FP32 plane's row-sum correction in the product
};

int solve(cie_hip_context ctx, cie_hip_csr m, const Request& q, cie_hip_solve_info* info) {
    if (!ctx || !m) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "pcg: null handle");
    CIE_HIP_REQUIRE_CONTEXT(ctx, m->context, "pcg");
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
    status = matrixWorkspace(m, 2 * vec64 + (q.mixed ? 4 : 3) * vecR, &workspace);
    if (status) return status;
    char* base = static_cast<char*>(workspace);
    double* solution = reinterpret_cast<double*>(base);
    double* temp64 = reinterpret_cast<double*>(base + vec64);
    void* r = base + 2 * vec64;
    void* p = base + 2 * vec64 + vecR;
    void* qv = base + 2 * vec64 + 2 * vecR;
    float* correction = q.mixed ? reinterpret_cast<float*>(base + 2 * vec64 + 3 * vecR) : nullptr;

    SolverState init{};
    init.absoluteTolerance = q.settings.absoluteResidual;
    init.relativeTolerance = q.settings.relativeResidual;
    init.maxIterations = q.settings.iterationCount;
    init.updateThreshold = q.refinementFactor * q.refinementFactor;
    init.restart = q.restart;
    init.history = ctx->history;
    init.historyCapacity = uint64_t(ctx->historyCapacity);
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

    const dim3 vgrid(vectorBlocks(n)), vblock(kVectorThreads);
    const int updateCount = int(vgrid.x);
    double seconds = 0.0;
    ProductProfile profile;
    int trip = 0;  // iterations enqueued; its parity picks the rho slot
    if (!q.mixed) {
        double* r64 = static_cast<double*>(r);
        double* p64 = static_cast<double*>(p);
        double* q64 = static_cast<double*>(qv);
        status = launchStart<double>(m, solution, q.b, r64, m->inverseDiagonal64, p64);
        if (status) return status;
        status = runLoop(ctx, timer, [&]() -> int {
            const int slot = trip++ & 1;
            PartialSource sums{};
            int s = launchProduct<double>(m, m->entries64, m->launch64, p64, q64, profile, &sums);
            if (s) return s;
            hipLaunchKernelGGL((updateKernel<double, double>), vgrid, vblock, 0, ctx->stream, slot, n, solution, r64,
                               p64, q64, m->inverseDiagonal64, sums.count, sums.partials, ctx->scratch.vectorPartials,
                               ctx->scratch.state);
            hipLaunchKernelGGL((directionKernel<double, false>), vgrid, vblock, 0, ctx->stream, slot, n, p64, r64,
                               m->inverseDiagonal64, updateCount, ctx->scratch.vectorPartials, ctx->scratch.state);
            return CIE_HIP_OK;
        }, &seconds);
    } else {
        float* r32 = static_cast<float*>(r);
        float* p32 = static_cast<float*>(p);
        float* q32 = static_cast<float*>(qv);
        status = launchStart<float>(m, solution, q.b, r32, m->inverseDiagonal32, p32);
        if (status) return status;
        hipLaunchKernelGGL((vectorKernel<float>), dim3(vectorGrid(n)), dim3(kReduceBlock), 0, ctx->stream, int(kFill),
                           n, correction, static_cast<const float*>(nullptr), 0.0f);
        // One trip: q = A32 p (+ c p); the product's level sums with, after a
        // trip that folded, the FP64 true residual replacing r; the update
        // (FP64 convergence test first when r was replaced); the direction.
        status = runLoop(ctx, timer, [&]() -> int {
            const int slot = trip++ & 1;
            int products = 0;
            int s = launchProduct<float>(m, m->entries32, m->launch32, p32, q32, profile, nullptr,
                                         q.rowSumCorrection ? m->rowCorrection32 : nullptr, &products);
            if (s) return s;
            PartialSource sums{};
            int refineCount = 0;
            s = launchLevelRefine(m, slot ^ 1, products, solution, q.b, r32, &sums, &refineCount);
            if (s) return s;
            hipLaunchKernelGGL(mixedUpdateKernel, vgrid, vblock, 0, ctx->stream, slot, n, solution, correction, r32,
                               p32, q32, m->inverseDiagonal32, sums.count, sums.partials, refineCount,
                               ctx->scratch.refinePartials, ctx->scratch.vectorPartials, ctx->scratch.state);
            hipLaunchKernelGGL((directionKernel<float, true>), vgrid, vblock, 0, ctx->stream, slot, n, p32, r32,
                               m->inverseDiagonal32, updateCount, ctx->scratch.vectorPartials, ctx->scratch.state);
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
    // Loop trips that ran the product kernel: iteration (+1 for the converged
    // trip; the mixed form detects convergence in the trip after the one that
    // folded, +2); none when the start check ended the solve.
    const uint64_t converging = final.done == kConverged ? (q.mixed ? 2 : 1) : 0;
    const uint64_t trips = final.startExit ? 0 : final.iteration + converging;
    result.spmvCount = trips + 1 + final.refinements;
    result.initialResidual = final.initialNorm;
    result.reportedIterations = final.startExit ? 0 : final.iteration + (final.done == kConverged ? 1 : 0);
    if (ctx->profiling) {
        const int worked = int(std::min<uint64_t>(trips, uint64_t(cie_hip_context_s::kProfileEvents)));
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
        if (q.mixed && final.done == kMaxIterations) {
            // Stopped by the cap: report the FP64 true residual of the returned x, not the FP32 recursion's.
            result.stats.absoluteResidual = norm;
            result.stats.relativeResidual = result.trueRelativeResidual;
        }
    }
    // ConjugateGradients_impl.hpp:100-111 return before the output step of
    // :190-198 when the initial residual already meets the tolerance: x keeps
    // its value whatever inScale and outScale are.
    if (!final.startExit) {
        status = writeOutput(ctx, q.inScale, q.outScale, q.x, solution, n);
        if (status) return status;
    }
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
    Request q{inScale, outScale, b, x, size, *settings, false, 0.0, 0, trueResidual != 0, false};
    return solve(context, matrix, q, info);
}

int cie_hip_pcg_mixed(cie_hip_context context, cie_hip_csr matrix, double inScale, const double* b, double outScale,
                      double* x, int64_t size, const cie_hip_mixed_settings* settings, cie_hip_solve_info* info) {
    if (!settings) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_pcg_mixed: null settings");
    double factor = settings->refinementFactor;
    if (!(factor > 0.0 && factor < 1.0)) factor = CIE_HIP_DEFAULT_REFINEMENT_FACTOR;
    Request q{inScale, outScale, b, x, size, settings->settings, true, factor, settings->restart ? 1 : 0, true,
              (settings->flags & CIE_HIP_MIXED_PLAIN_FP32) == 0};
    return solve(context, matrix, q, info);
}

}  // extern "C"

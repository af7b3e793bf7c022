// Device kernels of libcie_hip (CDNA4 / gfx950, wave64).
//
// Determinism: every sum below has a fixed association order that depends
// only on the problem size and the launch shape, never on scheduling:
//   * a CSR row is split over L lanes; lane l sums entries l, l+L, l+2L, ...
//     in increasing order, then an xor-butterfly combines the L lane sums
//     (every lane of the group ends with the same bits, since a+b == b+a);
//   * a grid reduction writes one partial per block (in-block order fixed),
//     and the last block to arrive sums the partials in index order. Only an
//     integer ticket counter is updated atomically, never a value.
#pragma once

#include "common.hpp"

#include <cfloat>

namespace ciehip {

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

// After thread 0 has stored this block's partials: returns true (in every
// thread) in the last block to arrive, which may then read all partials.
__device__ __forceinline__ bool arriveLast(unsigned* counter) {
    __shared__ int last;
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned ticket = atomicAdd(counter, 1u);
        last = ticket == gridDim.x - 1;
    }
    __syncthreads();
    if (last) __threadfence();
    return last;
}

__device__ __forceinline__ void resetCounter(unsigned* counter) {
    if (threadIdx.x == 0) atomicExch(counter, 0u);
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

// Blocks are dealt to the 8 XCDs round-robin; give each XCD a contiguous run
// of chunks so that its L2 sees a contiguous window of the gathered vector.
__device__ __forceinline__ int xcdChunk() {
    const int grid = gridDim.x;
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
                column[u] = __builtin_nontemporal_load(columnIndices + k);
                entry[u] = __builtin_nontemporal_load(entries + k);
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

struct RowChunks {
    int rowBegin, rowEnd;
};

__device__ __forceinline__ RowChunks rowChunk(int rows, int rowsPerChunk, int chunk) {
    const int begin = min(rows, chunk * rowsPerChunk);
    return {begin, min(rows, begin + rowsPerChunk)};
}

// out = inScale * out + outScale * A * in
template <class TV, class TM, int L, int U, int BS>
__global__ __launch_bounds__(BS) void csrProductKernel(int rows, int rowsPerChunk, const int32_t* __restrict__ rowExtents,
                                                       const int32_t* __restrict__ columnIndices,
                                                       const TM* __restrict__ entries, const TV* __restrict__ in,
                                                       TV* __restrict__ out, TV inScale, TV outScale) {
    const RowChunks range = rowChunk(rows, rowsPerChunk, xcdChunk());
    const int group = threadIdx.x / L;
    const int lane = threadIdx.x % L;
    for (int row0 = range.rowBegin; row0 < range.rowEnd; row0 += BS / L) {
        const int row = row0 + group;
        if (row < range.rowEnd) {
            const TV sum = rowProduct<TV, TM, L, U>(row, lane, rowExtents, columnIndices, entries, in);
            if (lane == 0) out[row] = inScale == TV(0) ? outScale * sum : inScale * out[row] + outScale * sum;
        }
    }
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

// Two-stage deterministic inner product; the result lands in *result (device).
template <class T>
__global__ __launch_bounds__(kReduceBlock) void innerProductKernel(int64_t n, const T* __restrict__ left,
                                                                   const T* __restrict__ right, T* partials,
                                                                   unsigned* counter, T* result) {
    __shared__ T lds[kReduceBlock / kWave];
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    T sum = T(0);
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) sum += left[i] * right[i];
    sum = blockSum<kReduceBlock>(sum, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = sum;
    if (!arriveLast(counter)) return;
    const T total = sumPartials<kReduceBlock>(partials, gridDim.x, 1, lds);
    if (threadIdx.x == 0) {
        *result = total;
        resetCounter(counter);
    }
}

// makeDiagonalOperator: inverse = 1 / a_ii; the first failing row (zero or
// missing diagonal) is recorded with atomicMin in failure[0] / failure[1].
template <class T>
__global__ __launch_bounds__(kReduceBlock) void inverseDiagonalKernel(int rows, const int32_t* __restrict__ rowExtents,
                                                                      const int32_t* __restrict__ columnIndices,
                                                                      const double* __restrict__ entries,
                                                                      T* __restrict__ inverse,
                                                                      unsigned long long* failure) {
    const int stride = gridDim.x * blockDim.x;
    for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < rows; row += stride) {
        // Columns are sorted: binary search for the diagonal.
        int lo = rowExtents[row], hi = rowExtents[row + 1];
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (columnIndices[mid] < row) lo = mid + 1; else hi = mid;
        }
        const bool found = lo < rowExtents[row + 1] && columnIndices[lo] == row;
        const double diagonal = found ? entries[lo] : 0.0;
        if (!found) atomicMin(failure + 1, (unsigned long long)row);
        else if (diagonal == 0.0) atomicMin(failure + 0, (unsigned long long)row);
        inverse[row] = found && diagonal != 0.0 ? T(1.0 / diagonal) : T(0);
    }
}

template <class TS, class TD>
__global__ __launch_bounds__(kReduceBlock) void convertKernel(int64_t n, const TS* __restrict__ source,
                                                              TD* __restrict__ target) {
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) target[i] = TD(source[i]);
}

}  // namespace ciehip

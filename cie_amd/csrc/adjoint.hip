// Adjoint gradient of a stiffness matrix assembled from per-element
// conductivities, K(k) = sum_e k_e P_e^T K_e0 P_e (the element integrand is
// linear in the cell modulus: integrands/impl/LinearIsotropicStiffnessIntegrand_impl.hpp:74,
// scattered by graph/impl/Assembler_impl.hpp:156-194). For J(u(k)) with
// K u = f, the adjoint lambda solves K^T lambda = dJ/du and
//     dJ/dk_e = -lambda_e^T K_e0 u_e,
// which is the transpose gather of the assembly scatter. No counterpart in the
// reference (it has no C++ adjoint).
#include "dispatch.hpp"
#include "kernels.hpp"

using namespace ciehip;

namespace {

constexpr int kGradientBlock = 128;
constexpr int kMaxLdsElementMatrix = 64;  // element matrices up to 64 x 64 are staged in LDS

// One thread per element; the element matrix is read in the same order by all
// threads (LDS broadcast when it fits). Sum order: for a, for b -- the oracle's.
template <bool kStaged>
__global__ __launch_bounds__(kGradientBlock) void elementGradientKernel(
    int64_t elements, int nodes, const int32_t* __restrict__ dofs, const double* __restrict__ elementMatrix,
    const double* __restrict__ scale, const double* __restrict__ lambda, const double* __restrict__ u,
    double* __restrict__ gradient) {
    __shared__ double staged[kStaged ? kMaxLdsElementMatrix * kMaxLdsElementMatrix : 1];
    const double* matrix = elementMatrix;
    if (kStaged) {
        for (int i = threadIdx.x; i < nodes * nodes; i += blockDim.x) staged[i] = elementMatrix[i];
        __syncthreads();
        matrix = staged;
    }
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < elements; e += stride) {
        const int32_t* d = dofs + e * nodes;
        double sum = 0.0;
        for (int a = 0; a < nodes; ++a) {
            const int32_t da = d[a];
            if (da < 0) continue;
            double row = 0.0;
            for (int b = 0; b < nodes; ++b) {
                const int32_t db = d[b];
                if (db >= 0) row += matrix[a * nodes + b] * u[db];
            }
            sum += lambda[da] * row;
        }
        const double value = -sum;
        gradient[e] = scale ? scale[e] * value : value;
    }
}

// entries64[slots[i, a, b]] += factor[i] * K_e0[a, b] for each listed element;
// one thread per (a, b) pair of one element at a time so that elements that
// share slots are applied in list order.
__global__ __launch_bounds__(kReduceBlock) void addElementMatricesKernel(
    int64_t count, int nodes, const int64_t* __restrict__ slots, const double* __restrict__ elementMatrix,
    const double* __restrict__ factors, double* __restrict__ entries64, float* __restrict__ entries32) {
    const int pairs = nodes * nodes;
    for (int64_t i = 0; i < count; ++i) {
        for (int k = threadIdx.x; k < pairs; k += blockDim.x) {
            const int64_t slot = slots[i * pairs + k];
            if (slot < 0) continue;
            const double value = entries64[slot] + factors[i] * elementMatrix[k];
            entries64[slot] = value;
            if (entries32) entries32[slot] = float(value);
        }
        __syncthreads();
    }
}

}  // namespace

extern "C" {

int cie_hip_element_gradient(cie_hip_context context, int64_t elements, int32_t nodesPerElement, const int32_t* dofs,
                             const double* elementMatrix, const double* scale, const double* lambda, const double* u,
                             double* gradient) {
    if (!context || elements < 0 || nodesPerElement < 1 ||
        (elements && (!dofs || !elementMatrix || !lambda || !u || !gradient)))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_element_gradient: invalid arguments");
    if (!elements) return CIE_HIP_OK;
    CIE_HIP_CHECK(hipSetDevice(context->device));
    const int64_t blocks = std::min<int64_t>(4096, (elements + kGradientBlock - 1) / kGradientBlock);
    if (nodesPerElement <= kMaxLdsElementMatrix)
        hipLaunchKernelGGL((elementGradientKernel<true>), dim3(blocks), dim3(kGradientBlock), 0, context->stream,
                           elements, nodesPerElement, dofs, elementMatrix, scale, lambda, u, gradient);
    else
        hipLaunchKernelGGL((elementGradientKernel<false>), dim3(blocks), dim3(kGradientBlock), 0, context->stream,
                           elements, nodesPerElement, dofs, elementMatrix, scale, lambda, u, gradient);
    CIE_HIP_CHECK(hipGetLastError());
    CIE_HIP_CHECK(hipStreamSynchronize(context->stream));
    return CIE_HIP_OK;
}

int cie_hip_csr_add_element_matrices(cie_hip_csr matrix, int64_t count, int32_t nodesPerElement, const int64_t* slots,
                                     const double* elementMatrix, const double* factors) {
    if (!matrix || count < 0 || nodesPerElement < 1 || (count && (!slots || !elementMatrix || !factors)))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_csr_add_element_matrices: invalid arguments");
    if (!matrix->ownsEntries64)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT,
                        "cie_hip_csr_add_element_matrices: the FP64 entries are borrowed, not owned");
    if (!count) return CIE_HIP_OK;
    cie_hip_context ctx = matrix->context;
    CIE_HIP_CHECK(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(addElementMatricesKernel, dim3(1), dim3(kReduceBlock), 0, ctx->stream, count, nodesPerElement,
                       slots, elementMatrix, factors, matrix->entries64, matrix->entries32);
    CIE_HIP_CHECK(hipGetLastError());
    CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return entriesChanged(matrix);  // the Jacobi diagonal and the column-block values are rebuilt on next use
}

}  // extern "C"

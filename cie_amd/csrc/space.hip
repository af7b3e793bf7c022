// LinalgSpaceLike operations on device vectors (solvers/inc/LinalgSpace.hpp:10-35),
// following DefaultSpace.cpp:50-187 / SYCLSpace.cpp:114-313, and the Jacobi
// DiagonalOperator::product (DiagonalOperator.cpp:37-53).
#include "dispatch.hpp"
#include "kernels.hpp"

#include <mutex>
#include <vector>

using namespace ciehip;

namespace {

// The no-allocation dot product's block partials: one array per device, held
// until the process ends. 2048 doubles is the largest grid vectorGrid gives
// (16 KiB). In host code `scratch` only names the symbol; its device address
// comes from hipGetSymbolAddress, once per device (staticScratchAddress).
constexpr int kStaticPartials = 2048;
static __device__ __align__(256) double scratch[kStaticPartials];

std::mutex staticScratchMutex;
std::vector<double*> staticScratchAddresses;  // by device ordinal; null until looked up

}  // namespace

namespace ciehip {

// Looks up this device's copy of `scratch` (the device must be current) and
// keeps the address; cie_hip_context_create calls it once per device.
int staticScratchAddress(int device, double** address) {
    std::lock_guard<std::mutex> lock(staticScratchMutex);
    if (device < 0) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "static scratch: invalid device");
    if (size_t(device) >= staticScratchAddresses.size()) staticScratchAddresses.resize(size_t(device) + 1, nullptr);
    if (!staticScratchAddresses[size_t(device)]) {
        void* pointer = nullptr;
        const hipError_t found = hipGetSymbolAddress(&pointer, HIP_SYMBOL(scratch));
        if (found != hipSuccess || !pointer)
            return setError(CIE_HIP_ERROR_HIP, std::string("static scratch: hipGetSymbolAddress: ") +
                                                   hipGetErrorString(found));
        staticScratchAddresses[size_t(device)] = static_cast<double*>(pointer);
    }
    if (address) *address = staticScratchAddresses[size_t(device)];
    return CIE_HIP_OK;
}

}  // namespace ciehip

namespace {

template <class T>
int vectorOp(cie_hip_context ctx, int op, T* target, const T* source, int64_t size, T value, const char* name) {
    if (!ctx || size < 0 || (size && !target) || (size && (op == kScale || op == kAdd || op == kAssign) && !source))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, std::string(name) + ": invalid arguments");
    if (!size) return CIE_HIP_OK;
    CIE_HIP_CHECK(hipSetDevice(ctx->device));
    hipLaunchKernelGGL((vectorKernel<T>), dim3(vectorGrid(size)), dim3(kReduceBlock), 0, ctx->stream, op, size, target,
                       source, value);
    CIE_HIP_CHECK(hipGetLastError());
    return CIE_HIP_OK;
}

template <class T>
int innerProduct(cie_hip_context ctx, const T* left, const T* right, int64_t size, T* result, const char* name) {
    if (!ctx || size < 0 || !result || (size && (!left || !right)))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, std::string(name) + ": invalid arguments");
    if (!size) {
        *result = T(0);
        return CIE_HIP_OK;
    }
    CIE_HIP_CHECK(hipSetDevice(ctx->device));
    T* partials = reinterpret_cast<T*>(ctx->scratch.partials);
    T* deviceResult = reinterpret_cast<T*>(ctx->scratch.scalars);
    const int grid = vectorGrid(size);
    hipLaunchKernelGGL((innerProductKernel<T>), dim3(grid), dim3(kReduceBlock), 0, ctx->stream, size, left, right,
                       partials);
    hipLaunchKernelGGL((innerProductFinishKernel<T>), dim3(1), dim3(kReduceBlock), 0, ctx->stream, grid, partials,
                       deviceResult);
    CIE_HIP_CHECK(hipGetLastError());
    T* pinned = reinterpret_cast<T*>(ctx->scratch.pinned);
    CIE_HIP_CHECK(hipMemcpyAsync(pinned, deviceResult, sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    *result = *pinned;
    return CIE_HIP_OK;
}

template <class T>
int diagonalProduct(cie_hip_context ctx, const T* inverse, T inScale, const T* in, T outScale, T* out, int64_t size,
                    const char* name) {
    if (!ctx || size < 0 || (size && (!inverse || !in || !out)))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, std::string(name) + ": invalid arguments");
    if (!size) return CIE_HIP_OK;
    CIE_HIP_CHECK(hipSetDevice(ctx->device));
    hipLaunchKernelGGL((diagonalProductKernel<T>), dim3(vectorGrid(size)), dim3(kReduceBlock), 0, ctx->stream, size,
                       inverse, inScale, in, outScale, out);
    CIE_HIP_CHECK(hipGetLastError());
    return CIE_HIP_OK;
}

}  // namespace

extern "C" {

int cie_hip_inner_product_f64(cie_hip_context c, const double* l, const double* r, int64_t n, double* result) {
    return innerProduct<double>(c, l, r, n, result, "cie_hip_inner_product_f64");
}

int cie_hip_inner_product_f64_static(cie_hip_context ctx, const double* left, const double* right, int64_t size,
                                     double* result) {
    if (!ctx || size < 0 || !result || (size && (!left || !right)))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_inner_product_f64_static: invalid arguments");
    const int grid = size ? vectorGrid(size) : 1;
    if (grid > kStaticPartials)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT,
                        "cie_hip_inner_product_f64_static: " + std::to_string(grid) + " block partials exceed the " +
                            std::to_string(kStaticPartials) + " of the static scratch");
    double* partials = nullptr;
    {
        std::lock_guard<std::mutex> lock(staticScratchMutex);
        if (size_t(ctx->device) < staticScratchAddresses.size()) partials = staticScratchAddresses[size_t(ctx->device)];
    }
    if (!partials)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT,
                        "cie_hip_inner_product_f64_static: no static scratch address for this device");
    CIE_HIP_CHECK(hipSetDevice(ctx->device));
    if (size) {
        hipLaunchKernelGGL((innerProductKernel<double>), dim3(grid), dim3(kReduceBlock), 0, ctx->stream, size, left,
                           right, partials);
        hipLaunchKernelGGL((innerProductFinishKernel<double>), dim3(1), dim3(kReduceBlock), 0, ctx->stream, grid,
                           static_cast<const double*>(partials), result);
    } else {
        // The empty sum: the finish kernel over zero partials writes +0 (a kernel, not a memset).
        hipLaunchKernelGGL((innerProductFinishKernel<double>), dim3(1), dim3(kReduceBlock), 0, ctx->stream, 0,
                           static_cast<const double*>(partials), result);
    }
    CIE_HIP_CHECK(hipGetLastError());
    return CIE_HIP_OK;
}
int cie_hip_scale_f64(cie_hip_context c, double* t, const double* s, int64_t n, double a) {
    return vectorOp<double>(c, kScale, t, s, n, a, "cie_hip_scale_f64");
}
int cie_hip_scale_uniform_f64(cie_hip_context c, double* t, int64_t n, double a) {
    return vectorOp<double>(c, kScaleUniform, t, nullptr, n, a, "cie_hip_scale_uniform_f64");
}
int cie_hip_add_f64(cie_hip_context c, double* t, const double* s, int64_t n, double a) {
    return vectorOp<double>(c, kAdd, t, s, n, a, "cie_hip_add_f64");
}
int cie_hip_assign_f64(cie_hip_context c, double* t, const double* s, int64_t n) {
    return vectorOp<double>(c, kAssign, t, s, n, 0.0, "cie_hip_assign_f64");
}
int cie_hip_fill_f64(cie_hip_context c, double* t, int64_t n, double v) {
    return vectorOp<double>(c, kFill, t, nullptr, n, v, "cie_hip_fill_f64");
}
int cie_hip_inner_product_f32(cie_hip_context c, const float* l, const float* r, int64_t n, float* result) {
    return innerProduct<float>(c, l, r, n, result, "cie_hip_inner_product_f32");
}
int cie_hip_scale_f32(cie_hip_context c, float* t, const float* s, int64_t n, float a) {
    return vectorOp<float>(c, kScale, t, s, n, a, "cie_hip_scale_f32");
}
int cie_hip_scale_uniform_f32(cie_hip_context c, float* t, int64_t n, float a) {
    return vectorOp<float>(c, kScaleUniform, t, nullptr, n, a, "cie_hip_scale_uniform_f32");
}
int cie_hip_add_f32(cie_hip_context c, float* t, const float* s, int64_t n, float a) {
    return vectorOp<float>(c, kAdd, t, s, n, a, "cie_hip_add_f32");
}
int cie_hip_assign_f32(cie_hip_context c, float* t, const float* s, int64_t n) {
    return vectorOp<float>(c, kAssign, t, s, n, 0.0f, "cie_hip_assign_f32");
}
int cie_hip_fill_f32(cie_hip_context c, float* t, int64_t n, float v) {
    return vectorOp<float>(c, kFill, t, nullptr, n, v, "cie_hip_fill_f32");
}
int cie_hip_diagonal_product_f64(cie_hip_context c, const double* inv, double inScale, const double* in, double outScale,
                                 double* out, int64_t n) {
    return diagonalProduct<double>(c, inv, inScale, in, outScale, out, n, "cie_hip_diagonal_product_f64");
}
int cie_hip_diagonal_product_f32(cie_hip_context c, const float* inv, float inScale, const float* in, float outScale,
                                 float* out, int64_t n) {
    return diagonalProduct<float>(c, inv, inScale, in, outScale, out, n, "cie_hip_diagonal_product_f32");
}

}  // extern "C"

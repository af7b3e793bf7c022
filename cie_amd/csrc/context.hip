// Contexts, device memory and error reporting of the C ABI.
#include "common.hpp"

#include <cstring>

namespace ciehip {

namespace {
thread_local std::string g_lastError;
}

int setError(int status, const std::string& message) {
    g_lastError = message;
    return status;
}

int ensurePartials(cie_hip_context ctx, int slots) {
    Scratch& s = ctx->scratch;
    if (slots <= s.partialSlots) return CIE_HIP_OK;
    CIE_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    double* grown = nullptr;
    if (hipMalloc(&grown, sizeof(double) * size_t(slots) * 4) != hipSuccess)
        return setError(CIE_HIP_ERROR_OUT_OF_MEMORY, "partials allocation failed");
    CIE_HIP_CHECK(hipFree(s.partials));
    s.partials = grown;
    s.partialSlots = slots;
    return CIE_HIP_OK;
}

}  // namespace ciehip

using ciehip::setError;

extern "C" {

const char* cie_hip_last_error(void) { return ciehip::g_lastError.c_str(); }

const char* cie_hip_status_string(int status) {
    switch (status) {
    case CIE_HIP_OK: return "ok";
    case CIE_HIP_ERROR_INVALID_ARGUMENT: return "invalid argument";
    case CIE_HIP_ERROR_SHAPE: return "incompatible sizes";
    case CIE_HIP_ERROR_NOT_SPD: return "p^T A p vanished (matrix not positive definite)";
    case CIE_HIP_ERROR_BREAKDOWN: return "preconditioned norm vanished (r^T z <= 0)";
    case CIE_HIP_ERROR_ZERO_DIAGONAL: return "explicit zero on the diagonal";
    case CIE_HIP_ERROR_MISSING_DIAGONAL: return "missing diagonal entry";
    case CIE_HIP_ERROR_HIP: return "HIP runtime error";
    case CIE_HIP_ERROR_OUT_OF_MEMORY: return "out of device memory";
    case CIE_HIP_ERROR_NO_DEVICE: return "no HIP device";
    case CIE_HIP_ERROR_NOT_IMPLEMENTED: return "not implemented for these arguments";
    default: return "unknown status";
    }
}

int cie_hip_device_count(int* count) {
    if (!count) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_device_count: null output");
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess) devices = 0;
    *count = devices;
    return CIE_HIP_OK;
}

int cie_hip_context_create(int device, void* stream, cie_hip_context* context) {
    if (!context) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_context_create: null output");
    *context = nullptr;
    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0)
        return setError(CIE_HIP_ERROR_NO_DEVICE, "cie_hip_context_create: no HIP device visible");
    if (device < 0 || device >= devices)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_context_create: device index out of range");
    CIE_HIP_CHECK(hipSetDevice(device));
    // Set-up of this device's static scratch (the no-allocation dot product's
    // partials): its address, looked up once per device with the device current.
    if (const int status = ciehip::staticScratchAddress(device, nullptr)) return status;
    auto* ctx = new cie_hip_context_s;
    ctx->device = device;
    hipDeviceProp_t properties;
    if (hipGetDeviceProperties(&properties, device) != hipSuccess) {
        delete ctx;
        return setError(CIE_HIP_ERROR_HIP, "cie_hip_context_create: hipGetDeviceProperties failed");
    }
    ctx->computeUnits = properties.multiProcessorCount;
    if (stream) {
        ctx->stream = static_cast<hipStream_t>(stream);
    } else {
        if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
            delete ctx;
            return setError(CIE_HIP_ERROR_HIP, "cie_hip_context_create: hipStreamCreate failed");
        }
        ctx->ownsStream = true;
    }
    auto& s = ctx->scratch;
    s.partialSlots = ciehip::kMaxPartials;
    bool ok = hipMalloc(&s.partials, sizeof(double) * ciehip::kMaxPartials * 4) == hipSuccess
           && hipMalloc(&s.level, sizeof(double) * ciehip::kLevelBlocks * 2) == hipSuccess
           && hipMalloc(&s.vectorPartials, sizeof(double) * ciehip::kVectorPartials * 2) == hipSuccess
           && hipMalloc(&s.refinePartials, sizeof(double) * ciehip::kMaxPartials * 2) == hipSuccess
           && hipMalloc(&s.scalars, sizeof(double) * 64) == hipSuccess
           && hipMalloc(&s.state, sizeof(ciehip::SolverState)) == hipSuccess
           && hipHostMalloc(&s.pinned, sizeof(double) * 64) == hipSuccess
           && hipHostMalloc(&s.pinnedState, 2 * sizeof(ciehip::SolverState)) == hipSuccess;
    if (!ok) {
        cie_hip_context_destroy(ctx);
        return setError(CIE_HIP_ERROR_OUT_OF_MEMORY, "cie_hip_context_create: scratch allocation failed");
    }
    *context = ctx;
    return CIE_HIP_OK;
}

}  // extern "C"

namespace {

void destroyContext(cie_hip_context context) {
    (void)hipSetDevice(context->device);
    if (context->stream) (void)hipStreamSynchronize(context->stream);
    auto& s = context->scratch;
    if (s.partials) (void)hipFree(s.partials);
    if (s.level) (void)hipFree(s.level);
    if (s.vectorPartials) (void)hipFree(s.vectorPartials);
    if (s.refinePartials) (void)hipFree(s.refinePartials);
    if (s.scalars) (void)hipFree(s.scalars);
    if (s.state) (void)hipFree(s.state);
    if (s.pinned) (void)hipHostFree(s.pinned);
    if (s.pinnedState) (void)hipHostFree(s.pinnedState);
    for (hipEvent_t& event : context->profileEvents)
        if (event) (void)hipEventDestroy(event);
    if (context->ownsStream && context->stream) (void)hipStreamDestroy(context->stream);
    delete context;
}

}  // namespace

namespace ciehip {

void holdContext(cie_hip_context context) { ++context->holds.holders; }

void dropContext(cie_hip_context context) {
    if (--context->holds.holders == 0 && context->holds.destroyRequested) destroyContext(context);
}

}  // namespace ciehip

extern "C" {

int cie_hip_context_destroy(cie_hip_context context) {
    if (!context || context->holds.destroyRequested) return CIE_HIP_OK;
    if (context->holds.holders > 0) {
        context->holds.destroyRequested = true;  // released with the last object created on it
        return CIE_HIP_OK;
    }
    destroyContext(context);
    return CIE_HIP_OK;
}

int cie_hip_context_synchronize(cie_hip_context context) {
    if (!context) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_context_synchronize: null context");
    CIE_HIP_CHECK(hipStreamSynchronize(context->stream));
    return CIE_HIP_OK;
}

int cie_hip_context_stream(cie_hip_context context, void** stream) {
    if (!context || !stream) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_context_stream: null argument");
    *stream = context->stream;
    return CIE_HIP_OK;
}

int cie_hip_context_set_profiling(cie_hip_context context, int enable) {
    if (!context) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_context_set_profiling: null context");
    if (enable) {
        CIE_HIP_CHECK(hipSetDevice(context->device));
        for (hipEvent_t& event : context->profileEvents)
            if (!event) CIE_HIP_CHECK(hipEventCreate(&event));
    }
    context->profiling = enable != 0;
    return CIE_HIP_OK;
}

int cie_hip_context_set_residual_history(cie_hip_context context, double* history, int64_t capacity) {
    if (!context) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_context_set_residual_history: null context");
    if (history && capacity < 0)
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_context_set_residual_history: negative capacity");
    context->history = history && capacity > 0 ? history : nullptr;
    context->historyCapacity = context->history ? capacity : 0;
    return CIE_HIP_OK;
}

int cie_hip_context_kernel_seconds(cie_hip_context context, double* seconds) {
    if (!context || !seconds) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_context_kernel_seconds: null argument");
    *seconds = context->kernelSeconds;
    return CIE_HIP_OK;
}

int cie_hip_malloc(cie_hip_context context, size_t bytes, void** pointer) {
    if (!context || !pointer) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_malloc: null argument");
    *pointer = nullptr;
    CIE_HIP_CHECK(hipSetDevice(context->device));
    CIE_HIP_CHECK(hipMalloc(pointer, bytes ? bytes : 1));
    return CIE_HIP_OK;
}

int cie_hip_free(cie_hip_context context, void* pointer) {
    if (!context) return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_free: null context");
    if (!pointer) return CIE_HIP_OK;
    CIE_HIP_CHECK(hipStreamSynchronize(context->stream));
    CIE_HIP_CHECK(hipFree(pointer));
    return CIE_HIP_OK;
}

int cie_hip_memcpy(cie_hip_context context, void* destination, const void* source, size_t bytes, int kind) {
    if (!context || (bytes && (!destination || !source)))
        return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_memcpy: null argument");
    if (!bytes) return CIE_HIP_OK;
    hipMemcpyKind direction;
    switch (kind) {
    case 0: direction = hipMemcpyHostToDevice; break;
    case 1: direction = hipMemcpyDeviceToHost; break;
    case 2: direction = hipMemcpyDeviceToDevice; break;
    default: return setError(CIE_HIP_ERROR_INVALID_ARGUMENT, "cie_hip_memcpy: kind must be 0, 1 or 2");
    }
    CIE_HIP_CHECK(hipMemcpyAsync(destination, source, bytes, direction, context->stream));
    CIE_HIP_CHECK(hipStreamSynchronize(context->stream));
    return CIE_HIP_OK;
}

}  // extern "C"

// fg_runtime.hip -- library state, error reporting, parameter upload and the
// device-memory/timing helpers of the C-ABI.
#include <cstdarg>
#include <mutex>

#include "fg_internal.hpp"

namespace fg {

State& state() {
    static State s;
    return s;
}

namespace {
thread_local int t_err = FG_OK;
thread_local char t_msg[512] = "";
}  // namespace

void set_error(int code, const char* fmt, ...) {
    t_err = code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(t_msg, sizeof(t_msg), fmt, ap);
    va_end(ap);
}

bool check_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    set_error(e == hipErrorOutOfMemory ? FG_ERR_OUT_OF_MEMORY : FG_ERR_HIP, "%s: %s", what,
              hipGetErrorString(e));
    return false;
}

bool check_launch(const char* what) { return check_hip(hipGetLastError(), what); }

// Fill kernel used instead of hipMemsetAsync so that captured graphs hold
// only kernel nodes.
__global__ void k_fill_u32(uint32_t* __restrict__ dst, uint32_t value, uint64_t count) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    // 16-byte stores for the aligned bulk, scalar stores for the rest.
    if (((uintptr_t)dst & 15u) == 0) {
        const uint64_t n4 = count / 4;
        uint4* d4 = reinterpret_cast<uint4*>(dst);
        const uint4 v = make_uint4(value, value, value, value);
        for (uint64_t k = i; k < n4; k += stride) d4[k] = v;
        for (uint64_t k = n4 * 4 + i; k < count; k += stride) dst[k] = value;
    } else {
        for (uint64_t k = i; k < count; k += stride) dst[k] = value;
    }
}

// Streaming copy, 16 bytes per lane: the reference kernel for scaling the
// HBM byte counters (a copy of B bytes reads B and writes B).
__global__ void k_copy_u4(uint4* __restrict__ dst, const uint4* __restrict__ src, uint64_t count) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride) dst[k] = src[k];
}

void launch_fill_u32(uint32_t* dst, uint32_t value, uint64_t count, hipStream_t s) {
    if (count == 0) return;
    uint64_t work = ((uintptr_t)dst & 15u) == 0 ? (count + 3) / 4 : count;
    unsigned blocks = (unsigned)((work + 255) / 256);
    if (blocks > 8192) blocks = 8192;
    if (blocks == 0) blocks = 1;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, s, dst, value, count);
}

}  // namespace fg

using namespace fg;

extern "C" {

const char* fg_version(void) { return "fluidgrid_amd 0.1 (gfx950)"; }

int fg_last_error(void) {
    int e = t_err;
    t_err = FG_OK;
    return e;
}

int fg_peek_error(void) { return t_err; }

void fg_record_error(int code, const char* message) { set_error(code, "%s", message ? message : ""); }

const char* fg_last_error_string(void) { return t_msg; }

int fg_set_stream(void* s) {
    state().stream = reinterpret_cast<hipStream_t>(s);
    return FG_OK;
}

void* fg_get_stream(void) { return reinterpret_cast<void*>(state().stream); }

int fg_set_key_mode(int mode) {
    if (mode != FG_KEY_REF_XOR && mode != FG_KEY_LINEAR) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_set_key_mode: unknown mode %d", mode);
        return FG_ERR_INVALID_ARGUMENT;
    }
    state().key_mode = mode;
    return FG_OK;
}

int fg_get_key_mode(void) { return state().key_mode; }

int fg_set_math_mode(int mode) {
    if (mode != FG_MATH_FAST && mode != FG_MATH_REFERENCE) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_set_math_mode: unknown mode %d", mode);
        return FG_ERR_INVALID_ARGUMENT;
    }
    state().math_mode = mode;
    return FG_OK;
}

int fg_get_math_mode(void) { return state().math_mode; }

int fg_set_tuning(int sphBlock, int sortDigitBits) {
    if (sphBlock != 64 && sphBlock != 128 && sphBlock != 256 && sphBlock != 512) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_set_tuning: SPH block %d not in {64,128,256,512}", sphBlock);
        return FG_ERR_INVALID_ARGUMENT;
    }
    if (sortDigitBits < 6 || sortDigitBits > 8) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_set_tuning: sort digit bits %d outside 6..8", sortDigitBits);
        return FG_ERR_INVALID_ARGUMENT;
    }
    state().sph_block = sphBlock;
    state().sort_digit_bits = sortDigitBits;
    return FG_OK;
}

int fg_set_schedule(int flatDensity, int flatForces) {
    state().flat_density = flatDensity != 0;
    state().flat_forces = flatForces != 0;
    return FG_OK;
}

int fg_set_tiled(int enable) {
    state().tiled = enable != 0;
    return FG_OK;
}

int fg_set_contact_defer(int enable) {
    state().defer_contact = enable != 0;
    return FG_OK;
}

int fg_get_contact_defer(int* enable) {
    if (!enable) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_get_contact_defer: null pointer");
        return FG_ERR_INVALID_ARGUMENT;
    }
    *enable = state().defer_contact ? 1 : 0;
    return FG_OK;
}

int fg_set_precision(int mode) {
    if (mode != FG_PRECISION_FP32 && mode != FG_PRECISION_MIXED) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_set_precision: unknown mode %d", mode);
        return FG_ERR_INVALID_ARGUMENT;
    }
    state().precision = mode;
    return FG_OK;
}

int fg_get_precision(void) { return state().precision; }

int fg_get_tiled(int* enable) {
    if (enable) *enable = state().tiled ? 1 : 0;
    return FG_OK;
}

int fg_get_schedule(int* flatDensity, int* flatForces) {
    if (flatDensity) *flatDensity = state().flat_density ? 1 : 0;
    if (flatForces) *flatForces = state().flat_forces ? 1 : 0;
    return FG_OK;
}

int fg_get_tuning(int* sphBlock, int* sortDigitBits) {
    if (sphBlock) *sphBlock = state().sph_block;
    if (sortDigitBits) *sortDigitBits = state().sort_digit_bits;
    return FG_OK;
}

// K_FluidSim.cu:591-597
void gInitializeCommonFluidParams(fg_common_fluid_params p) {
    state().common = p;
    state().have_common = true;
}

// K_FluidSim.cu:602-609
void gInitializeKernelParams(fg_kernels_cache k) {
    state().kernels = k;
    state().have_kernels = true;
}

int fg_get_common_fluid_params(fg_common_fluid_params* out) {
    if (!out) return FG_ERR_INVALID_ARGUMENT;
    *out = state().common;
    return state().have_common ? FG_OK : FG_ERR_NOT_INITIALIZED;
}

int fg_get_kernels_cache(fg_kernels_cache* out) {
    if (!out) return FG_ERR_INVALID_ARGUMENT;
    *out = state().kernels;
    return state().have_kernels ? FG_OK : FG_ERR_NOT_INITIALIZED;
}

int fg_device_count(int* count) {
    if (!count) return FG_ERR_INVALID_ARGUMENT;
    *count = 0;
    hipError_t e = hipGetDeviceCount(count);
    if (e != hipSuccess) {
        *count = 0;
        (void)hipGetLastError();
        return FG_OK;
    }
    return FG_OK;
}

int fg_device_name(int device, char* buf, size_t len) {
    hipDeviceProp_t prop;
    if (!check_hip(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties")) return t_err;
    snprintf(buf, len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return FG_OK;
}

int fg_malloc(void** ptr, size_t bytes) {
    if (!ptr) return FG_ERR_INVALID_ARGUMENT;
    *ptr = nullptr;
    if (bytes == 0) return FG_OK;
    return check_hip(hipMalloc(ptr, bytes), "hipMalloc") ? FG_OK : t_err;
}

int fg_free(void* ptr) {
    if (!ptr) return FG_OK;
    return check_hip(hipFree(ptr), "hipFree") ? FG_OK : t_err;
}

int fg_memcpy_htod(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return FG_OK;
    return check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, state().stream), "htod") &&
                   check_hip(hipStreamSynchronize(state().stream), "htod sync")
               ? FG_OK
               : t_err;
}

int fg_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return FG_OK;
    return check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, state().stream), "dtoh") &&
                   check_hip(hipStreamSynchronize(state().stream), "dtoh sync")
               ? FG_OK
               : t_err;
}

int fg_memcpy_dtod(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return FG_OK;
    return check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, state().stream), "dtod")
               ? FG_OK
               : t_err;
}

int fg_memset_u32(unsigned int* dst, unsigned int value, size_t count) {
    if (count && !dst) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_memset_u32: null destination");
        return FG_ERR_INVALID_ARGUMENT;
    }
    launch_fill_u32(dst, value, count, state().stream);
    return check_launch("fill") ? FG_OK : t_err;
}

int fg_copy_bytes(void* dst, const void* src, size_t bytes) {
    if (bytes % 16 || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15)) {
        set_error(FG_ERR_INVALID_ARGUMENT, "fg_copy_bytes: size and pointers must be 16-byte aligned");
        return FG_ERR_INVALID_ARGUMENT;
    }
    const uint64_t count = bytes / 16;
    if (count == 0) return FG_OK;
    unsigned blocks = (unsigned)((count + 255) / 256);
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(k_copy_u4, dim3(blocks), dim3(256), 0, state().stream, static_cast<uint4*>(dst),
                       static_cast<const uint4*>(src), count);
    return check_launch("copy") ? FG_OK : t_err;
}

int fg_synchronize(void) {
    return check_hip(hipStreamSynchronize(state().stream), "hipStreamSynchronize") ? FG_OK : t_err;
}

int fg_event_create(void** ev) {
    if (!ev) return FG_ERR_INVALID_ARGUMENT;
    hipEvent_t e;
    if (!check_hip(hipEventCreate(&e), "hipEventCreate")) return t_err;
    *ev = reinterpret_cast<void*>(e);
    return FG_OK;
}

int fg_event_destroy(void* ev) {
    return check_hip(hipEventDestroy(reinterpret_cast<hipEvent_t>(ev)), "hipEventDestroy") ? FG_OK : t_err;
}

int fg_event_record(void* ev) {
    return check_hip(hipEventRecord(reinterpret_cast<hipEvent_t>(ev), state().stream), "hipEventRecord")
               ? FG_OK
               : t_err;
}

int fg_event_record_stream(void* ev, void* stream) {
    return check_hip(hipEventRecord(reinterpret_cast<hipEvent_t>(ev), reinterpret_cast<hipStream_t>(stream)),
                     "hipEventRecord")
               ? FG_OK
               : t_err;
}

int fg_event_elapsed_ms(void* start, void* stop, float* ms) {
    if (!ms) return FG_ERR_INVALID_ARGUMENT;
    if (!check_hip(hipEventSynchronize(reinterpret_cast<hipEvent_t>(stop)), "hipEventSynchronize")) return t_err;
    return check_hip(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start),
                                         reinterpret_cast<hipEvent_t>(stop)),
                     "hipEventElapsedTime")
               ? FG_OK
               : t_err;
}

}  // extern "C"

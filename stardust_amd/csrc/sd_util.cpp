// Error plumbing and thin HIP runtime wrappers of the C ABI.
// Replaces CUDA_ERR_CHECK (src/engine/cuda/cuda_utils.hpp:8-14), which exits the
// process; every failure here is returned as an sd_status with a message.
#include <stdarg.h>
#include <stdio.h>

#include "sd_internal.hpp"

namespace sd {

static thread_local char g_last_error[512] = "no error";

sd_status set_error(sd_status code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

sd_status hip_error(hipError_t e, const char* what, const char* file, int line) {
    return set_error(SD_ERR_HIP, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
}

}  // namespace sd

extern "C" const char* sd_last_error(void) { return sd::g_last_error; }

extern "C" const char* sd_version(void) { return "stardust-amd 0.1 (gfx950)"; }

extern "C" sd_status sd_device_count(int* count) {
    SD_REQUIRE(count != nullptr, "sd_device_count: NULL");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e == hipErrorNoDevice) {
        *count = 0;
        return SD_OK;
    }
    SD_HIP(e);
    *count = c;
    return SD_OK;
}

extern "C" sd_status sd_malloc(void** dptr, size_t bytes) {
    SD_REQUIRE(dptr != nullptr, "sd_malloc: NULL out pointer");
    *dptr = nullptr;
    if (bytes == 0) bytes = 16;
    SD_HIP(hipMalloc(dptr, bytes));
    return SD_OK;
}

extern "C" sd_status sd_free(void* dptr) {
    if (dptr) SD_HIP(hipFree(dptr));
    return SD_OK;
}

extern "C" sd_status sd_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return SD_OK;
    SD_REQUIRE(dst && src, "sd_memcpy_htod: NULL pointer");
    SD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    SD_HIP(hipStreamSynchronize((hipStream_t)stream));
    return SD_OK;
}

extern "C" sd_status sd_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return SD_OK;
    SD_REQUIRE(dst && src, "sd_memcpy_dtoh: NULL pointer");
    SD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    SD_HIP(hipStreamSynchronize((hipStream_t)stream));
    return SD_OK;
}

extern "C" sd_status sd_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return SD_OK;
    SD_REQUIRE(dst && src, "sd_memcpy_dtod: NULL pointer");
    // a plain copy kernel (graph-safe; also the known-size reference for HBM counters)
    if (bytes % 4 == 0 && ((uintptr_t)dst % 4) == 0 && ((uintptr_t)src % 4) == 0)
        return sd::launch_copy_u32((uint32_t*)dst, (const uint32_t*)src, bytes / 4, (hipStream_t)stream);
    SD_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return SD_OK;
}

extern "C" sd_status sd_memset_u32(void* dst, uint32_t value, size_t count, void* stream) {
    if (count == 0) return SD_OK;
    SD_REQUIRE(dst, "sd_memset_u32: NULL pointer");
    return sd::launch_fill_u32((uint32_t*)dst, value, count, (hipStream_t)stream);
}

extern "C" sd_status sd_stream_create(void** stream) {
    SD_REQUIRE(stream, "sd_stream_create: NULL");
    hipStream_t s;
    SD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
    return SD_OK;
}

extern "C" sd_status sd_stream_destroy(void* stream) {
    if (stream) SD_HIP(hipStreamDestroy((hipStream_t)stream));
    return SD_OK;
}

extern "C" sd_status sd_stream_synchronize(void* stream) {
    SD_HIP(hipStreamSynchronize((hipStream_t)stream));
    return SD_OK;
}

extern "C" sd_status sd_event_create(void** ev) {
    SD_REQUIRE(ev, "sd_event_create: NULL");
    hipEvent_t e;
    SD_HIP(hipEventCreate(&e));
    *ev = (void*)e;
    return SD_OK;
}

extern "C" sd_status sd_event_destroy(void* ev) {
    if (ev) SD_HIP(hipEventDestroy((hipEvent_t)ev));
    return SD_OK;
}

extern "C" sd_status sd_event_record(void* ev, void* stream) {
    SD_REQUIRE(ev, "sd_event_record: NULL event");
    SD_HIP(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
    return SD_OK;
}

extern "C" sd_status sd_event_elapsed_ms(void* ev_start, void* ev_stop, float* ms) {
    SD_REQUIRE(ev_start && ev_stop && ms, "sd_event_elapsed_ms: NULL");
    SD_HIP(hipEventSynchronize((hipEvent_t)ev_stop));
    SD_HIP(hipEventElapsedTime(ms, (hipEvent_t)ev_start, (hipEvent_t)ev_stop));
    return SD_OK;
}

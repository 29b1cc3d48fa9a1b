// Shared helpers for the HIP translation units of libstardust_hip.so.
// Error plumbing (sd_status + thread-local message), launch checking and a few
// wave64 primitives used by several kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include "../../include/sd_api.h"

namespace sd {

// Records msg as the calling thread's last error and returns code.
sd_status set_error(sd_status code, const char* fmt, ...);
sd_status hip_error(hipError_t e, const char* what, const char* file, int line);

constexpr int WAVE = 64;

inline int div_up(long a, long b) { return (int)((a + b - 1) / b); }

// Lane helpers (gfx950 runs wave64 only).
__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() {
    return (lane_id() == 0) ? 0ull : (~0ull >> (64 - lane_id()));
}

// Inclusive wave scan (wave64) of a 32-bit unsigned value.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        uint32_t u = __shfl_up(v, d, WAVE);
        if (lane_id() >= d) v += u;
    }
    return v;
}

// Exclusive block scan over blockDim.x <= 1024 threads (one value per thread).
// wsum must hold blockDim.x/64 entries of LDS. Returns the exclusive prefix and
// writes the block total to *total (all threads).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* wsum, uint32_t* total) {
    const int w = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    uint32_t inc = wave_inclusive_scan(v);
    if (lane_id() == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t base = 0, tot = 0;
    for (int i = 0; i < nw; ++i) {
        uint32_t s = wsum[i];
        if (i < w) base += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return base + inc - v;
}

}  // namespace sd

#define SD_HIP(call)                                                            \
    do {                                                                        \
        hipError_t e_ = (call);                                                 \
        if (e_ != hipSuccess) return ::sd::hip_error(e_, #call, __FILE__, __LINE__); \
    } while (0)

#define SD_LAUNCH_CHECK() SD_HIP(hipGetLastError())

#define SD_REQUIRE(cond, ...)                                                   \
    do {                                                                        \
        if (!(cond)) return ::sd::set_error(SD_ERR_ARG, __VA_ARGS__);           \
    } while (0)

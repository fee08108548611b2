// Shared helpers of the libksm HIP sources (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "ksm.h"

namespace ksm {

// Records a formatted message for ksm_last_error() and returns `code`.
int fail(int code, const char *fmt, ...);

// Maps a HIP error to KSM_ERR_HIP with the call site in the message.
int hip_fail(hipError_t e, const char *where);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

constexpr size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

constexpr int kWave = 64;

// Sum over the 64 lanes of a wave (every lane gets the result).
__device__ __forceinline__ float wave_sum_f(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over each 32-lane half of a wave (lanes of a half get its result).
__device__ __forceinline__ float half_sum_f(float v) {
    for (int o = 16; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Maximum over each 32-lane half of a wave; a NaN on a lane makes its half's
// result NaN.
__device__ __forceinline__ float half_max_nan(float v) {
    for (int o = 16; o > 0; o >>= 1) {
        const float w = __shfl_xor(v, o, 64);
        v = (v != v || w != w) ? __builtin_nanf("") : fmaxf(v, w);
    }
    return v;
}

// Maximum over the 64 lanes of a wave; a NaN on any lane makes the result NaN.
__device__ __forceinline__ float wave_max_nan(float v) {
    for (int o = 32; o > 0; o >>= 1) {
        const float w = __shfl_xor(v, o, 64);
        v = (v != v || w != w) ? __builtin_nanf("") : fmaxf(v, w);
    }
    return v;
}

// Workspace carving: successive 256-byte aligned sub-buffers.
struct Carver {
    char *base;
    size_t off = 0;
    explicit Carver(void *p) : base(static_cast<char *>(p)) {}
    template <class T> T *take(size_t n) {
        off = align_up(off, 256);
        T *p = reinterpret_cast<T *>(base + off);
        off += n * sizeof(T);
        return p;
    }
};

}  // namespace ksm

#define KSM_LAUNCH_CHECK(where)                                   \
    do {                                                          \
        hipError_t e_ = hipGetLastError();                        \
        if (e_ != hipSuccess) return ksm::hip_fail(e_, where);    \
    } while (0)

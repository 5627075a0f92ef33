// Internal helpers shared by the translation units of libstenpy_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>
#include <string>

#include "stenpy_hip.h"

namespace sp {

// Thread-local message of the last failure (returned by sp_last_error()).
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define SP_HIP_CHECK(expr)                                                                 \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::sp::fail(SP_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define SP_REQUIRE(cond, msg)                                       \
    do {                                                            \
        if (!(cond)) return ::sp::fail(SP_EINVAL, std::string(msg)); \
    } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline int64_t numel(const sp_tensor* t) {
    int64_t n = 1;
    for (int i = 0; i < t->ndim; ++i) n *= t->shape[i];
    return n;
}

inline bool is_contiguous(const sp_tensor* t) {
    int64_t expect = 1;
    for (int i = t->ndim - 1; i >= 0; --i) {
        if (t->shape[i] != 1 && t->stride[i] != expect) return false;
        expect *= t->shape[i];
    }
    return true;
}

// Launch check after a kernel launch (captures configuration errors synchronously).
inline int launch_status(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(SP_EHIP, std::string(what) + ": " + hipGetErrorString(e));
    return SP_OK;
}

inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

// ---- device helpers -------------------------------------------------------------------------
__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
    return __uint_as_float(((uint32_t)h) << 16);
}
// Round-to-nearest-even FP32 -> BF16 (NaN stays NaN).
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

template <typename T> struct Store;
template <> struct Store<float> {
    static __device__ __forceinline__ void put(float* p, int64_t i, float v) { p[i] = v; }
};
template <> struct Store<uint16_t> {
    static __device__ __forceinline__ void put(uint16_t* p, int64_t i, float v) {
        p[i] = f32_to_bf16(v);
    }
};

// Neumaier compensated accumulator (FP32 pair).
struct Comp32 {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float x) {
        float t = s + x;
        if (fabsf(s) >= fabsf(x)) c += (s - t) + x; else c += (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ float value() const { return s + c; }
};

}  // namespace sp

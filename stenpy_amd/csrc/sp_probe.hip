// HBM streaming probes: the bandwidth this GPU sustains for a given read:write mix, used as
// the measured ceiling beside the 8 TB/s roofline.
//
//   SP_PROBE_COPY   out[i] = a[i]                          1 read : 1 write
//   SP_PROBE_4R1W   4 reads : 1 write; f (n floats) and u (3n floats) read, n floats written:
//                   the byte mix of the stencil graph laplacian(f) + u.grad(f) + |curl u|^2
//                   (16 B read, 4 B written per voxel).  u is read as three contiguous thirds
//                   so that every load is fully coalesced: the probe measures the mix, not an
//                   access pattern
//   SP_PROBE_4R1W_AOS the same mix with u channel-last (voxel-major, 3 floats per voxel), the
//                   graph's own layout: a thread's four voxels read 48 contiguous bytes of u as
//                   three 16-byte loads, so each load instruction covers a third of its wave's
//                   3 KiB and the three together cover it whole
//   SP_PROBE_READ   per-block sums of a[i]                 reads only (one double per block)
//   SP_PROBE_WRITE  out[i] = i-dependent constant           writes only
//
// Each mix has two launch shapes: one float4 group per thread over the whole range ("flat"),
// and a grid-stride loop over 8 blocks per CU, one group per iteration ("persistent").  The
// caller times both and keeps the faster.  Every access is a 16-byte vector load or store,
// so n must be a multiple of 4 and the pointers 16-byte aligned.
#include <hip/hip_runtime.h>

#include "sp_common.h"
#include "stenpy_hip.h"

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ float4 add4(float4 a, float4 b) {
    return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
}

// four voxels' f plus 12 floats of u -> four sums (the probe prices bytes, not arithmetic)
__device__ __forceinline__ float4 mix4(float4 f, float4 u0, float4 u1, float4 u2) {
    return make_float4(f.x + (u0.x + u0.y + u0.z), f.y + (u0.w + u1.x + u1.y),
                       f.z + (u1.z + u1.w + u2.x), f.w + (u2.y + u2.z + u2.w));
}

__global__ void __launch_bounds__(kThreads) k_copy(const float4* __restrict__ a,
                                                   float4* __restrict__ o, int64_t n4,
                                                   int persistent) {
    const int64_t stride = persistent ? (int64_t)gridDim.x * kThreads : n4;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n4; i += stride) o[i] = a[i];
}

__global__ void __launch_bounds__(kThreads) k_4r1w(const float4* __restrict__ f,
                                                   const float4* __restrict__ u,
                                                   float4* __restrict__ o, int64_t n4,
                                                   int persistent) {
    const int64_t stride = persistent ? (int64_t)gridDim.x * kThreads : n4;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n4; i += stride)
        o[i] = mix4(f[i], u[i], u[n4 + i], u[2 * n4 + i]);
}

__global__ void __launch_bounds__(kThreads) k_4r1w_aos(const float4* __restrict__ f,
                                                       const float4* __restrict__ u,
                                                       float4* __restrict__ o, int64_t n4,
                                                       int persistent) {
    const int64_t stride = persistent ? (int64_t)gridDim.x * kThreads : n4;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n4; i += stride)
        o[i] = mix4(f[i], u[3 * i], u[3 * i + 1], u[3 * i + 2]);
}

__global__ void __launch_bounds__(kThreads) k_read(const float4* __restrict__ a,
                                                   double* __restrict__ part, int64_t n4,
                                                   int persistent) {
    __shared__ float s_part[kThreads / 64];
    const int64_t stride = persistent ? (int64_t)gridDim.x * kThreads : n4;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n4; i += stride)
        acc = add4(acc, a[i]);
    float v = (acc.x + acc.y) + (acc.z + acc.w);
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int w = 0; w < kThreads / 64; ++w) t += s_part[w];
        part[blockIdx.x] = (double)t;
    }
}

__global__ void __launch_bounds__(kThreads) k_write(float4* __restrict__ o, int64_t n4,
                                                    int persistent) {
    const int64_t stride = persistent ? (int64_t)gridDim.x * kThreads : n4;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n4; i += stride) {
        const float v = (float)(i & 1023);
        o[i] = make_float4(v, v + 1.f, v + 2.f, v + 3.f);
    }
}

}  // namespace

using namespace sp;

extern "C" int64_t sp_probe_blocks(int persistent, int64_t n, int n_cu) {
    const int64_t flat = (n / 4 + kThreads - 1) / kThreads;
    const int64_t grid = persistent ? (int64_t)(n_cu > 0 ? n_cu : 256) * 8 : flat;
    return grid < flat ? grid : flat;
}

extern "C" int sp_probe_stream(int mix, int persistent, const float* a, const float* u,
                               float* out, double* partials, int64_t n, int n_cu, void* stream) {
    SP_REQUIRE(a && n > 0 && (n & 3) == 0 && ((uintptr_t)a & 15) == 0,
               "probe: n must be a positive multiple of 4 and the input 16-byte aligned");
    const int64_t n4 = n / 4;
    const unsigned g = (unsigned)sp_probe_blocks(persistent, n, n_cu);
    hipStream_t s = as_stream(stream);
    switch (mix) {
        case SP_PROBE_COPY:
            SP_REQUIRE(out && ((uintptr_t)out & 15) == 0, "probe: out must be 16-byte aligned");
            hipLaunchKernelGGL(k_copy, dim3(g), dim3(kThreads), 0, s, (const float4*)a,
                               (float4*)out, n4, persistent);
            break;
        case SP_PROBE_4R1W:
            SP_REQUIRE(out && u && ((uintptr_t)out & 15) == 0 && ((uintptr_t)u & 15) == 0,
                       "probe: u and out must be 16-byte aligned");
            hipLaunchKernelGGL(k_4r1w, dim3(g), dim3(kThreads), 0, s, (const float4*)a,
                               (const float4*)u, (float4*)out, n4, persistent);
            break;
        case SP_PROBE_4R1W_AOS:
            SP_REQUIRE(out && u && ((uintptr_t)out & 15) == 0 && ((uintptr_t)u & 15) == 0,
                       "probe: u and out must be 16-byte aligned");
            hipLaunchKernelGGL(k_4r1w_aos, dim3(g), dim3(kThreads), 0, s, (const float4*)a,
                               (const float4*)u, (float4*)out, n4, persistent);
            break;
        case SP_PROBE_WRITE:
            SP_REQUIRE(((uintptr_t)a & 15) == 0, "probe: out must be 16-byte aligned");
            hipLaunchKernelGGL(k_write, dim3(g), dim3(kThreads), 0, s, (float4*)a, n4, persistent);
            break;
        case SP_PROBE_READ:
            SP_REQUIRE(partials, "probe: partials (one double per block) required");
            hipLaunchKernelGGL(k_read, dim3(g), dim3(kThreads), 0, s, (const float4*)a, partials,
                               n4, persistent);
            break;
        default:
            return fail(SP_EINVAL, "probe: unknown mix");
    }
    return launch_status("sp_probe_stream");
}

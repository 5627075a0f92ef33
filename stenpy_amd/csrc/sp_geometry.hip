// Exact Euclidean distance transform (stenpy_engine.py:3115-3156, CPU branch: scipy's
// distance_transform_edt of (t == 0) with per-axis sampling).  Voxels with t != 0 are the
// background (distance 0); every voxel with t == 0 gets the distance to the nearest background
// voxel.
//
// Separable min-plus passes in FP64, axis 0 first: after the pass over axis d every voxel holds
// the smallest sum of squared per-axis offsets over the axes done so far,
//     D_d(i) = min_j D_{d-1}(j) + ((i_d - j_d) s_d)^2    (j on i's line along axis d),
// starting from D_{-1} = 0 on the background and +inf elsewhere.  The sum is accumulated in the
// order scipy forms it ((c0 + c1) + c2 with c_d = (offset_d * s_d)^2), and each voxel scans its
// line outwards from itself until the squared offset alone reaches the best sum found, so the
// work is proportional to the distances.  A field with no background has no finite distance:
// every voxel gets +inf.
#include <algorithm>
#include <cmath>
#include <string>

#include "sp_common.h"

namespace sp {
namespace {

__global__ void k_edt_init(const float* __restrict__ x, double* __restrict__ d, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        d[i] = x[i] == 0.0f ? INFINITY : 0.0;
}

// one axis: ``len`` points spaced ``stride`` elements apart, sampling s
__global__ void k_edt_pass(const double* __restrict__ in, double* __restrict__ out, int64_t len,
                           int64_t stride, double s, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t pos = (i / stride) % len;
        double best = in[i];
        for (int64_t r = 1; r < len; ++r) {
            const double o = __dmul_rn((double)r, s);
            const double c = __dmul_rn(o, o);
            if (!(c < best)) break;
            if (pos - r >= 0) best = fmin(best, __dadd_rn(in[i - r * stride], c));
            if (pos + r < len) best = fmin(best, __dadd_rn(in[i + r * stride], c));
        }
        out[i] = best;
    }
}

__global__ void k_edt_sqrt(const double* __restrict__ d, float* __restrict__ out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (float)sqrt(d[i]);
}

unsigned grid_of(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 16384); }

}  // namespace
}  // namespace sp

using namespace sp;

extern "C" int sp_distance_transform(const sp_tensor* x, sp_tensor* out, const double* dx, void* stream) {
    SP_REQUIRE(x && out && x->ptr && out->ptr && dx, "sp_distance_transform: null argument");
    SP_REQUIRE(x->ndim >= 1 && x->ndim <= 3, "sp_distance_transform: 1-D, 2-D or 3-D fields only");
    SP_REQUIRE(x->dtype == SP_F32 && out->dtype == SP_F32, "sp_distance_transform: FP32 fields only");
    SP_REQUIRE(is_contiguous(x) && is_contiguous(out), "sp_distance_transform: contiguous tensors only");
    SP_REQUIRE(out->ndim == x->ndim, "sp_distance_transform: output rank differs");
    for (int i = 0; i < x->ndim; ++i) SP_REQUIRE(out->shape[i] == x->shape[i], "sp_distance_transform: shapes differ");
    const int64_t n = numel(x);
    if (n == 0) return SP_OK;
    hipStream_t st = as_stream(stream);
    double *a = nullptr, *b = nullptr;
    SP_HIP_CHECK(hipMallocAsync((void**)&a, sizeof(double) * n, st));
    hipError_t eb = hipMallocAsync((void**)&b, sizeof(double) * n, st);
    int rc = eb == hipSuccess ? SP_OK : fail(SP_EHIP, std::string("hipMallocAsync: ") + hipGetErrorString(eb));
    if (!rc) {
        k_edt_init<<<grid_of(n), 256, 0, st>>>((const float*)x->ptr, a, n);
        rc = launch_status("edt init");
    }
    int64_t stride = n;
    for (int d = 0; d < x->ndim && !rc; ++d) {
        const int64_t len = x->shape[d];
        stride /= len;
        k_edt_pass<<<grid_of(n), 256, 0, st>>>(a, b, len, stride, dx[d], n);
        rc = launch_status("edt pass");
        double* t = a; a = b; b = t;
    }
    if (!rc) {
        k_edt_sqrt<<<grid_of(n), 256, 0, st>>>(a, (float*)out->ptr, n);
        rc = launch_status("edt sqrt");
    }
    for (double* p : {a, b}) {
        if (!p) continue;
        hipError_t e = hipFreeAsync(p, st);
        if (!rc && e != hipSuccess) rc = fail(SP_EHIP, std::string("hipFreeAsync: ") + hipGetErrorString(e));
    }
    return rc;
}

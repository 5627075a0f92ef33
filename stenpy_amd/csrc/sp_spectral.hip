// Spectral operators (stenpy_engine.py:2871-3012): n-D FFTs and spectral derivatives.
//
// The transforms run through rocFFT in double precision (complex-to-complex, in place, on a
// stream-ordered FP64 workspace); the hand-written kernels around them promote the FP32 input,
// apply the spectral multiplier and round the result once to FP32.  The multiplier restates
// the reference's wave numbers exactly: k = 2 pi * fftfreq(n, dx) where fftfreq is evaluated in
// FP32 (torch's default dtype) as float(j) * float(1 / (n dx)), j = i or i - n, and then
// widened to FP64 (stenpy_engine.py:2939-2946).
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <string>

#include <rocfft/rocfft.h>

#include "sp_common.h"

namespace sp {
namespace {

constexpr double kTwoPi = 6.283185307179586;  // 2 * math.pi

struct PlanKey {
    int ndim;
    size_t len[3];
    int inverse;
    int device;
    bool operator<(const PlanKey& o) const {
        if (ndim != o.ndim) return ndim < o.ndim;
        for (int i = 0; i < 3; ++i)
            if (len[i] != o.len[i]) return len[i] < o.len[i];
        if (inverse != o.inverse) return inverse < o.inverse;
        return device < o.device;
    }
};

struct Plan {
    rocfft_plan plan = nullptr;
    size_t work = 0;
};

std::mutex g_mu;
std::map<PlanKey, Plan> g_plans;
bool g_setup = false;

int rocfft_check(rocfft_status s, const char* what) {
    if (s != rocfft_status_success) return fail(SP_EHIP, std::string(what) + ": rocFFT status " + std::to_string((int)s));
    return SP_OK;
}

// Plan of an in-place double-precision complex transform over all axes (cached per shape,
// direction and device; plans live until the process ends).
int get_plan(int ndim, const int64_t* shape, int inverse, Plan* out) {
    int dev = 0;
    SP_HIP_CHECK(hipGetDevice(&dev));
    PlanKey key{ndim, {1, 1, 1}, inverse, dev};
    for (int i = 0; i < ndim; ++i) key.len[i] = (size_t)shape[ndim - 1 - i];  // fastest first
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_setup) {
        if (int rc = rocfft_check(rocfft_setup(), "rocfft_setup")) return rc;
        g_setup = true;
    }
    auto it = g_plans.find(key);
    if (it == g_plans.end()) {
        Plan p;
        if (int rc = rocfft_check(rocfft_plan_create(&p.plan, rocfft_placement_inplace,
                                                     inverse ? rocfft_transform_type_complex_inverse
                                                             : rocfft_transform_type_complex_forward,
                                                     rocfft_precision_double, (size_t)ndim, key.len, 1,
                                                     nullptr),
                                  "rocfft_plan_create"))
            return rc;
        if (int rc = rocfft_check(rocfft_plan_get_work_buffer_size(p.plan, &p.work), "work size")) return rc;
        it = g_plans.emplace(key, p).first;
    }
    *out = it->second;
    return SP_OK;
}

int run_plan(const Plan& p, double2* z, hipStream_t st) {
    void* work = nullptr;
    if (p.work) SP_HIP_CHECK(hipMallocAsync(&work, p.work, st));
    rocfft_execution_info info = nullptr;
    int rc = rocfft_check(rocfft_execution_info_create(&info), "execution info");
    if (!rc) rc = rocfft_check(rocfft_execution_info_set_stream(info, st), "set stream");
    if (!rc && p.work) rc = rocfft_check(rocfft_execution_info_set_work_buffer(info, work, p.work), "work buffer");
    void* bufs[1] = {z};
    if (!rc) rc = rocfft_check(rocfft_execute(p.plan, bufs, nullptr, info), "rocfft_execute");
    if (info) rocfft_execution_info_destroy(info);
    if (work) {
        hipError_t e = hipFreeAsync(work, st);
        if (!rc && e != hipSuccess) rc = fail(SP_EHIP, std::string("hipFreeAsync: ") + hipGetErrorString(e));
    }
    return rc;
}

// x (FP32 real, or interleaved FP32 complex) -> FP64 complex
__global__ void k_load(const float* __restrict__ x, int cplx, double2* __restrict__ z, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        z[i] = cplx ? make_double2((double)x[2 * i], (double)x[2 * i + 1]) : make_double2((double)x[i], 0.0);
}

// scale and round once: complex FP32 pairs (mode 0) or the real part in FP32 (mode 1)
__global__ void k_store(const double2* __restrict__ z, float* __restrict__ out, int mode, double scale,
                        int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 v = z[i];
        if (mode == 0) {
            out[2 * i] = (float)(v.x * scale);
            out[2 * i + 1] = (float)(v.y * scale);
        } else {
            out[i] = (float)(v.x * scale);
        }
    }
}

struct KAxes {
    int64_t n[3];   // as a 3-D field (leading axes of length 1 for 1-D / 2-D inputs)
    float s[3];     // float(1 / (n dx)) per axis
};

__device__ __forceinline__ double wave_number(int64_t i, int64_t n, float s) {
    const int64_t j = i < (n + 1) / 2 ? i : i - n;
    return __dmul_rn((double)__fmul_rn((float)j, s), kTwoPi);
}

// z *= i k_dim (kind 0) or z *= -(k0^2 + k1^2 + k2^2) (kind 1), products rounded separately
// as torch evaluates them (no contraction)
__global__ void k_kmul(double2* __restrict__ z, const KAxes ax, int kind, int dim, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i2 = i % ax.n[2], i1 = (i / ax.n[2]) % ax.n[1], i0 = i / (ax.n[2] * ax.n[1]);
        const double2 v = z[i];
        if (kind == 0) {
            const int64_t ii = dim == 0 ? i0 : (dim == 1 ? i1 : i2);
            const double k = wave_number(ii, ax.n[dim], ax.s[dim]);
            z[i] = make_double2(-__dmul_rn(k, v.y), __dmul_rn(k, v.x));
        } else {
            const double k0 = wave_number(i0, ax.n[0], ax.s[0]);
            const double k1 = wave_number(i1, ax.n[1], ax.s[1]);
            const double k2 = wave_number(i2, ax.n[2], ax.s[2]);
            const double m = -__dadd_rn(__dadd_rn(__dmul_rn(k0, k0), __dmul_rn(k1, k1)), __dmul_rn(k2, k2));
            z[i] = make_double2(__dmul_rn(m, v.x), __dmul_rn(m, v.y));
        }
    }
}

unsigned grid_for(int64_t n) { return (unsigned)std::min<int64_t>(((n + 255) / 256), 8192); }

int check_field(const sp_tensor* t, const char* what) {
    SP_REQUIRE(t && t->ptr, std::string(what) + ": null tensor");
    SP_REQUIRE(t->ndim >= 1 && t->ndim <= 3, std::string(what) + ": 1-D, 2-D or 3-D fields only");
    SP_REQUIRE(is_contiguous(t), std::string(what) + ": contiguous tensors only");
    return SP_OK;
}

bool same_shape(const sp_tensor* a, const sp_tensor* b) {
    if (a->ndim != b->ndim) return false;
    for (int i = 0; i < a->ndim; ++i)
        if (a->shape[i] != b->shape[i]) return false;
    return true;
}

}  // namespace
}  // namespace sp

using namespace sp;

extern "C" int sp_fft(const sp_tensor* x, sp_tensor* out, int inverse, int norm, void* stream) {
    if (int rc = check_field(x, "sp_fft")) return rc;
    if (int rc = check_field(out, "sp_fft out")) return rc;
    SP_REQUIRE(same_shape(x, out), "sp_fft: input and output shapes differ");
    SP_REQUIRE(x->dtype == SP_F32 || x->dtype == SP_C64, "sp_fft: input must be FP32 or complex FP32");
    SP_REQUIRE(out->dtype == SP_C64 || (inverse && out->dtype == SP_F32),
               "sp_fft: output must be complex FP32 (or FP32, the real part, for the inverse)");
    SP_REQUIRE(norm >= SP_NORM_BACKWARD && norm <= SP_NORM_FORWARD, "sp_fft: unknown norm");
    const int64_t n = numel(x);
    if (n == 0) return SP_OK;
    hipStream_t st = as_stream(stream);
    Plan plan;
    if (int rc = get_plan(x->ndim, x->shape, inverse, &plan)) return rc;
    double2* z = nullptr;
    SP_HIP_CHECK(hipMallocAsync((void**)&z, sizeof(double2) * n, st));
    k_load<<<grid_for(n), 256, 0, st>>>((const float*)x->ptr, x->dtype == SP_C64, z, n);
    int rc = launch_status("fft load");
    if (!rc) rc = run_plan(plan, z, st);
    if (!rc) {
        // torch.fft norm: "backward" scales the inverse by 1/n, "forward" the forward one,
        // "ortho" both by 1/sqrt(n)
        const double scale = norm == SP_NORM_ORTHO ? 1.0 / std::sqrt((double)n)
                             : ((norm == SP_NORM_FORWARD) != (inverse != 0) ? 1.0 / (double)n : 1.0);
        k_store<<<grid_for(n), 256, 0, st>>>(z, (float*)out->ptr, out->dtype == SP_F32 ? 1 : 0, scale, n);
        rc = launch_status("fft store");
    }
    hipError_t e = hipFreeAsync(z, st);
    if (!rc && e != hipSuccess) rc = fail(SP_EHIP, std::string("hipFreeAsync: ") + hipGetErrorString(e));
    return rc;
}

extern "C" int sp_spectral(const sp_tensor* x, sp_tensor* out, const double* dx, int kind, int dim,
                           void* stream) {
    if (int rc = check_field(x, "sp_spectral")) return rc;
    if (int rc = check_field(out, "sp_spectral out")) return rc;
    SP_REQUIRE(same_shape(x, out), "sp_spectral: input and output shapes differ");
    SP_REQUIRE(x->dtype == SP_F32 && out->dtype == SP_F32, "sp_spectral: FP32 fields only");
    SP_REQUIRE(kind == SP_SPECTRAL_GRADIENT || kind == SP_SPECTRAL_LAPLACIAN, "sp_spectral: unknown kind");
    SP_REQUIRE(dx, "sp_spectral: null spacing");
    const int nd = x->ndim;
    SP_REQUIRE(kind != SP_SPECTRAL_GRADIENT || (dim >= 0 && dim < nd), "sp_spectral: dim out of range");
    const int64_t n = numel(x);
    if (n == 0) return SP_OK;
    KAxes ax;
    for (int a = 0; a < 3; ++a) {
        const int src = a - (3 - nd);
        ax.n[a] = src >= 0 ? x->shape[src] : 1;
        const double d = src >= 0 ? dx[src] : 1.0;
        SP_REQUIRE(d != 0.0, "sp_spectral: zero spacing");
        ax.s[a] = (float)(1.0 / ((double)ax.n[a] * d));
    }
    hipStream_t st = as_stream(stream);
    Plan fwd, inv;
    if (int rc = get_plan(nd, x->shape, 0, &fwd)) return rc;
    if (int rc = get_plan(nd, x->shape, 1, &inv)) return rc;
    double2* z = nullptr;
    SP_HIP_CHECK(hipMallocAsync((void**)&z, sizeof(double2) * n, st));
    k_load<<<grid_for(n), 256, 0, st>>>((const float*)x->ptr, 0, z, n);
    int rc = launch_status("spectral load");
    if (!rc) rc = run_plan(fwd, z, st);
    if (!rc) {
        k_kmul<<<grid_for(n), 256, 0, st>>>(z, ax, kind, dim + (3 - nd), n);
        rc = launch_status("spectral multiply");
    }
    if (!rc) rc = run_plan(inv, z, st);
    if (!rc) {
        // forward and inverse with the same norm scale by 1/n in total, whatever the norm
        k_store<<<grid_for(n), 256, 0, st>>>(z, (float*)out->ptr, 1, 1.0 / (double)n, n);
        rc = launch_status("spectral store");
    }
    hipError_t e = hipFreeAsync(z, st);
    if (!rc && e != hipSuccess) rc = fail(SP_EHIP, std::string("hipFreeAsync: ") + hipGetErrorString(e));
    return rc;
}

// Spectral operators (stenpy_engine.py:2871-3012): n-D FFTs and spectral derivatives.
//
// The transforms run through rocFFT in double precision on a stream-ordered FP64 workspace:
// complex-to-complex in place for the FFT ops, real-to-complex / complex-to-real (half the
// work) for the spectral derivatives.  The hand-written kernels around them promote the FP32
// input, apply the spectral multiplier and round the result once to FP32.  The multiplier restates
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
    int type;       // rocfft_transform_type
    int device;
    bool operator<(const PlanKey& o) const {
        if (ndim != o.ndim) return ndim < o.ndim;
        for (int i = 0; i < 3; ++i)
            if (len[i] != o.len[i]) return len[i] < o.len[i];
        if (type != o.type) return type < o.type;
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

// Double-precision plan over all axes of the given type: complex forward / inverse in place,
// real forward / inverse out of place (cached per shape, type and device; plans live until the
// process ends).
int get_plan(int ndim, const int64_t* shape, rocfft_transform_type type, Plan* out) {
    int dev = 0;
    SP_HIP_CHECK(hipGetDevice(&dev));
    PlanKey key{ndim, {1, 1, 1}, (int)type, dev};
    for (int i = 0; i < ndim; ++i) key.len[i] = (size_t)shape[ndim - 1 - i];  // fastest first
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_setup) {
        if (int rc = rocfft_check(rocfft_setup(), "rocfft_setup")) return rc;
        g_setup = true;
    }
    auto it = g_plans.find(key);
    if (it == g_plans.end()) {
        Plan p;
        const bool cplx = type == rocfft_transform_type_complex_forward ||
                          type == rocfft_transform_type_complex_inverse;
        if (int rc = rocfft_check(rocfft_plan_create(&p.plan, cplx ? rocfft_placement_inplace : rocfft_placement_notinplace,
                                                     type, rocfft_precision_double, (size_t)ndim, key.len, 1,
                                                     nullptr),
                                  "rocfft_plan_create"))
            return rc;
        if (int rc = rocfft_check(rocfft_plan_get_work_buffer_size(p.plan, &p.work), "work size")) return rc;
        it = g_plans.emplace(key, p).first;
    }
    *out = it->second;
    return SP_OK;
}

int run_plan(const Plan& p, void* in, void* out, hipStream_t st) {
    void* work = nullptr;
    if (p.work) SP_HIP_CHECK(hipMallocAsync(&work, p.work, st));
    rocfft_execution_info info = nullptr;
    int rc = rocfft_check(rocfft_execution_info_create(&info), "execution info");
    if (!rc) rc = rocfft_check(rocfft_execution_info_set_stream(info, st), "set stream");
    if (!rc && p.work) rc = rocfft_check(rocfft_execution_info_set_work_buffer(info, work, p.work), "work buffer");
    void* ibuf[1] = {in};
    void* obuf[1] = {out};
    if (!rc) rc = rocfft_check(rocfft_execute(p.plan, ibuf, out ? obuf : nullptr, info), "rocfft_execute");
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
    int64_t h2;     // stored length of the last axis: n[2] / 2 + 1 (half spectrum)
    float s[3];     // float(1 / (n dx)) per axis
};

// FP32 real -> FP64 real
__global__ void k_load_real(const float* __restrict__ x, double* __restrict__ r, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        r[i] = (double)x[i];
}

__global__ void k_store_real(const double* __restrict__ r, float* __restrict__ out, double scale, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (float)(r[i] * scale);
}

__device__ __forceinline__ double wave_number(int64_t i, int64_t n, float s) {
    const int64_t j = i < (n + 1) / 2 ? i : i - n;
    return __dmul_rn((double)__fmul_rn((float)j, s), kTwoPi);
}

// On the half spectrum (n0, n1, n2 / 2 + 1) of a real field: z *= i k_dim (kind 0) or
// z *= -(k0^2 + k1^2 + k2^2) (kind 1), products rounded separately as torch evaluates them (no
// contraction).  For kind 0 the Nyquist plane of axis dim (even length) is zeroed: there
// i k_dim X is anti-Hermitian, so its share of ifftn is purely imaginary and the reference's
// .real drops it; everywhere else the product is Hermitian and the complex-to-real inverse
// returns exactly that real part.
__global__ void k_kmul(double2* __restrict__ z, const KAxes ax, int kind, int dim, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i2 = i % ax.h2, i1 = (i / ax.h2) % ax.n[1], i0 = i / (ax.h2 * ax.n[1]);
        const double2 v = z[i];
        if (kind == 0) {
            const int64_t ii = dim == 0 ? i0 : (dim == 1 ? i1 : i2);
            const int64_t nd = ax.n[dim];
            const double k = wave_number(ii, nd, ax.s[dim]);
            z[i] = ((nd & 1) == 0 && ii == nd / 2) ? make_double2(0.0, 0.0)
                                                     : make_double2(-__dmul_rn(k, v.y), __dmul_rn(k, v.x));
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
    if (int rc = get_plan(x->ndim, x->shape, inverse ? rocfft_transform_type_complex_inverse
                                                     : rocfft_transform_type_complex_forward, &plan))
        return rc;
    double2* z = nullptr;
    SP_HIP_CHECK(hipMallocAsync((void**)&z, sizeof(double2) * n, st));
    k_load<<<grid_for(n), 256, 0, st>>>((const float*)x->ptr, x->dtype == SP_C64, z, n);
    int rc = launch_status("fft load");
    if (!rc) rc = run_plan(plan, z, nullptr, st);
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
    ax.h2 = ax.n[2] / 2 + 1;
    const int64_t nh = ax.n[0] * ax.n[1] * ax.h2;
    hipStream_t st = as_stream(stream);
    Plan fwd, inv;
    if (int rc = get_plan(nd, x->shape, rocfft_transform_type_real_forward, &fwd)) return rc;
    if (int rc = get_plan(nd, x->shape, rocfft_transform_type_real_inverse, &inv)) return rc;
    double* r = nullptr;
    double2* z = nullptr;
    SP_HIP_CHECK(hipMallocAsync((void**)&r, sizeof(double) * n, st));
    hipError_t ea = hipMallocAsync((void**)&z, sizeof(double2) * nh, st);
    int rc = ea == hipSuccess ? SP_OK : fail(SP_EHIP, std::string("hipMallocAsync: ") + hipGetErrorString(ea));
    if (!rc) {
        k_load_real<<<grid_for(n), 256, 0, st>>>((const float*)x->ptr, r, n);
        rc = launch_status("spectral load");
    }
    if (!rc) rc = run_plan(fwd, r, z, st);
    if (!rc) {
        k_kmul<<<grid_for(nh), 256, 0, st>>>(z, ax, kind, dim + (3 - nd), nh);
        rc = launch_status("spectral multiply");
    }
    if (!rc) rc = run_plan(inv, z, r, st);
    if (!rc) {
        // forward and inverse with the same norm scale by 1/n in total, whatever the norm
        k_store_real<<<grid_for(n), 256, 0, st>>>(r, (float*)out->ptr, 1.0 / (double)n, n);
        rc = launch_status("spectral store");
    }
    for (void* b : {(void*)z, (void*)r}) {
        if (!b) continue;
        hipError_t e = hipFreeAsync(b, st);
        if (!rc && e != hipSuccess) rc = fail(SP_EHIP, std::string("hipFreeAsync: ") + hipGetErrorString(e));
    }
    return rc;
}

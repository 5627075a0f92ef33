// Statistics and integrals outside the fused path (CPU branches of stenpy_engine.py):
//   cumulative_integral  trapezoid cumulative sum along one axis, leading 0   (2462-2483)
//   surface_integral     sum_i w_i sum_c t[i,c] n[i,c]                         (2486-2496)
//   covariance           dot(a - mean a, b - mean b) / max(n - 1, 1)           (2575-2592)
//   correlation          dot(a', b') / (|a'| |b'| + 1e-15)                     (2595-2615)
//   entropy              -sum p log(p + eps), p = |t| / (sum |t| + eps)        (2638-2669)
// Every sum is accumulated in FP64 from the FP32 inputs.  Global sums use per-block partial
// sums and one fixed-order combine (deterministic, no atomics); results stay on the device.
#include <algorithm>
#include <cmath>
#include <string>

#include "sp_common.h"

namespace sp {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxParts = 2048;

__device__ __forceinline__ double block_sum(double v, double* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] += sh[t + s];
        __syncthreads();
    }
    return sh[0];
}

// per-block partial sums of f(i) over [0, n)
template <class F>
__global__ void __launch_bounds__(kThreads) k_partials(F f, int64_t n, double* __restrict__ part) {
    __shared__ double sh[kThreads];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads)
        acc += f(i);
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void __launch_bounds__(kThreads) k_combine(const double* __restrict__ part, int np,
                                                     double* __restrict__ out) {
    __shared__ double sh[kThreads];
    double acc = 0.0;
    for (int i = threadIdx.x; i < np; i += kThreads) acc += part[i];
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) *out = s;
}

int parts_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + kThreads - 1) / kThreads, kMaxParts)); }

template <class F>
int global_sum(F f, int64_t n, double* part, double* out, hipStream_t st) {
    const int np = parts_for(n);
    k_partials<<<np, kThreads, 0, st>>>(f, n, part);
    if (int rc = launch_status("partial sums")) return rc;
    k_combine<<<1, kThreads, 0, st>>>(part, np, out);
    return launch_status("combine");
}

struct AbsF {
    const float* x;
    __device__ double operator()(int64_t i) const { return fabs((double)x[i]); }
};
struct ValF {
    const float* x;
    __device__ double operator()(int64_t i) const { return (double)x[i]; }
};
// (x - mx)(y - my) with the means read from the device
struct CrossF {
    const float* x; const float* y; const double* mx; const double* my;
    __device__ double operator()(int64_t i) const { return ((double)x[i] - *mx) * ((double)y[i] - *my); }
};
struct EntF {
    const float* x; const double* total; double eps;
    __device__ double operator()(int64_t i) const {
        const double p = fabs((double)x[i]) / (*total + eps);
        return -p * log(p + eps);
    }
};
// sum_c t[i, c] n[i, c] (t broadcast over c when tc == 1), times w[i]
struct FluxF {
    const float* t; const float* nr; const float* w; int c; int tc;
    __device__ double operator()(int64_t i) const {
        double s = 0.0;
        for (int k = 0; k < c; ++k) s += (double)t[i * tc + (tc == 1 ? 0 : k)] * (double)nr[i * c + k];
        return s * (double)w[i];
    }
};

__global__ void k_scale_mean(double* v, int64_t n) { *v = *v / (double)n; }

// covariance / correlation from {S_ab, S_aa, S_bb}
__global__ void k_pair_final(const double* s, int kind, int64_t n, double* out) {
    if (kind == 0) *out = s[0] / (double)(n - 1 > 1 ? n - 1 : 1);
    else *out = s[0] / (sqrt(s[1]) * sqrt(s[2]) + 1e-15);
}

// ---- line kernels: one thread per line of ``len`` points spaced ``stride`` apart ----
struct Lines {
    int64_t len, stride, nlines;
    __device__ int64_t base(int64_t l) const { return (l / stride) * stride * len + l % stride; }
};

__global__ void k_cumtrapz(const float* __restrict__ x, float* __restrict__ out, Lines L, double h) {
    for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < L.nlines; l += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = L.base(l);
        double acc = 0.0;
        out[b] = 0.0f;
        double prev = (double)x[b];
        for (int64_t k = 1; k < L.len; ++k) {
            const double cur = (double)x[b + k * L.stride];
            acc += (prev + cur) * h;
            out[b + k * L.stride] = (float)acc;
            prev = cur;
        }
    }
}

// out[l] = -sum p log(p + eps) along the line, p = |x| / (sum |x| + eps)
__global__ void k_entropy_lines(const float* __restrict__ x, float* __restrict__ out, Lines L, double eps) {
    for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < L.nlines; l += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = L.base(l);
        double tot = 0.0;
        for (int64_t k = 0; k < L.len; ++k) tot += fabs((double)x[b + k * L.stride]);
        double acc = 0.0;
        for (int64_t k = 0; k < L.len; ++k) {
            const double p = fabs((double)x[b + k * L.stride]) / (tot + eps);
            acc += -p * log(p + eps);
        }
        out[l] = (float)acc;
    }
}

__global__ void k_zero(float* __restrict__ out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = 0.0f;
}

unsigned line_grid(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384)); }

int require_f32(const sp_tensor* t, const char* what) {
    SP_REQUIRE(t && t->ptr, std::string(what) + ": null tensor");
    SP_REQUIRE(t->dtype == SP_F32, std::string(what) + ": FP32 tensors only");
    SP_REQUIRE(is_contiguous(t), std::string(what) + ": contiguous tensors only");
    return SP_OK;
}

int lines_of(const sp_tensor* x, int dim, Lines* L) {
    SP_REQUIRE(dim >= 0 && dim < x->ndim, "dim out of range");
    int64_t inner = 1;
    for (int i = dim + 1; i < x->ndim; ++i) inner *= x->shape[i];
    L->len = x->shape[dim];
    L->stride = inner;
    L->nlines = numel(x) / std::max<int64_t>(L->len, 1);
    return SP_OK;
}

struct Scratch {
    double* p = nullptr;
    hipStream_t st;
    explicit Scratch(hipStream_t s) : st(s) {}
    int alloc(size_t n) {
        hipError_t e = hipMallocAsync((void**)&p, sizeof(double) * n, st);
        return e == hipSuccess ? SP_OK : fail(SP_EHIP, std::string("hipMallocAsync: ") + hipGetErrorString(e));
    }
    int release(int rc) {
        if (p) {
            hipError_t e = hipFreeAsync(p, st);
            if (!rc && e != hipSuccess) rc = fail(SP_EHIP, std::string("hipFreeAsync: ") + hipGetErrorString(e));
        }
        return rc;
    }
};

}  // namespace
}  // namespace sp

using namespace sp;

extern "C" int sp_cumulative_trapezoid(const sp_tensor* x, sp_tensor* out, int dim, double dx, void* stream) {
    if (int rc = require_f32(x, "sp_cumulative_trapezoid")) return rc;
    if (int rc = require_f32(out, "sp_cumulative_trapezoid out")) return rc;
    SP_REQUIRE(x->ndim == out->ndim && numel(x) == numel(out), "sp_cumulative_trapezoid: shapes differ");
    Lines L;
    if (int rc = lines_of(x, dim, &L)) return rc;
    if (numel(x) == 0) return SP_OK;
    hipStream_t st = as_stream(stream);
    if (L.len < 2) {
        k_zero<<<line_grid(numel(x)), 256, 0, st>>>((float*)out->ptr, numel(x));
        return launch_status("zero");
    }
    k_cumtrapz<<<line_grid(L.nlines), 256, 0, st>>>((const float*)x->ptr, (float*)out->ptr, L, dx * 0.5);
    return launch_status("cumulative trapezoid");
}

extern "C" int sp_surface_integral(const sp_tensor* t, const sp_tensor* normals, const sp_tensor* weights,
                                   double* result, void* stream) {
    if (int rc = require_f32(t, "sp_surface_integral(t)")) return rc;
    if (int rc = require_f32(normals, "sp_surface_integral(normals)")) return rc;
    if (int rc = require_f32(weights, "sp_surface_integral(area_weights)")) return rc;
    SP_REQUIRE(result, "sp_surface_integral: null result");
    SP_REQUIRE(normals->ndim >= 1, "sp_surface_integral: normals need a component axis");
    const int c = (int)normals->shape[normals->ndim - 1];
    const int64_t m = numel(normals) / std::max(c, 1);
    SP_REQUIRE(numel(weights) == m, "sp_surface_integral: area_weights must have the normals' shape without the last axis");
    const int tc = numel(t) == numel(normals) ? c : 1;
    SP_REQUIRE(numel(t) == m * tc, "sp_surface_integral: t must have the normals' shape (or a last axis of 1)");
    hipStream_t st = as_stream(stream);
    Scratch s(st);
    int rc = s.alloc(kMaxParts);
    if (!rc) rc = global_sum(FluxF{(const float*)t->ptr, (const float*)normals->ptr, (const float*)weights->ptr, c, tc},
                             m, s.p, result, st);
    return s.release(rc);
}

extern "C" int sp_pair_stat(int kind, const sp_tensor* a, const sp_tensor* b, double* result, void* stream) {
    if (int rc = require_f32(a, "sp_pair_stat(a)")) return rc;
    if (int rc = require_f32(b, "sp_pair_stat(b)")) return rc;
    SP_REQUIRE(kind == SP_COVARIANCE || kind == SP_CORRELATION, "sp_pair_stat: unknown kind");
    SP_REQUIRE(numel(a) == numel(b), "sp_pair_stat: shape mismatch");
    SP_REQUIRE(result, "sp_pair_stat: null result");
    const int64_t n = numel(a);
    SP_REQUIRE(n > 0, "sp_pair_stat: empty input");
    hipStream_t st = as_stream(stream);
    Scratch s(st);
    // [0, kMaxParts): partials; then ma, mb, S_ab, S_aa, S_bb
    int rc = s.alloc(kMaxParts + 5);
    double* part = s.p;
    double* v = s.p + kMaxParts;
    const float* pa = (const float*)a->ptr;
    const float* pb = (const float*)b->ptr;
    if (!rc) rc = global_sum(ValF{pa}, n, part, v + 0, st);
    if (!rc) rc = global_sum(ValF{pb}, n, part, v + 1, st);
    if (!rc) {
        k_scale_mean<<<1, 1, 0, st>>>(v + 0, n);
        k_scale_mean<<<1, 1, 0, st>>>(v + 1, n);
        rc = launch_status("means");
    }
    if (!rc) rc = global_sum(CrossF{pa, pb, v + 0, v + 1}, n, part, v + 2, st);
    if (!rc && kind == SP_CORRELATION) rc = global_sum(CrossF{pa, pa, v + 0, v + 0}, n, part, v + 3, st);
    if (!rc && kind == SP_CORRELATION) rc = global_sum(CrossF{pb, pb, v + 1, v + 1}, n, part, v + 4, st);
    if (!rc) {
        k_pair_final<<<1, 1, 0, st>>>(v + 2, kind, n, result);
        rc = launch_status("pair statistic");
    }
    return s.release(rc);
}

extern "C" int sp_entropy(const sp_tensor* x, int dim, double eps, double* result, sp_tensor* out, void* stream) {
    if (int rc = require_f32(x, "sp_entropy")) return rc;
    hipStream_t st = as_stream(stream);
    if (dim >= 0) {
        if (int rc = require_f32(out, "sp_entropy out")) return rc;
        Lines L;
        if (int rc = lines_of(x, dim, &L)) return rc;
        SP_REQUIRE(numel(out) == L.nlines, "sp_entropy: output must have the input's shape without dim");
        if (L.nlines == 0) return SP_OK;
        k_entropy_lines<<<line_grid(L.nlines), 256, 0, st>>>((const float*)x->ptr, (float*)out->ptr, L, eps);
        return launch_status("entropy lines");
    }
    SP_REQUIRE(result, "sp_entropy: null result");
    const int64_t n = numel(x);
    Scratch s(st);
    int rc = s.alloc(kMaxParts + 1);
    if (!rc) rc = global_sum(AbsF{(const float*)x->ptr}, n, s.p, s.p + kMaxParts, st);
    if (!rc) rc = global_sum(EntF{(const float*)x->ptr, s.p + kMaxParts, eps}, n, s.p, result, st);
    return s.release(rc);
}

// Global reductions (stenpy_engine.py:2340-2382 sum/mean/norm_l2/min_max, 2546-2572 variance,
// 2415-2459 integrate, 2692-2696 the determinant's any(|diag| < 1e-12) test).
//
// Stage 1: a fixed grid of at most kMaxBlocks blocks; each thread walks a grid-stride range
// and keeps compensated FP32 sums (two-sum error terms; squares split exactly with an FMA,
// so variance = (S2 - S1^2/n)/(n-1) is formed in FP64 at the end); lanes are merged with
// cross-lane shuffles and waves through
// LDS; one FP64 record per block goes to the workspace.
// Stage 2: one block merges the records in FP64 in a fixed order.  Both stages have a fixed
// reduction order, so results are bit-reproducible (the reference's GPU variance uses float
// atomics and is not).
#include <mutex>

#include "sp_common.h"

using namespace sp;

namespace {

constexpr int kMaxBlocks = 2048;
constexpr int kThreads = 256;
constexpr int kRec = 4;  // doubles per block record

struct Acc {
    float s, c;       // compensated sum of the primary quantity (x, x^2 or w x)
    float s2, c2;     // compensated sum of x^2 (variance)
    float lo, hi;     // min / max
};

// The compensated helpers turn FMA contraction off: a product contracted into the running sum
// would leave no separately rounded term for the two-sum and the FMA split to correct.
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
#pragma clang fp contract(off)
    s = a + b;
    float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// a.s/a.c += x and a.s2/a.c2 += x^2, the square split exactly with an FMA
__device__ __forceinline__ void add_sum(float& s, float& c, float x) {
#pragma clang fp contract(off)
    float t, e;
    two_sum(s, x, t, e);
    s = t;
    c += e;
}
__device__ __forceinline__ void add_sq(float& s, float& c, float x) {
#pragma clang fp contract(off)
    const float q = x * x;
    add_sum(s, c, q);
    c += fmaf(x, x, -q);
}
__device__ __forceinline__ void add_prod(float& s, float& c, float x, float w) {
#pragma clang fp contract(off)
    const float q = x * w;
    add_sum(s, c, q);
    c += fmaf(x, w, -q);
}

__device__ __forceinline__ void merge(Acc& a, const Acc& b) {
#pragma clang fp contract(off)
    float s, e;
    two_sum(a.s, b.s, s, e);
    a.s = s;
    a.c = a.c + b.c + e;
    two_sum(a.s2, b.s2, s, e);
    a.s2 = s;
    a.c2 = a.c2 + b.c2 + e;
    a.lo = fminf(a.lo, b.lo);
    a.hi = fmaxf(a.hi, b.hi);
}

__device__ __forceinline__ Acc shfl_acc(const Acc& a, int lane_mask) {
    Acc b;
    b.s = __shfl_xor(a.s, lane_mask);
    b.c = __shfl_xor(a.c, lane_mask);
    b.s2 = __shfl_xor(a.s2, lane_mask);
    b.c2 = __shfl_xor(a.c2, lane_mask);
    b.lo = __shfl_xor(a.lo, lane_mask);
    b.hi = __shfl_xor(a.hi, lane_mask);
    return b;
}

// Per-axis quadrature weight of index i on an axis of n points (reference Simpson/trapezoid,
// stenpy_engine.py:2415-2457) as an integer numerator k over 6: the weight is k * dx / 6.  The
// product of three numerators (at most 8^3) is exact in FP32, each term x * k is split exactly
// with an FMA, and dx0 dx1 dx2 / 216 multiplies the block record once, in FP64.
__device__ __forceinline__ float quad_k(int64_t i, int64_t n, int method) {
    if (method == 1 || n < 3) {  // trapezoid (also Simpson's n < 3 branch)
        if (n < 2) return 0.f;
        return (i == 0 || i == n - 1) ? 3.f : 6.f;
    }
    const int64_t m = (n % 2 == 0) ? n - 1 : n;  // Simpson points
    float k = 0.f;
    if (i < m) k = (i == 0 || i == m - 1) ? 2.f : ((i % 2 == 1) ? 8.f : 4.f);
    if (n % 2 == 0 && i >= n - 2) k += 3.f;  // trapezoid on the last interval
    return k;
}

struct IntegrateGeom {
    int64_t n0, n1, n2;
    double scale;  // dx0 dx1 dx2 / 216
    int method;
};

__global__ void __launch_bounds__(kThreads) k_reduce_stage1(int kind, const float* __restrict__ x,
                                                            int64_t n, double* __restrict__ ws,
                                                            IntegrateGeom ig) {
    Acc a{0.f, 0.f, 0.f, 0.f, INFINITY, -INFINITY};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (kind == SP_R_ANYSMALLDIAG) {
        // x is (..., 3, 3); n counts matrices
        for (int64_t v = t0; v < n; v += stride) {
            const float* m = x + v * 9;
            if (fabsf(m[0]) < 1e-12f || fabsf(m[4]) < 1e-12f || fabsf(m[8]) < 1e-12f) a.hi = 1.f;
        }
    } else if (kind == SP_R_INTEGRATE) {
        for (int64_t i = t0; i < n; i += stride) {
            const int64_t k = i % ig.n2;
            const int64_t r = i / ig.n2;
            const int64_t j = r % ig.n1;
            const int64_t z = r / ig.n1;
            const float w = (quad_k(z, ig.n0, ig.method) * quad_k(j, ig.n1, ig.method)) *
                            quad_k(k, ig.n2, ig.method);
            add_prod(a.s, a.c, x[i], w);
        }
    } else {
        for (int64_t i = t0; i < n; i += stride) {
            const float v = x[i];
            if (kind == SP_R_SUM || kind == SP_R_MEAN) {
                add_sum(a.s, a.c, v);
            } else if (kind == SP_R_NORM2) {
                add_sq(a.s, a.c, v);
            } else if (kind == SP_R_VAR || kind == SP_R_VAR_BIASED) {
                add_sum(a.s, a.c, v);
                add_sq(a.s2, a.c2, v);
            } else {  // min/max; NaN propagates like torch.min/max
                if (v != v) { a.lo = v; a.hi = v; }
                else if (a.lo == a.lo) { a.lo = fminf(a.lo, v); a.hi = fmaxf(a.hi, v); }
            }
        }
    }
    // wave reduction
    for (int mask = 32; mask >= 1; mask >>= 1) {
        Acc b = shfl_acc(a, mask);
        if (kind == SP_R_MINMAX && (a.lo != a.lo || b.lo != b.lo)) {
            a.lo = a.hi = NAN;
        } else {
            merge(a, b);
        }
    }
    __shared__ Acc wave_acc[kThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wave_acc[wave] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        Acc r = wave_acc[0];
        for (int w = 1; w < kThreads / 64; ++w) {
            if (kind == SP_R_MINMAX && (r.lo != r.lo || wave_acc[w].lo != wave_acc[w].lo))
                r.lo = r.hi = NAN;
            else
                merge(r, wave_acc[w]);
        }
        double* rec = ws + (int64_t)blockIdx.x * kRec;
        if (kind == SP_R_VAR || kind == SP_R_VAR_BIASED) {
            rec[0] = r.s;
            rec[1] = r.c;
            rec[2] = r.s2;
            rec[3] = r.c2;
        } else if (kind == SP_R_MINMAX || kind == SP_R_ANYSMALLDIAG) {
            rec[0] = r.lo;
            rec[1] = r.hi;
            rec[2] = rec[3] = 0.0;
        } else {
            const double sc = kind == SP_R_INTEGRATE ? ig.scale : 1.0;
            rec[0] = (double)r.s * sc;
            rec[1] = (double)r.c * sc;
            rec[2] = rec[3] = 0.0;
        }
    }
}

struct Acc64 {
    double s, c, s2, c2, lo, hi;
};

__device__ __forceinline__ void nsum64(double& s, double& c, double x) {
    const double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x; else c += (x - t) + s;
    s = t;
}

__device__ __forceinline__ void merge64(Acc64& a, const Acc64& b, bool minmax) {
    if (minmax) {
        if (a.lo != a.lo || b.lo != b.lo) { a.lo = a.hi = NAN; return; }
        a.lo = fmin(a.lo, b.lo);
        a.hi = fmax(a.hi, b.hi);
        return;
    }
    nsum64(a.s, a.c, b.s);
    a.c += b.c;
    nsum64(a.s2, a.c2, b.s2);
    a.c2 += b.c2;
}

constexpr int kCThreads = 1024;   // threads of a final-combine block

__device__ __forceinline__ Acc64 load_record(int kind, bool mm, const double* __restrict__ rec) {
    Acc64 r{0, 0, 0, 0, INFINITY, -INFINITY};
    if (kind == SP_R_VAR || kind == SP_R_VAR_BIASED) { r.s = rec[0]; r.c = rec[1]; r.s2 = rec[2]; r.c2 = rec[3]; }
    else if (mm) { r.lo = rec[0]; r.hi = rec[1]; }
    else { r.s = rec[0]; r.c = rec[1]; }
    return r;
}

// Final combine of ``nblocks`` block records by one block of kCThreads threads.  Thread t
// merges records t, t + kCThreads, ... into four independent accumulators (their loads are in
// flight together), then the block merges the partials in a fixed tree: the result does not
// depend on timing.
__device__ __forceinline__ void combine_records(int kind, const double* __restrict__ ws, int nblocks,
                                                int64_t n, double* __restrict__ result) {
    const bool mm = (kind == SP_R_MINMAX || kind == SP_R_ANYSMALLDIAG);
    Acc64 a[4];
    for (int k = 0; k < 4; ++k) a[k] = Acc64{0, 0, 0, 0, INFINITY, -INFINITY};
    int b = threadIdx.x;
    for (; b + 3 * kCThreads < nblocks; b += 4 * kCThreads) {
        Acc64 r[4];
        #pragma unroll
        for (int k = 0; k < 4; ++k) r[k] = load_record(kind, mm, ws + (int64_t)(b + k * kCThreads) * kRec);
        #pragma unroll
        for (int k = 0; k < 4; ++k) merge64(a[k], r[k], mm);
    }
    for (; b < nblocks; b += kCThreads) merge64(a[0], load_record(kind, mm, ws + (int64_t)b * kRec), mm);
    merge64(a[0], a[1], mm);
    merge64(a[2], a[3], mm);
    merge64(a[0], a[2], mm);
    __shared__ Acc64 part[kCThreads];
    part[threadIdx.x] = a[0];
    __syncthreads();
    for (int step = kCThreads / 2; step >= 1; step >>= 1) {
        if (threadIdx.x < step) merge64(part[threadIdx.x], part[threadIdx.x + step], mm);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const Acc64 r = part[0];
        switch (kind) {
            case SP_R_SUM: case SP_R_INTEGRATE: result[0] = r.s + r.c; break;
            case SP_R_MEAN: result[0] = (r.s + r.c) / (double)n; break;
            case SP_R_NORM2: result[0] = sqrt(r.s + r.c); break;
            case SP_R_VAR:
            case SP_R_VAR_BIASED: {
                const double s1 = r.s + r.c, s2 = r.s2 + r.c2;
                const double m2 = s2 - s1 * (s1 / (double)n);
                result[0] = m2 / (double)(kind == SP_R_VAR ? n - 1 : n);
                break;
            }
            case SP_R_MINMAX: result[0] = r.lo; result[1] = r.hi; break;
            case SP_R_ANYSMALLDIAG: result[0] = r.hi > 0.0 ? 1.0 : 0.0; break;
        }
    }
}

__global__ void __launch_bounds__(kCThreads) k_reduce_stage2(int kind, const double* __restrict__ ws,
                                                            int nblocks, int64_t n,
                                                            double* __restrict__ result) {
    combine_records(kind, ws, nblocks, n, result);
}

// one block per descriptor: the final combines of many reductions in one launch
__global__ void __launch_bounds__(kCThreads) k_reduce_combine_batch(const sp_combine_desc* __restrict__ d) {
    const sp_combine_desc e = d[blockIdx.x];
    combine_records(e.kind, e.records, e.nrecords, e.count, e.result);
}

// up to SP_COMBINE_MULTI_MAX descriptors passed by value (no device buffer; capturable)
struct CombineArgs { sp_combine_desc d[SP_COMBINE_MULTI_MAX]; };
__global__ void __launch_bounds__(kCThreads) k_reduce_combine_multi(const CombineArgs a) {
    const sp_combine_desc e = a.d[blockIdx.x];
    combine_records(e.kind, e.records, e.nrecords, e.count, e.result);
}

}  // namespace

extern "C" size_t sp_reduce_workspace_bytes(int64_t) {
    return (size_t)kMaxBlocks * kRec * sizeof(double);
}

extern "C" int sp_reduce(int kind, const sp_tensor* x, double* result, void* workspace,
                         const double* wparams, void* stream) {
    SP_REQUIRE(x && x->dtype == SP_F32, "reduce: input must be f32");
    SP_REQUIRE(is_contiguous(x), "reduce: input must be contiguous");
    SP_REQUIRE(result && workspace, "reduce: result and workspace are required");
    SP_REQUIRE(kind >= SP_R_SUM && kind <= SP_R_ANYSMALLDIAG, "reduce: unknown kind");
    int64_t n = numel(x);
    IntegrateGeom ig{1, 1, 1, 1.0, 0};
    if (kind == SP_R_INTEGRATE) {
        SP_REQUIRE(x->ndim == 3 && wparams, "integrate: expected a 3-D field and weights");
        ig = IntegrateGeom{x->shape[0], x->shape[1], x->shape[2],
                           wparams[0] * wparams[1] * wparams[2] / 216.0, (int)wparams[3]};
    }
    if (kind == SP_R_ANYSMALLDIAG) {
        SP_REQUIRE(x->ndim >= 2 && x->shape[x->ndim - 1] == 3 && x->shape[x->ndim - 2] == 3,
                   "any-small-diag: expected (..., 3, 3)");
        n /= 9;
    }
    SP_REQUIRE(n > 0 || kind == SP_R_SUM || kind == SP_R_ANYSMALLDIAG || kind == SP_R_INTEGRATE,
               "reduce: empty input");
    int64_t blocks = (n + kThreads * 8 - 1) / (kThreads * 8);
    if (blocks < 1) blocks = 1;
    if (blocks > kMaxBlocks) blocks = kMaxBlocks;
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_reduce_stage1, dim3((unsigned)blocks), dim3(kThreads), 0, st, kind,
                       (const float*)x->ptr, n, (double*)workspace, ig);
    int rc = launch_status("sp_reduce(stage1)");
    if (rc) return rc;
    hipLaunchKernelGGL(k_reduce_stage2, dim3(1), dim3(kCThreads), 0, st, kind,
                       (const double*)workspace, (int)blocks, n, result);
    return launch_status("sp_reduce(stage2)");
}

// ---- allocation-free variant ---------------------------------------------------------------
// The block records of sp_reduce_noalloc live in this file-scope array instead of a caller
// workspace: kMaxBlocks records of kRec doubles (64 KiB), held by every GPU the library runs
// on until the process ends.  The stage-1 grid is clamped to kMaxBlocks, so any input size
// fits, and every launched block writes all of its record before stage 2 reads it, so the
// array needs no clearing between calls.  Calls that overlap on one GPU share it and race.
static __device__ __align__(256) double scratch[kMaxBlocks * kRec];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of ``scratch`` on the current device, looked up once per device (host code
// must not pass the array's name itself to a kernel: it is a host-side stand-in).
int scratch_for_current_device(void** out) {
    int dev = 0;
    SP_HIP_CHECK(hipGetDevice(&dev));
    SP_REQUIRE(dev >= 0 && dev < kMaxDevices, "reduce_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        SP_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        SP_REQUIRE(p != nullptr, "reduce_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SP_OK;
}

}  // namespace

extern "C" size_t sp_reduce_noalloc_scratch_bytes(void) { return sizeof(scratch); }

extern "C" int sp_reduce_noalloc(int kind, const sp_tensor* x, double* result,
                                 const double* wparams, void* stream) {
    void* ws = nullptr;
    int rc = scratch_for_current_device(&ws);
    if (rc) return rc;
    SP_REQUIRE(x && x->dtype == SP_F32 && is_contiguous(x),
               "reduce_noalloc: input must be contiguous f32");
    // the records sp_reduce would need for this input, against the array's capacity
    const int64_t n = numel(x) / (kind == SP_R_ANYSMALLDIAG ? 9 : 1);
    int64_t blocks = (n + kThreads * 8 - 1) / (kThreads * 8);
    if (blocks > kMaxBlocks) blocks = kMaxBlocks;
    SP_REQUIRE((size_t)(blocks < 1 ? 1 : blocks) * kRec * sizeof(double) <= sizeof(scratch),
               "reduce_noalloc: the input needs more block records than the scratch array holds");
    return sp_reduce(kind, x, result, ws, wparams, stream);
}

extern "C" int sp_reduce_combine(int kind, const double* records, int nrecords, int64_t count,
                                 double* result, void* stream) {
    SP_REQUIRE(records && result && nrecords > 0, "reduce_combine: bad arguments");
    SP_REQUIRE(kind >= SP_R_SUM && kind <= SP_R_ANYSMALLDIAG, "reduce_combine: unknown kind");
    hipLaunchKernelGGL(k_reduce_stage2, dim3(1), dim3(kCThreads), 0, as_stream(stream), kind,
                       records, nrecords, count, result);
    return launch_status("sp_reduce_combine");
}

extern "C" int sp_reduce_combine_batch(const sp_combine_desc* descs, int n, void* stream) {
    SP_REQUIRE(descs && n > 0, "reduce_combine_batch: bad arguments");
    hipLaunchKernelGGL(k_reduce_combine_batch, dim3((unsigned)n), dim3(kCThreads), 0,
                       as_stream(stream), descs);
    return launch_status("sp_reduce_combine_batch");
}

extern "C" int sp_reduce_combine_multi(const sp_combine_desc* descs, int n, void* stream) {
    SP_REQUIRE(descs && n > 0 && n <= SP_COMBINE_MULTI_MAX, "reduce_combine_multi: bad arguments");
    CombineArgs a;
    for (int i = 0; i < n; ++i) a.d[i] = descs[i];
    hipLaunchKernelGGL(k_reduce_combine_multi, dim3((unsigned)n), dim3(kCThreads), 0,
                       as_stream(stream), a);
    return launch_status("sp_reduce_combine_multi");
}

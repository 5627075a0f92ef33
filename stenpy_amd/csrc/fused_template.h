// Device-side support for the run-time generated fused stencil kernels.
//
// Pasted verbatim (after sp_sym3.h) at the top of every hipRTC source, so it must stay
// self-contained.  The generated part defines the per-graph evaluation functions; this file
// provides coordinate transforms, LDS index mapping, BF16 stores, compensated block
// reductions and the reduction record layout shared with sp_reduce.hip's final combine
// (4 doubles per block: {s, c} sums, {s1, c1, s2, c2} variance, {lo, hi} min/max).
#pragma once

#ifndef INFINITY
#define INFINITY __builtin_inff()
#endif
#ifndef NAN
#define NAN __builtin_nanf("")
#endif

typedef long long i64;
typedef unsigned long long u64;

// ---- coordinates --------------------------------------------------------------------------
__device__ __forceinline__ int wrapc(int c, int n) { return c < 0 ? c + n : (c >= n ? c - n : c); }
// any-period wrap, used for halo loads (always a valid index)
__device__ __forceinline__ int wrapm(int c, int n) { int r = c % n; return r < 0 ? r + n : r; }
__device__ __forceinline__ int clampc(int c, int n) { return c < 0 ? 0 : (c >= n ? n - 1 : c); }
// radius-1 mirror without repeating the edge (size-1 axes are zero padded by the caller)
__device__ __forceinline__ int reflc(int c, int n) {
    return c < 0 ? (-c < n - 1 ? -c : n - 1) : (c >= n ? (2 * (n - 1) - c > 0 ? 2 * (n - 1) - c : 0) : c);
}
__device__ __forceinline__ bool inside(int c, int n) { return c >= 0 && c < n; }

// Tile-local index of absolute coordinate c for a tile starting at t0 with T cells and halo H
// on an axis of length n.  Coordinates reached through a periodic wrap are mapped back to
// the near side; the result is clamped so that discarded (select-masked) lanes stay inside
// the LDS array.
__device__ __forceinline__ int lmap(int c, int t0, int T, int H, int n) {
    int l = c - t0;
    if (l >= T + H) l -= n;
    if (l < -H) l += n;
    l = l < -H ? -H : l;
    l = l > T + H - 1 ? T + H - 1 : l;
    return l + H;
}

// Orders one wave's LDS accesses across its lanes (a cross-lane exchange through LDS).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---- stores -----------------------------------------------------------------------------------
__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ void st_f32(float* p, i64 i, float v) { __builtin_nontemporal_store(v, p + i); }
__device__ __forceinline__ void st_bf16(unsigned short* p, i64 i, float v) { p[i] = f2bf(v); }

// ---- compensated accumulators ----------------------------------------------------------------
// Each function that forms an error term turns FMA contraction off: contracting x * k into the
// running sum (t = fma(x, k, s)) would leave t without the separately rounded product that the
// two-sum and the FMA split of the product assume.
struct RSum {  // Neumaier sum
    float s, c;
    __device__ __forceinline__ void init() { s = 0.f; c = 0.f; }
    __device__ __forceinline__ void add(float x) {
#pragma clang fp contract(off)
        float t = s + x;
        float bb = t - s;
        c += (s - (t - bb)) + (x - bb);
        s = t;
    }
};
__device__ __forceinline__ void rsum_add_prod(RSum& a, float x, float k) {  // a += x*k exactly split
#pragma clang fp contract(off)
    const float q = x * k;
    a.add(q);
    a.c += fmaf(x, k, -q);
}
struct RVar {  // compensated first and second moments (exact squares via FMA)
    RSum s1, s2;
    __device__ __forceinline__ void init() { s1.init(); s2.init(); }
    __device__ __forceinline__ void add(float x) {
#pragma clang fp contract(off)
        s1.add(x);
        const float q = x * x;
        s2.add(q);
        s2.c += fmaf(x, x, -q);
    }
};
__device__ __forceinline__ void rsum_add_sq(RSum& a, float x) {  // a += x*x exactly split
#pragma clang fp contract(off)
    const float q = x * x;
    a.add(q);
    a.c += fmaf(x, x, -q);
}
struct RMinMax {
    float lo, hi;
    __device__ __forceinline__ void init() { lo = INFINITY; hi = -INFINITY; }
    __device__ __forceinline__ void add(float x) {
        if (x != x) { lo = x; hi = x; }
        else if (lo == lo) { lo = fminf(lo, x); hi = fmaxf(hi, x); }
    }
};

__device__ __forceinline__ void merge_sum(RSum& a, float s, float c) {
#pragma clang fp contract(off)
    float t = a.s + s;
    float bb = t - a.s;
    float e = (a.s - (t - bb)) + (s - bb);
    a.s = t;
    a.c = a.c + c + e;
}
__device__ __forceinline__ void merge_mm(RMinMax& a, float lo, float hi) {
    if (a.lo != a.lo || lo != lo) { a.lo = a.hi = NAN; return; }
    a.lo = fminf(a.lo, lo);
    a.hi = fmaxf(a.hi, hi);
}

// Block-wide reductions; ``scratch`` holds at least blockDim.x/64 * 4 floats.  Thread 0 of
// the block writes the 4-double record.
__device__ __forceinline__ void block_sum(RSum a, float* scratch, double* rec, double scale = 1.0) {
    for (int m = 32; m >= 1; m >>= 1) merge_sum(a, __shfl_xor(a.s, m), __shfl_xor(a.c, m));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) { scratch[2 * w] = a.s; scratch[2 * w + 1] = a.c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        RSum r; r.s = scratch[0]; r.c = scratch[1];
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) merge_sum(r, scratch[2 * i], scratch[2 * i + 1]);
        rec[0] = (double)r.s * scale; rec[1] = (double)r.c * scale; rec[2] = 0.0; rec[3] = 0.0;
    }
}
__device__ __forceinline__ void block_var(RVar a, float* scratch, double* rec) {
    for (int m = 32; m >= 1; m >>= 1) {
        merge_sum(a.s1, __shfl_xor(a.s1.s, m), __shfl_xor(a.s1.c, m));
        merge_sum(a.s2, __shfl_xor(a.s2.s, m), __shfl_xor(a.s2.c, m));
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
        scratch[4 * w] = a.s1.s; scratch[4 * w + 1] = a.s1.c;
        scratch[4 * w + 2] = a.s2.s; scratch[4 * w + 3] = a.s2.c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        RSum r1, r2;
        r1.s = scratch[0]; r1.c = scratch[1]; r2.s = scratch[2]; r2.c = scratch[3];
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) {
            merge_sum(r1, scratch[4 * i], scratch[4 * i + 1]);
            merge_sum(r2, scratch[4 * i + 2], scratch[4 * i + 3]);
        }
        rec[0] = r1.s; rec[1] = r1.c; rec[2] = r2.s; rec[3] = r2.c;
    }
}
__device__ __forceinline__ void block_mm(RMinMax a, float* scratch, double* rec) {
    for (int m = 32; m >= 1; m >>= 1) merge_mm(a, __shfl_xor(a.lo, m), __shfl_xor(a.hi, m));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) { scratch[2 * w] = a.lo; scratch[2 * w + 1] = a.hi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        RMinMax r; r.lo = scratch[0]; r.hi = scratch[1];
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) merge_mm(r, scratch[2 * i], scratch[2 * i + 1]);
        rec[0] = r.lo; rec[1] = r.hi; rec[2] = 0.0; rec[3] = 0.0;
    }
}

// Per-axis quadrature weight as an integer numerator k over 6, the weight being k * dx / 6
// (Simpson with a trailing trapezoid for even n; trapezoid for n < 3 or method 1), the weights
// of stenpy_engine.py:2415-2457 applied separably.  The product of three axes' numerators (at
// most 8^3) is exact in FP32; dx0 dx1 dx2 / 216 multiplies the block record once, in FP64.
__device__ __forceinline__ float quad_k(int i, int n, int method) {
    if (method == 1 || n < 3) {
        if (n < 2) return 0.f;
        return (i == 0 || i == n - 1) ? 3.f : 6.f;
    }
    const int m = (n % 2 == 0) ? n - 1 : n;
    float k = 0.f;
    if (i < m) k = (i == 0 || i == m - 1) ? 2.f : ((i % 2 == 1) ? 8.f : 4.f);
    if (n % 2 == 0 && i >= n - 2) k += 3.f;
    return k;
}

// torch.gradient one-sided formulas (interior ((x+)-(x-))/h/2)
__device__ __forceinline__ float tg_mid(float xp, float xm, float h) { return ((xp - xm) / h) * 0.5f; }
__device__ __forceinline__ float tg_left2(float x0, float x1, float x2, float h) {
    return ((-1.5f * x0 + 2.0f * x1) - 0.5f * x2) / h;
}
__device__ __forceinline__ float tg_right2(float xm2, float xm1, float x0, float h) {
    return ((0.5f * xm2 - 2.0f * xm1) + 1.5f * x0) / h;
}
__device__ __forceinline__ float tg_left1(float x0, float x1, float h) { return (x1 - x0) / h; }
__device__ __forceinline__ float tg_right1(float xm1, float x0, float h) { return (x0 - xm1) / h; }

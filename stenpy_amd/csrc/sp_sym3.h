// Device routines for per-voxel 3x3 tensor algebra.  Self-contained (no includes) so the
// run-time compiled fused kernels can paste this file into their source unchanged.
//
// Matrices are row-major float[9].  Semantics follow the reference CPU path:
//   trace        sum of the diagonal, folded left            (stenpy_engine.py:2674-2680)
//   determinant  det(A + reg I); nan -> 0, +-inf -> +-1e30, clamp to +-1e30 (2683-2712)
//   eigvalsh     eigenvalues of the lower triangle, ascending (2715-2721)
//   deviatoric   A - tr(A)/3 I; non-finite -> 0                (2852-2867)
//   inverse      adjugate / det                               (2724-2730)
#pragma once

#define SP_SYM3_FP64_BAND 0.05f  // 1-|r| below this: near-double-degenerate -> FP64 branch

__device__ __forceinline__ float sym3_trace(const float* a) { return (a[0] + a[4]) + a[8]; }

// a*d - b*c with one rounding error (Kahan's FMA scheme).
__device__ __forceinline__ float det2_fma(float a, float b, float c, float d) {
    const float w = b * c;
    const float e = fmaf(-b, c, w);
    const float f = fmaf(a, d, -w);
    return f + e;
}

__device__ __forceinline__ float clamp_det(float d) {
    if (d != d) return 0.0f;
    if (d > 1e30f) return 1e30f;
    if (d < -1e30f) return -1e30f;
    return d;
}

__device__ __forceinline__ float sym3_det_regularised(const float* m, float reg) {
    const float a00 = m[0] + reg, a11 = m[4] + reg, a22 = m[8] + reg;
    const float c0 = det2_fma(a11, m[5], m[7], a22);
    const float c1 = det2_fma(m[3], m[5], m[6], a22);
    const float c2 = det2_fma(m[3], a11, m[6], m[7]);
    return clamp_det((a00 * c0 - m[1] * c1) + m[2] * c2);
}

__device__ __forceinline__ void sort3(float* e) {
    float t;
    if (e[0] > e[1]) { t = e[0]; e[0] = e[1]; e[1] = t; }
    if (e[1] > e[2]) { t = e[1]; e[1] = e[2]; e[2] = t; }
    if (e[0] > e[1]) { t = e[0]; e[0] = e[1]; e[1] = t; }
}

// Eigenvalues of a symmetric 3x3 (lower triangle, ascending, as eigvalsh), trigonometric closed
// form: q = tr/3, p = sqrt(|A - qI|^2 / 6), r = det((A - qI)/p) / 2, eigenvalues
// q + 2p cos(phi/3 + 2k pi/3) with phi = acos(r).  The angle is taken as
// acos(r) = 2 atan2(sqrt(1 - r), sqrt(1 + r)), accurate where acos(r) itself is not (|r| -> 1).
// Near a double eigenvalue (1 - |r| < SP_SYM3_FP64_BAND) r formed in FP32 has lost the digits
// that matter, so there the moments q, p, r and the final combination are recomputed in FP64
// (products of the FP32 entries); the trigonometry is FP32 in both branches.  A diagonal
// matrix returns its sorted diagonal exactly.
template <typename T>
__device__ __forceinline__ void sym3_moments(const float* m, T& q, T& p, T& r) {
    const T a00 = m[0], a11 = m[4], a22 = m[8];
    const T a10 = m[3], a20 = m[6], a21 = m[7];
    q = (a00 + a11 + a22) / T(3);
    const T b00 = a00 - q, b11 = a11 - q, b22 = a22 - q;
    const T p1 = a10 * a10 + a20 * a20 + a21 * a21;
    const T p2 = b00 * b00 + b11 * b11 + b22 * b22 + T(2) * p1;
    p = sqrt(p2 / T(6));
    const T detB = b00 * (b11 * b22 - a21 * a21) - a10 * (a10 * b22 - a21 * a20) +
                   a20 * (a10 * a21 - b11 * a20);
    r = p > T(0) ? detB / (T(2) * p * p * p) : T(0);
    r = r < T(-1) ? T(-1) : (r > T(1) ? T(1) : r);
}
template <typename T>
__device__ __forceinline__ void sym3_from_moments(T q, T p, T r, float* e) {
    const float sm = sqrtf((float)(T(1) - r)), sp = sqrtf((float)(T(1) + r));
    const float phi = 2.0f * atan2f(sm, sp) / 3.0f;
    const T c2 = (T)cosf(phi), c0 = (T)cosf(phi + 2.0943951023931954923f);
    const T e2 = q + T(2) * p * c2, e0 = q + T(2) * p * c0;
    e[0] = (float)e0;
    e[1] = (float)(T(3) * q - e0 - e2);
    e[2] = (float)e2;
}
__device__ __forceinline__ void sym3_eigvalsh(const float* m, float* e) {
    if (m[3] == 0.0f && m[6] == 0.0f && m[7] == 0.0f) {
        e[0] = m[0]; e[1] = m[4]; e[2] = m[8];
        sort3(e);
        return;
    }
    float q, p, r;
    sym3_moments<float>(m, q, p, r);
    if (1.0f - fabsf(r) < SP_SYM3_FP64_BAND || !(r == r)) {
        double qd, pd, rd;
        sym3_moments<double>(m, qd, pd, rd);
        sym3_from_moments<double>(qd, pd, rd, e);
    } else {
        sym3_from_moments<float>(q, p, r, e);
    }
    sort3(e);
}

__device__ __forceinline__ float finite_or_zero(float x) {
    return (x == x && fabsf(x) <= 3.402823466e38f) ? x : 0.0f;
}

__device__ __forceinline__ void sym3_deviatoric(const float* m, float* d) {
    const float t = sym3_trace(m) / 3.0f;
#pragma unroll
    for (int i = 0; i < 9; ++i) d[i] = finite_or_zero((i % 4 == 0) ? m[i] - t : m[i]);
}

__device__ __forceinline__ void mat3_inverse(const float* m, float* inv) {
    const float c00 = det2_fma(m[4], m[5], m[7], m[8]);
    const float c01 = -det2_fma(m[3], m[5], m[6], m[8]);
    const float c02 = det2_fma(m[3], m[4], m[6], m[7]);
    const float det = (m[0] * c00 + m[1] * c01) + m[2] * c02;
    const float id = 1.0f / det;
    inv[0] = c00 * id;
    inv[3] = c01 * id;
    inv[6] = c02 * id;
    inv[1] = -det2_fma(m[1], m[2], m[7], m[8]) * id;
    inv[4] = det2_fma(m[0], m[2], m[6], m[8]) * id;
    inv[7] = -det2_fma(m[0], m[1], m[6], m[7]) * id;
    inv[2] = det2_fma(m[1], m[2], m[4], m[5]) * id;
    inv[5] = -det2_fma(m[0], m[2], m[3], m[5]) * id;
    inv[8] = det2_fma(m[0], m[1], m[3], m[4]) * id;
}

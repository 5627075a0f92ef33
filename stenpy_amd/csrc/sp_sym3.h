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

__device__ __forceinline__ float sym3_trace(const float* a) { return (a[0] + a[4]) + a[8]; }

// x / 3, correctly rounded, without the division sequence: q = x r with r = RN(1/3), then one
// FMA correction q + RN(x - 3q) r (Markstein's scheme for a divisor known in advance).  Checked
// over every finite FP32 x against the FP64 quotient rounded to FP32 (x / 3 is never within
// 2^-51 of an FP32 rounding midpoint, so that rounding is the correct one); the sign is x's
// (the correction turns -0 into +0), infinities and NaN return the product x r.  Four
// instructions instead of about a dozen.
__device__ __forceinline__ float div3f(float x) {
    const float r = 0.3333333432674407958984375f;   // RN(1/3)
    const float q = x * r;
    const float q2 = fmaf(fmaf(-q, 3.0f, x), r, q);
    return q2 == q2 ? copysignf(q2, x) : q;
}

// a*d - b*c with one rounding error (Kahan's FMA scheme).
__device__ __forceinline__ float det2_fma(float a, float b, float c, float d) {
    const float w = b * c;
    const float e = fmaf(-b, c, w);
    const float f = fmaf(a, d, -w);
    return f + e;
}

__device__ __forceinline__ float clamp_det(float d) {
    // NaN -> 0, then a clamp to [-1e30, 1e30] (one median-of-three on the GPU)
    return d == d ? fminf(fmaxf(d, -1e30f), 1e30f) : 0.0f;
}

__device__ __forceinline__ float sym3_det_regularised(const float* m, float reg) {
    const float a00 = m[0] + reg, a11 = m[4] + reg, a22 = m[8] + reg;
    const float c0 = det2_fma(a11, m[5], m[7], a22);
    const float c1 = det2_fma(m[3], m[5], m[6], a22);
    const float c2 = det2_fma(m[3], a11, m[6], m[7]);
    return clamp_det((a00 * c0 - m[1] * c1) + m[2] * c2);
}

// Square root for the closed form's internal terms: the hardware instruction (1 ulp) on the
// GPU instead of the correctly rounded sequence; the angle and the scale p already carry errors
// of a few ulp of |B|.
#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC_RTC__)
#define sym3_sqrt(x) __builtin_amdgcn_sqrtf(x)
#define sym3_rcp(x) __builtin_amdgcn_rcpf(x)
#define sym3_rsq(x) __builtin_amdgcn_rsqf(x)
#else
#define sym3_sqrt(x) sqrtf(x)
#define sym3_rcp(x) (1.0f / (x))
#define sym3_rsq(x) (1.0f / sqrtf(x))
#endif

// 1 / max(sqrt(s), eps) for a sum of squares s >= 0, the scale of a unit normal g / max(|g|, eps):
// the hardware reciprocal square root (1 ulp) refined by one Newton step in FP32,
// r = r0 + r0 (1/2 - (s r0) r0 / 2), with the product s r0 rounded and the rest fused.  Each
// component g_i r is then within about 2 ulp of g_i / max(|g|, eps), against about 1.5 ulp for the
// FP32 formula that rounds the norm and the quotient.  Below eps^2 the same step runs on eps^2
// (r = 1 / eps to about 1 ulp); an infinite s gives 0 (as the formula); a NaN s gives NaN.
__device__ __forceinline__ float unit_rnorm(float s, float eps) {
    const float sc = s < eps * eps ? eps * eps : s;
    const float r0 = sym3_rsq(sc);
    const float e = fmaf(-0.5f * (sc * r0), r0, 0.5f);
    const float r = fmaf(r0, e, r0);
    return s > 3.4028234663852886e38f ? 0.0f : r;
}

// atan2(y, x) for y >= 0 (or NaN), the closed form's angle in [0, pi]: a = min / max of |x| and
// y (the hardware reciprocal and one FMA correction), atan a = a + a^3 P(a^2) on [0, 1] with P of
// degree 8 fitted like the sine and cosine below (0.94 ulp in FP32), then pi/2 - r above the
// diagonal and pi - r for x < 0 (-0 included, as atan2).  NaN in either argument gives NaN.
__device__ __forceinline__ float atan2_upper(float y, float x) {
    const float ax = fabsf(x);
    const bool sw = y > ax;
    const float lo = sw ? ax : y, hi = sw ? y : ax;
    const float rc = sym3_rcp(hi);
    float a = lo * rc;
    a = fmaf(fmaf(-a, hi, lo), rc, a);
    a = hi == 0.0f ? 0.0f : a;                    // atan2(0, +-0): a = 0
    const float t = a * a;
    float p = fmaf(t, -1.789061236e-03f, 1.089788415e-02f);
    p = fmaf(p, t, -3.115503304e-02f);
    p = fmaf(p, t, 5.794504657e-02f);
    p = fmaf(p, t, -8.403480053e-02f);
    p = fmaf(p, t, 1.095253304e-01f);
    p = fmaf(p, t, -1.426439285e-01f);
    p = fmaf(p, t, 1.999857426e-01f);
    p = fmaf(p, t, -3.333330154e-01f);
    float r = fmaf(a * t, p, a);
    r = sw ? 1.57079637050628662f - r : r;
    return __builtin_signbit(x) ? 3.14159274101257324f - r : r;
}

// sin and cos of phi in [0, pi/3] (the closed form's angle), without range reduction:
// sin = x + x^3 S(x^2), cos = 1 - x^2/2 + x^4 C(x^2), S and C of degree 4 fitted by weighted
// least squares on Chebyshev nodes for relative error (fit error below 2^-30); evaluated in
// FP32 they are within 1.03 ulp (sin) and 1.26 ulp (cos) over the interval.
__device__ __forceinline__ void sincos_0_pi3(float x, float* sn, float* cs) {
    const float t = x * x;
    float s = fmaf(t, -2.453835179e-08f, 2.755117066e-06f);
    s = fmaf(s, t, -1.984123664e-04f);
    s = fmaf(s, t, 8.333332837e-03f);
    s = fmaf(s, t, -1.666666716e-01f);
    *sn = fmaf(x * t, s, x);
    float c = fmaf(t, 2.046258718e-09f, -2.755155037e-07f);
    c = fmaf(c, t, 2.480154944e-05f);
    c = fmaf(c, t, -1.388888923e-03f);
    c = fmaf(c, t, 4.166666791e-02f);
    *cs = fmaf(t * t, c, fmaf(-0.5f, t, 1.0f));
}

__device__ __forceinline__ void sort3(float* e) {
    float t;
    if (e[0] > e[1]) { t = e[0]; e[0] = e[1]; e[1] = t; }
    if (e[1] > e[2]) { t = e[1]; e[1] = e[2]; e[2] = t; }
    if (e[0] > e[1]) { t = e[0]; e[0] = e[1]; e[1] = t; }
}

// Eigenvalues of a symmetric 3x3 (lower triangle, ascending, as eigvalsh), all in FP32.
// Trigonometric closed form: with q = tr/3, B = A - qI, p = sqrt(tr(B^2) / 6) the eigenvalues
// are q + 2p cos(phi/3 + 2k pi/3), where cos(phi) = det(B) / (2p^3).  The angle is not taken
// from that quotient (acos loses every digit of a near-double eigenvalue) but as
//     phi = atan2(sqrt(D), 3 sqrt(3) det(B)),   D = prod_{i<j} (l_i - l_j)^2,
// using sin(phi)^2 = 1 - cos(phi)^2 = D / (4 (tr(B^2)/2)^3).  D is evaluated as a sum of
// squares, which has no cancellation: D = det(V^T V) for the Vandermonde V_ij = l_i^j, and
// V^T V is the Gram matrix of I, B, B^2 under <X, Y> = tr(XY).  Writing a symmetric X as the
// 6-vector (x00, x11, x22, sqrt2 x01, sqrt2 x02, sqrt2 x12), Cauchy-Binet turns that 3x3 Gram
// determinant into the sum of the squared 3x3 minors of the 6x3 matrix [I | B | B^2]; the 19
// nonzero minors group by how many diagonal rows they take (3, 2 or 1):
//     D = m3^2 + 2 sum_{i<j, o} ((b_j - b_i) c_o - (c_j - c_i) b_o)^2 + 12 sum_{o<o'} (b_o c_o' - c_o b_o')^2
// with b = B, c = B^2, i, j diagonal and o, o' off-diagonal positions.  Each minor carries an
// absolute error of a few ulp of |B|^3, so the angle is accurate to a few ulp for every
// spectrum, double and triple eigenvalues included.  B is first scaled exactly, by a power of
// two that brings its largest entry into [1, 2), which keeps |B|^6 inside the FP32 range.  A
// diagonal matrix returns its sorted diagonal exactly.
__device__ __forceinline__ void sym3_eigvalsh(const float* m, float* e) {
    if (m[3] == 0.0f && m[6] == 0.0f && m[7] == 0.0f) {
        e[0] = m[0]; e[1] = m[4]; e[2] = m[8];
        sort3(e);
        return;
    }
    const float q = div3f((m[0] + m[4]) + m[8]);
    float b0 = m[0] - q, b1 = m[4] - q, b2 = m[8] - q, o01 = m[3], o02 = m[6], o12 = m[7];
    const float mx = fmaxf(fmaxf(fmaxf(fabsf(b0), fabsf(b1)), fmaxf(fabsf(b2), fabsf(o01))),
                           fmaxf(fabsf(o02), fabsf(o12)));
    if (mx == 0.0f) {
        // a multiple of I (fmaxf skips NaN: an off-diagonal NaN still reaches the result)
        e[0] = e[1] = e[2] = q + 0.0f * ((o01 + o02) + o12);
        return;
    }
    // exact scaling by a power of two: the largest entry lands in [1, 2)
    int k;
    (void)frexpf(mx, &k);
    b0 = ldexpf(b0, 1 - k); b1 = ldexpf(b1, 1 - k); b2 = ldexpf(b2, 1 - k);
    o01 = ldexpf(o01, 1 - k); o02 = ldexpf(o02, 1 - k); o12 = ldexpf(o12, 1 - k);
    // C = B^2
    const float c0 = b0 * b0 + o01 * o01 + o02 * o02;
    const float c1 = o01 * o01 + b1 * b1 + o12 * o12;
    const float c2 = o02 * o02 + o12 * o12 + b2 * b2;
    const float c01 = b0 * o01 + o01 * b1 + o02 * o12;
    const float c02 = b0 * o02 + o01 * o12 + o02 * b2;
    const float c12 = o01 * o02 + b1 * o12 + o12 * b2;
    // minors with the three diagonal rows, two, and one
    const float m3 = ((b1 * c2 - b2 * c1) - (b0 * c2 - b2 * c0)) + (b0 * c1 - b1 * c0);
    const float db[3] = {b1 - b0, b2 - b0, b2 - b1}, dc[3] = {c1 - c0, c2 - c0, c2 - c1};
    const float bo[3] = {o01, o02, o12}, co[3] = {c01, c02, c12};
    float s2 = 0.0f;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            const float t = db[i] * co[o] - dc[i] * bo[o];
            s2 = fmaf(t, t, s2);
        }
    const float t01 = o01 * c02 - c01 * o02, t02 = o01 * c12 - c01 * o12, t12 = o02 * c12 - c02 * o12;
    const float s1 = (t01 * t01 + t02 * t02) + t12 * t12;
    const float disc = fmaf(m3, m3, fmaf(2.0f, s2, 12.0f * s1));
    const float detb = (b0 * (b1 * b2 - o12 * o12) - o01 * (o01 * b2 - o12 * o02)) +
                       o02 * (o01 * o12 - b1 * o02);
    const float phi = div3f(atan2_upper(sym3_sqrt(disc), 5.19615242270663188f * detb));
    // (c0 + c1 + c2) / 6 as div3f(.) / 2: exact halving, the sum being >= 1 after the scaling
    const float p2 = ldexpf(2.0f * sym3_sqrt(div3f((c0 + c1) + c2) * 0.5f), k - 1);
    float sn, cs;
    sincos_0_pi3(phi, &sn, &cs);
    e[2] = fmaf(p2, cs, q);
    // cos(phi + 2 pi / 3) = -cos(phi) / 2 - sin(phi) sqrt(3) / 2
    e[0] = fmaf(p2, fmaf(-0.86602540378443865f, sn, -0.5f * cs), q);
    e[1] = (3.0f * q - e[0]) - e[2];
    sort3(e);
}

__device__ __forceinline__ float finite_or_zero(float x) {
    return (x == x && fabsf(x) <= 3.402823466e38f) ? x : 0.0f;
}

__device__ __forceinline__ void sym3_deviatoric(const float* m, float* d) {
    const float t = div3f(sym3_trace(m));
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

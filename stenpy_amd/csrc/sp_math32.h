// FP32 exp, tanh, log, sin and cos for the generated and the point-wise kernels, accurate to
// 0.5 ulp + 2^-16 ulp: each result is the FP32 rounding of an FP64 value whose relative error is
// below 2^-45, so it is the correctly rounded result except where the exact value lies that
// close to a rounding midpoint.  They replace FP32 rounding of the FP64 library functions, which
// gives the same accuracy with longer code: those carry polynomials and extra-precise steps for
// 53-bit results (gfx950 code for one call: tanh 188 instructions against 73 here, log 117
// against 77, exp about equal).  They read no tables: a table load in a kernel that keeps global
// loads in flight makes each call wait for all of them.
//
//   exp(x)   x = k ln2 + r, |r| <= ln2/2, exp(x) = 2^k (1 + p(r)), p ~ expm1 of degree 10
//            (weighted least squares on Chebyshev nodes, relative error 2^-49; the other fits
//            below are made the same way).
//   tanh(x)  |tanh x| = E / (E + 2), E = expm1(2|x|) = (2^k - 1) + 2^k p(r), which keeps the
//            relative accuracy of E for small |x|; 1 / (E + 2) is the FP32 reciprocal refined
//            by one Newton step in FP64 (relative error 2^-46).
//   log(x)   x = 2^e m, m in [sqrt(1/2), sqrt(2)), f = m - 1 (exact), s = f / (2 + f) (same
//            reciprocal), log x = e ln2 + 2 atanh(s) with 2 atanh(s) = 2s + s^3 q(s^2), q of
//            degree 5 (relative error 2^-52).
//   sin(x), cos(x)  for |x| < 2^19: x = k pi/2 + r, |r| <= pi/4, with pi/2 in two parts (the
//            remainder's error is below 2^-89 absolute), then sin r = r + r^3 S(r^2) and
//            cos r = 1 - r^2/2 + r^4 C(r^2), S and C of degree 4 (errors 2^-47 relative and 2^-53),
//            picked and signed by k mod 4.  Larger arguments, infinities and NaN go to the FP64
//            library functions.
//
// Pasted verbatim into every hipRTC source and included by the plain point-wise kernels.  The
// same text builds as plain C++ for the host, where the FP32 reciprocal is a division, so the
// functions can be checked over all FP32 inputs.
#pragma once

#if defined(__HIPCC_RTC__) || defined(__HIPCC__)
#define SPM_FN __device__ __forceinline__
#define spm_rcp(v) __builtin_amdgcn_rcpf(v)
#else
#include <math.h>
#define SPM_FN static inline
#define spm_rcp(v) (1.0f / (v))
#endif

// k = round(x / ln2) and r = x - k ln2 for |x| <= 160.  The integer k is read from the low word
// of x / ln2 + 1.5 2^52; r is formed by one fused multiply-add with ln2 rounded to double,
// whose error (< 2^-54 per unit of k, |k| < 2^8) stays below 2^-46.  A NaN x gives a NaN r.
SPM_FN double spm_reduce(double x, int* k) {
    const double t = __builtin_fma(x, 0x1.71547652b82fep+0, 0x1.8p52);
    *k = (int)(unsigned)__builtin_bit_cast(unsigned long long, t);
    return __builtin_fma(t - 0x1.8p52, -0x1.62e42fefa39efp-1, x);
}

// 2^k for |k| < 1022, built in the exponent field
SPM_FN double spm_pow2(int k) {
    return __builtin_bit_cast(double, (unsigned long long)(unsigned)(k + 1023) << 52);
}

// expm1(r) for |r| <= ln2/2: r + r^2/2 + r^3 g(r), g of degree 7
SPM_FN double spm_expm1_poly(double r) {
    double g = __builtin_fma(r, 0x1.28167f74af9aep-22, 0x1.72dadaeb1a523p-19);
    g = __builtin_fma(g, r, 0x1.a01a33359b3aap-16);
    g = __builtin_fma(g, r, 0x1.a0199b64d8a49p-13);
    g = __builtin_fma(g, r, 0x1.6c16c152a45a6p-10);
    g = __builtin_fma(g, r, 0x1.111111216520fp-7);
    g = __builtin_fma(g, r, 0x1.555555555b074p-5);
    g = __builtin_fma(g, r, 0x1.555555555213dp-3);
    return __builtin_fma(r * r, __builtin_fma(g, r, 0.5), r);
}

// 1 / d for d in [1, 2^31]
SPM_FN double spm_recip(double d) {
    const double y = (double)spm_rcp((float)d);
    return __builtin_fma(y, __builtin_fma(-d, y, 1.0), y);
}

SPM_FN float spm_expf(float xf) {
    double x = (double)xf;
    x = x > 100.0 ? 100.0 : (x < -160.0 ? -160.0 : x);  // beyond these the FP32 result is inf / 0
    int k;
    const double r = spm_reduce(x, &k);
    const double s = spm_pow2(k);
    return (float)__builtin_fma(s, spm_expm1_poly(r), s);
}

SPM_FN float spm_tanhf(float xf) {
    const double a = __builtin_fabs((double)xf);
    int k;
    const double r = spm_reduce(a > 10.0 ? 20.0 : a + a, &k);  // tanh rounds to 1 beyond 9.02
    const double s = spm_pow2(k);
    const double E = __builtin_fma(s, spm_expm1_poly(r), s - 1.0);
    const float t = (float)__builtin_copysign(E * spm_recip(E + 2.0), (double)xf);
    return xf != xf ? xf : t;
}

SPM_FN float spm_logf(float xf) {
    unsigned u = __builtin_bit_cast(unsigned, xf);
    int e = 0;
    if (u < 0x00800000u) {  // subnormal: scale into the normal range
        u = __builtin_bit_cast(unsigned, xf * 0x1p23f);
        e = -23;
    }
    // m = x 2^-sh in [sqrt(1/2), sqrt(2)): the exponent field offset by sqrt(1/2)'s bits
    const int sh = (int)(u - 0x3f3504f3u) >> 23;
    e += sh;
    const float m = __builtin_bit_cast(float, u - ((unsigned)sh << 23));
    const double f = (double)(m - 1.0f);
    const double s = f * spm_recip(f + 2.0);
    const double z = s * s;
    double q = __builtin_fma(z, 0x1.5883d9d9dbc7fp-3, 0x1.732b0622c37c9p-3);
    q = __builtin_fma(q, z, 0x1.c72281feea166p-3);
    q = __builtin_fma(q, z, 0x1.249241634527cp-2);
    q = __builtin_fma(q, z, 0x1.999999a2b9f9fp-2);
    q = __builtin_fma(q, z, 0x1.5555555553898p-1);
    const double l = __builtin_fma(s * z, q, s + s);
    const float v = (float)__builtin_fma((double)e, 0x1.62e42fefa39efp-1, l);
    // zero, negative, infinite and NaN inputs
    return xf > 0.0f ? (xf == __builtin_inff() ? xf : v) : (xf == 0.0f ? -__builtin_inff() : __builtin_nanf(""));
}

// sin (c = 0) or cos (c = 1)
SPM_FN float spm_trigf(float xf, int c) {
    const double x = (double)xf;
    if (!(__builtin_fabs(x) < 0x1p19)) return (float)(c ? cos(x) : sin(x));
    const double t = __builtin_fma(x, 0x1.45f306dc9c883p-1, 0x1.8p52);
    const int k = (int)(unsigned)__builtin_bit_cast(unsigned long long, t) + c;
    const double kd = t - 0x1.8p52;
    const double r = __builtin_fma(kd, -0x1.1a62633145c07p-54, __builtin_fma(kd, -0x1.921fb54442d18p+0, x));
    const double z = r * r;
    double ps = __builtin_fma(z, -0x1.a963efc6db6e5p-26, 0x1.71d77395617ffp-19);
    ps = __builtin_fma(ps, z, -0x1.a019f9457a92ap-13);
    ps = __builtin_fma(ps, z, 0x1.1111110c8f29fp-7);
    ps = __builtin_fma(ps, z, -0x1.5555555552280p-3);
    double pc = __builtin_fma(z, 0x1.1bc3bffbe4db7p-29, -0x1.27e01c91eb775p-22);
    pc = __builtin_fma(pc, z, 0x1.a019fa875674fp-16);
    pc = __builtin_fma(pc, z, -0x1.6c16c167466abp-10);
    pc = __builtin_fma(pc, z, 0x1.5555555552f55p-5);
    const double v = (k & 1) ? __builtin_fma(z * z, pc, __builtin_fma(z, -0.5, 1.0)) : r * __builtin_fma(z, ps, 1.0);
    return (float)((k & 2) ? -v : v);
}
SPM_FN float spm_sinf(float x) { return spm_trigf(x, 0); }
SPM_FN float spm_cosf(float x) { return spm_trigf(x, 1); }

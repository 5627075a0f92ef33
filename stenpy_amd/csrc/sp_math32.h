// FP32 exp, expm1-based tanh and log for the run-time generated kernels, accurate to
// 0.5 ulp + 2^-18 ulp: each result is the FP32 rounding of an FP64 value whose relative error
// is below 2^-43, so it equals the correctly rounded result except where the exact value lies
// within 2^-43 (relative) of a rounding midpoint.  They replace FP32 rounding of the FP64
// library functions, which gives the same accuracy at about twice the FP64 work (the FP64
// functions carry polynomials long enough for 53-bit results and range checks FP32 inputs
// never need).
//
//   exp(x)   x = k ln2/64 + r, |r| <= ln2/128, exp(x) = 2^(k>>6) * 2^((k&63)/64) * exp(r), with
//            the 64 powers 2^(j/64) from a table (each the nearest double) and exp(r) - 1 as
//            its degree-4 Taylor polynomial (truncation r^5/120 < 2^-44).
//   tanh(x)  |tanh x| = E / (E + 2), E = expm1(2|x|) from the same reduction without the
//            cancellation of exp(2|x|) - 1: E = (s - 1) + s p(r), s = 2^(k/64), p of degree 5, and
//            1 / (E + 2) from the FP32 reciprocal refined by one Newton step (error 2^-45).
//   log(x)   x = 2^e m, m in [0.75, 1.5), c = 1/(1 + j/128) rounded to FP32 for j =
//            round(128 (m - 1)) in [-32, 64], r = m c - 1 exact to one rounding (|r| < 2^-7.6),
//            log x = e ln2 - log c + log1p(r), log1p(r) as its degree-6 Taylor polynomial and
//            -log c from a table (each the nearest double).
//
// Pasted verbatim into every hipRTC source.  The same text builds as plain C for the host,
// where the reciprocal is a division, so the functions can be checked over all FP32 inputs.
#pragma once

#ifdef __HIPCC_RTC__
#define SPM_FN __device__ __forceinline__
#define SPM_TAB __constant__
#define spm_rcp(v) __builtin_amdgcn_rcpf(v)
#else
#define SPM_FN static inline
#define SPM_TAB static const
#define spm_rcp(v) (1.0f / (v))
#endif

// 2^(j/64), j = 0..63
SPM_TAB double spm_exp2_64[64] = {
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0,
    0x1.0874518759bc8p+0, 0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0,
    0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0, 0x1.172b83c7d517bp+0,
    0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0,
    0x1.2d285a6e4030bp+0, 0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0,
    0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0, 0x1.3dea64c123422p+0,
    0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0,
    0x1.56f4736b527dap+0, 0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0,
    0x1.6247eb03a5585p+0, 0x1.6623882552225p+0, 0x1.6a09e667f3bcdp+0,
    0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0,
    0x1.868d99b4492edp+0, 0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0,
    0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0, 0x1.9c49182a3f090p+0,
    0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0,
    0x1.bcc1e904bc1d2p+0, 0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0,
    0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0, 0x1.d5818dcfba487p+0,
    0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0,
    0x1.fa7c1819e90d8p+0,
};
// c_j = 1 / (1 + j/128) rounded to FP32, j = -32..64
SPM_TAB float spm_log_c[97] = {
    0x1.5555560000000p+0f, 0x1.51d07e0000000p+0f, 0x1.4e5e0a0000000p+0f, 0x1.4afd6a0000000p+0f,
    0x1.47ae140000000p+0f, 0x1.446f860000000p+0f, 0x1.4141420000000p+0f, 0x1.3e22cc0000000p+0f,
    0x1.3b13b20000000p+0f, 0x1.3813820000000p+0f, 0x1.3521d00000000p+0f, 0x1.323e340000000p+0f,
    0x1.2f684c0000000p+0f, 0x1.2c9fb40000000p+0f, 0x1.29e4120000000p+0f, 0x1.27350c0000000p+0f,
    0x1.24924a0000000p+0f, 0x1.21fb780000000p+0f, 0x1.1f70480000000p+0f, 0x1.1cf06a0000000p+0f,
    0x1.1a7b960000000p+0f, 0x1.1811820000000p+0f, 0x1.15b1e60000000p+0f, 0x1.135c820000000p+0f,
    0x1.1111120000000p+0f, 0x1.0ecf560000000p+0f, 0x1.0c97140000000p+0f, 0x1.0a68100000000p+0f,
    0x1.0842100000000p+0f, 0x1.0624de0000000p+0f, 0x1.0410420000000p+0f, 0x1.0204080000000p+0f,
    0x1.0000000000000p+0f, 0x1.fc07f00000000p-1f, 0x1.f81f820000000p-1f, 0x1.f4465a0000000p-1f,
    0x1.f07c200000000p-1f, 0x1.ecc07c0000000p-1f, 0x1.e9131a0000000p-1f, 0x1.e573ac0000000p-1f,
    0x1.e1e1e20000000p-1f, 0x1.de5d6e0000000p-1f, 0x1.dae6080000000p-1f, 0x1.d77b660000000p-1f,
    0x1.d41d420000000p-1f, 0x1.d0cb580000000p-1f, 0x1.cd85680000000p-1f, 0x1.ca4b300000000p-1f,
    0x1.c71c720000000p-1f, 0x1.c3f8f00000000p-1f, 0x1.c0e0700000000p-1f, 0x1.bdd2b80000000p-1f,
    0x1.bacf920000000p-1f, 0x1.b7d6c40000000p-1f, 0x1.b4e81c0000000p-1f, 0x1.b203640000000p-1f,
    0x1.af286c0000000p-1f, 0x1.ac57020000000p-1f, 0x1.a98ef60000000p-1f, 0x1.a6d01a0000000p-1f,
    0x1.a41a420000000p-1f, 0x1.a16d400000000p-1f, 0x1.9ec8ea0000000p-1f, 0x1.9c2d140000000p-1f,
    0x1.99999a0000000p-1f, 0x1.970e500000000p-1f, 0x1.948b100000000p-1f, 0x1.920fb40000000p-1f,
    0x1.8f9c180000000p-1f, 0x1.8d30180000000p-1f, 0x1.8acb900000000p-1f, 0x1.886e600000000p-1f,
    0x1.8618620000000p-1f, 0x1.83c9780000000p-1f, 0x1.8181820000000p-1f, 0x1.7f40600000000p-1f,
    0x1.7d05f40000000p-1f, 0x1.7ad2200000000p-1f, 0x1.78a4c80000000p-1f, 0x1.767dce0000000p-1f,
    0x1.745d180000000p-1f, 0x1.7242880000000p-1f, 0x1.702e060000000p-1f, 0x1.6e1f760000000p-1f,
    0x1.6c16c20000000p-1f, 0x1.6a13ce0000000p-1f, 0x1.6816820000000p-1f, 0x1.661ec60000000p-1f,
    0x1.642c860000000p-1f, 0x1.623fa80000000p-1f, 0x1.6058160000000p-1f, 0x1.5e75bc0000000p-1f,
    0x1.5c98820000000p-1f, 0x1.5ac0560000000p-1f, 0x1.58ed240000000p-1f, 0x1.571ed40000000p-1f,
    0x1.5555560000000p-1f,
};
// -log(c_j), j = -32..64
SPM_TAB double spm_log_t[97] = {
    -0x1.269623134db8ap-2, -0x1.1bf99425a6b8cp-2, -0x1.1178e6c27e478p-2,
    -0x1.071385f4d5862p-2, -0x1.f991c3cb3b370p-3, -0x1.e530edde7100ep-3,
    -0x1.d10383e655e65p-3, -0x1.bd0874c3bd8abp-3, -0x1.a93ed8c8ad9cap-3,
    -0x1.95a5b2ef70165p-3, -0x1.823c18551a3bep-3, -0x1.6f01247756aaap-3,
    -0x1.5bf407b543db1p-3, -0x1.4913d2733b540p-3, -0x1.365fc6c159004p-3,
    -0x1.23d715e49c1f7p-3, -0x1.1178ee227e458p-3, -0x1.fe89129dbd565p-4,
    -0x1.da727838446a0p-4, -0x1.b6ac7c9ad5ad1p-4, -0x1.9335e4d594988p-4,
    -0x1.700d3deeac089p-4, -0x1.4d31165207eacp-4, -0x1.2aa0580471746p-4,
    -0x1.08599959e39a5p-4, -0x1.ccb7265ddb24dp-5, -0x1.894a8349fb262p-5,
    -0x1.466ad942de386p-5, -0x1.0415c89e74404p-5, -0x1.8492858c8c979p-6,
    -0x1.0205a38935667p-6, -0x1.01014f588de6dp-7, 0x0.0p+0,
    0x1.fe02b6b106791p-8, 0x1.fc0a890fc03e4p-7, 0x1.7b91acfd5b11cp-6,
    0x1.f82990e783380p-6, 0x1.39e86e1febd8dp-5, 0x1.77459be32dd23p-5,
    0x1.b42de091971d5p-5, 0x1.f0a30a01162a7p-5, 0x1.1653710a37ae3p-4,
    0x1.341d7461bd1ddp-4, 0x1.51b06dd061852p-4, 0x1.6f0d272e56b4dp-4,
    0x1.8c3465e319b45p-4, 0x1.a926d8a4ad570p-4, 0x1.c5e54bf5bc748p-4,
    0x1.e27074e2af2e8p-4, 0x1.fec9141dbeabbp-4, 0x1.0d77e8cd08e5ap-3,
    0x1.1b72b012f67a8p-3, 0x1.29552c41ff52ep-3, 0x1.371fc161e8f75p-3,
    0x1.44d2b38cb7d29p-3, 0x1.526e5e5a1b438p-3, 0x1.5ff3060a793d5p-3,
    0x1.6d60fce19d21fp-3, 0x1.7ab890410d909p-3, 0x1.87fa08620c915p-3,
    0x1.9525a80f456b8p-3, 0x1.a23bbffe2b567p-3, 0x1.af3c91880bffep-3,
    0x1.bc286be2d8cecp-3, 0x1.c8ff7a79a9a26p-3, 0x1.d5c21434fbb98p-3,
    0x1.e27075e2af2e7p-3, 0x1.ef0adfddc5940p-3, 0x1.fb918bd5e3e44p-3,
    0x1.04025b6b4d04ap-2, 0x1.0a3250a7390f0p-2, 0x1.1058bd1ae4ae2p-2,
    0x1.1675c97aba611p-2, 0x1.1c898b36999fdp-2, 0x1.22941e6cf7969p-2,
    0x1.2895a0bde86a4p-2, 0x1.2e8e2bee11d31p-2, 0x1.347ddb2987d59p-2,
    0x1.3a64c596945eap-2, 0x1.404309206a7e5p-2, 0x1.4618ba21c5ecap-2,
    0x1.4be5f937778a1p-2, 0x1.51aad7c2df82ep-2, 0x1.5767736c55a74p-2,
    0x1.5d1bda55809d0p-2, 0x1.62c82c939c7a3p-2, 0x1.686c8039b14b4p-2,
    0x1.6e08ec7aba1eap-2, 0x1.739d7e2bbd00ap-2, 0x1.792a545dd47a8p-2,
    0x1.7eaf83c82afc3p-2, 0x1.842d1c51e8b1bp-2, 0x1.89a33a8c14264p-2,
    0x1.8f11ea7b662d0p-2, 0x1.94793ee21170bp-2, 0x1.99d957617e08cp-2,
    0x1.9f323ccbf9854p-2,
};

// k = round(x 64/ln2) and r = x - k ln2/64 for |x| <= 160.  The integer k is read from the low
// word of x 64/ln2 + 1.5 2^52; r is formed by one fused multiply-add with ln2/64 rounded to
// double, whose error (< 2^-60 per unit of k, |k| < 2^14) stays below 2^-46.  A NaN x gives a
// NaN r and an arbitrary k.
SPM_FN double spm_reduce(double x, int* k) {
    const double t = __builtin_fma(x, 0x1.71547652b82fep+6, 0x1.8p52);
    *k = (int)(unsigned)__builtin_bit_cast(unsigned long long, t);
    return __builtin_fma(t - 0x1.8p52, -0x1.62e42fefa39efp-7, x);
}

// 2^(k/64) for |k| < 2^14 (the power of two is built in the exponent field)
SPM_FN double spm_pow2_64(int k) {
    const unsigned long long e = (unsigned long long)(unsigned)((k >> 6) + 1023) << 52;
    return spm_exp2_64[k & 63] * __builtin_bit_cast(double, e);
}

// exp(r) - 1 for |r| <= ln2/128: degree 4 (relative error r^4/120 < 2^-37, below 2^-44 once
// added to 1) and degree 5 (relative error r^5/720 < 2^-47, for expm1 itself)
SPM_FN double spm_expm1_poly(double r) {
    const double q = __builtin_fma(__builtin_fma(r, 0x1.5555555555555p-5, 0x1.5555555555555p-3), r, 0.5);
    return __builtin_fma(q * r, r, r);
}
SPM_FN double spm_expm1_poly5(double r) {
    double q = __builtin_fma(r, 0x1.1111111111111p-7, 0x1.5555555555555p-5);
    q = __builtin_fma(__builtin_fma(q, r, 0x1.5555555555555p-3), r, 0.5);
    return __builtin_fma(q * r, r, r);
}

SPM_FN float spm_expf(float xf) {
    double x = (double)xf;
    x = x > 100.0 ? 100.0 : (x < -160.0 ? -160.0 : x);  // beyond these the FP32 result is inf / 0
    int k;
    const double r = spm_reduce(x, &k);
    const double s = spm_pow2_64(k);
    return (float)__builtin_fma(s, spm_expm1_poly(r), s);
}

SPM_FN float spm_tanhf(float xf) {
    if (xf != xf) return xf;
    const double a = __builtin_fabs((double)xf);
    int k;
    const double r = spm_reduce(a > 10.0 ? 20.0 : a + a, &k);  // tanh rounds to 1 beyond 9.02
    const double s = spm_pow2_64(k);
    const double E = __builtin_fma(s, spm_expm1_poly5(r), s - 1.0);
    const double d = E + 2.0;
    const double y0 = (double)spm_rcp((float)d);
    const double y = __builtin_fma(y0, __builtin_fma(-d, y0, 1.0), y0);
    const double t = E * y;
    return (float)__builtin_copysign(t, (double)xf);
}

SPM_FN float spm_logf(float xf) {
    if (!(xf > 0.0f) || xf == __builtin_inff()) {
        if (xf == 0.0f) return -__builtin_inff();
        return xf == __builtin_inff() ? xf : __builtin_nanf("");
    }
    unsigned u = __builtin_bit_cast(unsigned, xf);
    int e = 0;
    if (u < 0x00800000u) {  // subnormal: scale into the normal range
        u = __builtin_bit_cast(unsigned, xf * 0x1p23f);
        e = -23;
    }
    // exponent relative to [0.75, 1.5): m = x 2^-e' with the field offset by 0.75's bits
    const int sh = (int)(u - 0x3f400000u) >> 23;
    e += sh;
    const float m = __builtin_bit_cast(float, u - ((unsigned)sh << 23));
    const int j = (int)__builtin_rintf((m - 1.0f) * 128.0f);
    const double r = __builtin_fma((double)m, (double)spm_log_c[j + 32], -1.0);
    double p = __builtin_fma(r, -0x1.5555555555555p-3, 0x1.999999999999ap-3);
    p = __builtin_fma(p, r, -0.25);
    p = __builtin_fma(p, r, 0x1.5555555555555p-2);
    p = __builtin_fma(p, r, -0.5);
    p = __builtin_fma(p * r, r, r);
    return (float)(__builtin_fma((double)e, 0x1.62e42fefa39efp-1, spm_log_t[j + 32]) + p);
}

// Host-side slab generator for the out-of-core path: planes [z0, z0 + nz) of the seeded
// Fourier field, written straight into (pinned) host memory by all host cores.
//
// f[i,j,k] = sum_m a_m cos(2 pi (k_m0 i/n0 + k_m1 j/n1 + k_m2 k/n2) + phi_m) + noise.
// The cosine is evaluated separably in FP64: per row (i, j) and mode m the complex
// coefficient c_m = a_m exp(i (2 pi (k_m0 i/n0 + k_m1 j/n1) + phi_m)) is formed once; along the
// row Re(c_m exp(i 2 pi k_m2 k/n2)) = Re(c_m) C_m[k] - Im(c_m) S_m[k] with tables C, S, i.e.
// two FMAs per voxel and mode.  The noise is the counter-based SplitMix64 term of the device
// generator.  Modes that share the x wavenumber k_m2 share one table row and one term per voxel.
// Results agree with the direct formula to FP64 rounding, i.e. the FP32 values match the device
// generator except, rarely, by one ulp.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

namespace {

inline uint64_t splitmix_at(uint64_t seed, uint64_t lin) {
    uint64_t z = seed + (lin + 1ull) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// The row loop vectorises over k once 64-bit multiplies and u64 -> double conversions are
// vector instructions (AVX-512 DQ); AVX2 and plain builds keep the scalar noise term.  A call
// covers columns [0, len) of one row segment; the tables C, S are offset to the segment's first
// column and keep their row stride ts (one row per distinct x wavenumber).  D, the number of
// distinct x wavenumbers, is a template argument so that the term loop unrolls.
#define SP_ROW_BODY                                                                                  \
    for (int64_t k = 0; k < len; ++k) {                                                              \
        double acc = 0.0;                                                                            \
        for (int m = 0; m < D; ++m) acc += cr[m] * C[m * ts + k] - ci[m] * S[m * ts + k];           \
        const double u = (double)(splitmix_at(seed, lin0 + (uint64_t)k) >> 11) * (1.0 / 9007199254740992.0); \
        out[k] = (float)(acc + (2.0 * u - 1.0) * 1e-3);                                              \
    }

using RowFn = void (*)(float*, int64_t, int64_t, const double*, const double*, const double*,
                       const double*, uint64_t, uint64_t);

template <int D>
__attribute__((target("avx512f,avx512dq,avx512vl,avx512bw,fma")))
void row_avx512dq(float* __restrict__ out, int64_t len, int64_t ts, const double* __restrict__ cr,
                  const double* __restrict__ ci, const double* __restrict__ C,
                  const double* __restrict__ S, uint64_t seed, uint64_t lin0) {
    SP_ROW_BODY
}

template <int D>
__attribute__((target("avx2,fma")))
void row_avx2(float* __restrict__ out, int64_t len, int64_t ts, const double* __restrict__ cr,
              const double* __restrict__ ci, const double* __restrict__ C,
              const double* __restrict__ S, uint64_t seed, uint64_t lin0) {
    SP_ROW_BODY
}

template <int D>
void row_plain(float* __restrict__ out, int64_t len, int64_t ts, const double* __restrict__ cr,
               const double* __restrict__ ci, const double* __restrict__ C,
               const double* __restrict__ S, uint64_t seed, uint64_t lin0) {
    SP_ROW_BODY
}

template <template <int> class Pick>
RowFn by_terms(int d) {
    switch (d) {
        case 1: return Pick<1>::fn;
        case 2: return Pick<2>::fn;
        case 3: return Pick<3>::fn;
        case 4: return Pick<4>::fn;
        case 5: return Pick<5>::fn;
        case 6: return Pick<6>::fn;
        case 7: return Pick<7>::fn;
        default: return Pick<8>::fn;
    }
}
template <int D> struct PickAvx512 { static constexpr RowFn fn = row_avx512dq<D>; };
template <int D> struct PickAvx2 { static constexpr RowFn fn = row_avx2<D>; };
template <int D> struct PickPlain { static constexpr RowFn fn = row_plain<D>; };

RowFn pick_row(int d) {
    __builtin_cpu_init();
    if (__builtin_cpu_supports("avx512dq") && __builtin_cpu_supports("avx512vl")) return by_terms<PickAvx512>(d);
    if (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma")) return by_terms<PickAvx2>(d);
    return by_terms<PickPlain>(d);
}

// Rows are taken in blocks of ROW_BLOCK and columns in segments of COL_SEG: the tables of one
// segment (8 modes x 2 x COL_SEG doubles = 32 KiB) stay in the core's L1 while every row of the
// block runs over it, instead of streaming both whole tables from L2 once per row.
constexpr int64_t ROW_BLOCK = 16;
constexpr int64_t COL_SEG = 256;

}  // namespace

extern "C" int sp_host_gen_fourier(float* out, const int64_t n[3], int64_t z0, int64_t nz,
                                   const int32_t* kk, const double* a, const double* phi,
                                   uint64_t seed, int threads) {
    if (!out || !n || !kk || !a || !phi || nz < 0 || z0 < 0 || z0 + nz > n[0]) return 1;
    const int64_t n0 = n[0], n1 = n[1], n2 = n[2];
    const double two_pi = 6.283185307179586476925286766559;
    // Modes that share an x wavenumber share their row tables: the row coefficients of such
    // modes are summed once per row (d <= 8 terms per voxel instead of 8).
    int wave[8], term[8], d = 0;
    for (int m = 0; m < 8; ++m) {
        int q = 0;
        while (q < d && wave[q] != kk[3 * m + 2]) ++q;
        if (q == d) wave[d++] = kk[3 * m + 2];
        term[m] = q;
    }
    std::vector<double> C((size_t)d * n2), S((size_t)d * n2);
    for (int q = 0; q < d; ++q)
        for (int64_t k = 0; k < n2; ++k) {
            const double t = two_pi * (double)wave[q] * (double)k / (double)n2;
            C[q * n2 + k] = std::cos(t);
            S[q * n2 + k] = std::sin(t);
        }
    const int64_t rows = nz * n1;
    const int64_t blocks = (rows + ROW_BLOCK - 1) / ROW_BLOCK;
    const RowFn row = pick_row(d);
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < blocks; ++b) {
        const int64_t r0 = b * ROW_BLOCK, r1 = r0 + ROW_BLOCK < rows ? r0 + ROW_BLOCK : rows;
        double cr[ROW_BLOCK][8], ci[ROW_BLOCK][8];
        for (int64_t r = r0; r < r1; ++r) {
            const int64_t i = z0 + r / n1, j = r % n1;
            for (int q = 0; q < d; ++q) cr[r - r0][q] = ci[r - r0][q] = 0.0;
            for (int m = 0; m < 8; ++m) {
                const double t = two_pi * ((double)kk[3 * m] * (double)i / (double)n0 +
                                           (double)kk[3 * m + 1] * (double)j / (double)n1) + phi[m];
                cr[r - r0][term[m]] += a[m] * std::cos(t);
                ci[r - r0][term[m]] += a[m] * std::sin(t);
            }
        }
        for (int64_t k0 = 0; k0 < n2; k0 += COL_SEG) {
            const int64_t len = k0 + COL_SEG < n2 ? COL_SEG : n2 - k0;
            for (int64_t r = r0; r < r1; ++r) {
                const int64_t i = z0 + r / n1, j = r % n1;
                row(out + r * n2 + k0, len, n2, cr[r - r0], ci[r - r0], C.data() + k0, S.data() + k0,
                    seed, (uint64_t)((i * n1 + j) * n2 + k0));
            }
        }
    }
    return 0;
}

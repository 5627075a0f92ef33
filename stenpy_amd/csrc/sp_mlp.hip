// Kernels of the schedule model: a small MLP trained on the autotuner's timings.
//
// The only dense contraction in the backend, so the only MFMA user: C = alpha * op(A) op(B)
// with bf16 operands and fp32 accumulation on v_mfma_f32_16x16x16_bf16.  A block of 256
// threads (4 wave64) computes a 64x64 tile of C; each wave owns a 32x32 quadrant made of
// 2x2 MFMA 16x16 tiles; K advances 16 at a time through LDS.  The remaining kernels are the
// element-wise pieces of forward/backward/Adam.
#include "sp_common.h"

using namespace sp;

namespace {

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int BM = 64, BN = 64, BK = 16;

__device__ __forceinline__ uint16_t ld_bf16(const uint16_t* p, int64_t i) { return p[i]; }

__global__ void __launch_bounds__(256) k_gemm_bf16(int ta, int tb, int M, int N, int K,
                                                   const uint16_t* __restrict__ A, int lda,
                                                   const uint16_t* __restrict__ B, int ldb,
                                                   float* __restrict__ C, int ldc, float alpha,
                                                   int accumulate) {
    __shared__ uint16_t As[BM][BK + 4];
    __shared__ uint16_t Bs[BK][BN + 4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    f32x4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += BK) {
        // stage A(m, k) and B(k, n) tiles, zero padded outside the matrices
        for (int e = tid; e < BM * BK; e += 256) {
            const int r = e / BK, c = e % BK;
            const int m = m0 + r, k = k0 + c;
            uint16_t v = 0;
            if (m < M && k < K) v = ta ? ld_bf16(A, (int64_t)k * lda + m) : ld_bf16(A, (int64_t)m * lda + k);
            As[r][c] = v;
        }
        for (int e = tid; e < BK * BN; e += 256) {
            const int r = e / BN, c = e % BN;
            const int k = k0 + r, n = n0 + c;
            uint16_t v = 0;
            if (k < K && n < N) v = tb ? ld_bf16(B, (int64_t)n * ldb + k) : ld_bf16(B, (int64_t)k * ldb + n);
            Bs[r][c] = v;
        }
        __syncthreads();
        // lane l holds A[row = l % 16][k = 4 (l / 16) + 0..3] and B[k = 4 (l / 16) + 0..3][col = l % 16]
        const int fr = lane & 15, fk = (lane >> 4) * 4;
        bf16x4 af[2], bfr[2];
        for (int i = 0; i < 2; ++i) {
            for (int t = 0; t < 4; ++t) af[i][t] = (short)As[wr + 16 * i + fr][fk + t];
            for (int t = 0; t < 4; ++t) bfr[i][t] = (short)Bs[fk + t][wc + 16 * i + fr];
        }
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i], bfr[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    // lane l holds C[row = 4 (l / 16) + t][col = l % 16], t = 0..3
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int t = 0; t < 4; ++t) {
                const int m = m0 + wr + 16 * i + (lane >> 4) * 4 + t;
                const int n = n0 + wc + 16 * j + (lane & 15);
                if (m < M && n < N) {
                    float v = alpha * acc[i][j][t];
                    float* dst = C + (int64_t)m * ldc + n;
                    *dst = accumulate ? *dst + v : v;
                }
            }
}

// z = x + bias (per column); out_f32 = z; out_bf16 = relu(z) (or z when relu == 0)
__global__ void __launch_bounds__(256) k_bias_act(const float* __restrict__ x,
                                                  const float* __restrict__ bias, int rows,
                                                  int cols, int relu, float* __restrict__ z,
                                                  uint16_t* __restrict__ a) {
    const int64_t n = (int64_t)rows * cols;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float v = x[i] + bias[i % cols];
        if (z) z[i] = v;
        if (a) a[i] = f32_to_bf16(relu ? fmaxf(v, 0.f) : v);
    }
}

// g_out = g_in * (z > 0) as fp32 and bf16 (ReLU backward)
__global__ void __launch_bounds__(256) k_relu_grad(const float* __restrict__ g,
                                                   const float* __restrict__ z, int64_t n,
                                                   float* __restrict__ gf, uint16_t* __restrict__ gb) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float v = z[i] > 0.f ? g[i] : 0.f;
        if (gf) gf[i] = v;
        if (gb) gb[i] = f32_to_bf16(v);
    }
}

// out[c] = sum_r x[r, c] (bias gradient), one thread per column, fixed order
__global__ void __launch_bounds__(256) k_colsum(const float* __restrict__ x, int rows, int cols,
                                                float* __restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f, comp = 0.f;
    for (int r = 0; r < rows; ++r) {
        const float v = x[(int64_t)r * cols + c];
        const float t = s + v;
        comp += fabsf(s) >= fabsf(v) ? (s - t) + v : (v - t) + s;
        s = t;
    }
    out[c] = s + comp;
}

// loss = mean((y - t)^2); grad = 2 (y - t) / n (fp32 and bf16)
__global__ void __launch_bounds__(256) k_mse(const float* __restrict__ y, const float* __restrict__ t,
                                             int n, float* __restrict__ grad,
                                             uint16_t* __restrict__ gradb, float* __restrict__ loss) {
    __shared__ float part[256];
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += 256) {
        const float d = y[i] - t[i];
        s += d * d;
        const float g = 2.f * d / (float)n;
        if (grad) grad[i] = g;
        if (gradb) gradb[i] = f32_to_bf16(g);
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st >= 1; st >>= 1) {
        if (threadIdx.x < st) part[threadIdx.x] += part[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0 && loss) loss[0] = part[0] / (float)n;
}

// Adam step on fp32 master weights, refreshing the bf16 copy used by the GEMMs
__global__ void __launch_bounds__(256) k_adam(float* __restrict__ w, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, int64_t n,
                                              float lr, float b1, float b2, float eps, float c1,
                                              float c2, float wd, uint16_t* __restrict__ wb) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i] + wd * w[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wn = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wn;
        if (wb) wb[i] = f32_to_bf16(wn);
    }
}

__global__ void __launch_bounds__(256) k_cast_bf16(const float* __restrict__ x, int64_t n,
                                                   uint16_t* __restrict__ y) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = f32_to_bf16(x[i]);
}

unsigned blocks_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

}  // namespace

extern "C" int sp_gemm_bf16(int ta, int tb, int M, int N, int K, const void* A, int lda,
                            const void* B, int ldb, float* C, int ldc, float alpha,
                            int accumulate, void* stream) {
    SP_REQUIRE(M > 0 && N > 0 && K > 0 && A && B && C, "gemm: bad arguments");
    SP_REQUIRE(lda > 0 && ldb > 0 && ldc >= N, "gemm: bad leading dimensions");
    dim3 grid(cdiv(N, BN), cdiv(M, BM));
    hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, as_stream(stream), ta, tb, M, N, K,
                       (const uint16_t*)A, lda, (const uint16_t*)B, ldb, C, ldc, alpha, accumulate);
    return launch_status("sp_gemm_bf16");
}

extern "C" int sp_mlp_bias_act(const float* x, const float* bias, int rows, int cols, int relu,
                               float* z, void* a_bf16, void* stream) {
    SP_REQUIRE(x && bias && rows > 0 && cols > 0, "bias_act: bad arguments");
    hipLaunchKernelGGL(k_bias_act, dim3(blocks_for((int64_t)rows * cols)), dim3(256), 0,
                       as_stream(stream), x, bias, rows, cols, relu, z, (uint16_t*)a_bf16);
    return launch_status("sp_mlp_bias_act");
}

extern "C" int sp_mlp_relu_grad(const float* g, const float* z, int64_t n, float* g_out,
                                void* g_out_bf16, void* stream) {
    SP_REQUIRE(g && z && n > 0, "relu_grad: bad arguments");
    hipLaunchKernelGGL(k_relu_grad, dim3(blocks_for(n)), dim3(256), 0, as_stream(stream), g, z, n,
                       g_out, (uint16_t*)g_out_bf16);
    return launch_status("sp_mlp_relu_grad");
}

extern "C" int sp_mlp_colsum(const float* x, int rows, int cols, float* out, void* stream) {
    SP_REQUIRE(x && out && rows > 0 && cols > 0, "colsum: bad arguments");
    hipLaunchKernelGGL(k_colsum, dim3(cdiv(cols, 256)), dim3(256), 0, as_stream(stream), x, rows,
                       cols, out);
    return launch_status("sp_mlp_colsum");
}

extern "C" int sp_mlp_mse(const float* y, const float* t, int n, float* grad, void* grad_bf16,
                          float* loss, void* stream) {
    SP_REQUIRE(y && t && n > 0, "mse: bad arguments");
    hipLaunchKernelGGL(k_mse, dim3(1), dim3(256), 0, as_stream(stream), y, t, n, grad,
                       (uint16_t*)grad_bf16, loss);
    return launch_status("sp_mlp_mse");
}

extern "C" int sp_mlp_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr,
                           float beta1, float beta2, float eps, int step, float weight_decay,
                           void* w_bf16, void* stream) {
    SP_REQUIRE(w && g && m && v && n > 0 && step > 0, "adam: bad arguments");
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(n)), dim3(256), 0, as_stream(stream), w, g, m, v, n,
                       lr, beta1, beta2, eps, c1, c2, weight_decay, (uint16_t*)w_bf16);
    return launch_status("sp_mlp_adam");
}

extern "C" int sp_cast_f32_bf16(const float* x, int64_t n, void* y, void* stream) {
    SP_REQUIRE(x && y && n >= 0, "cast: bad arguments");
    if (n == 0) return SP_OK;
    hipLaunchKernelGGL(k_cast_bf16, dim3(blocks_for(n)), dim3(256), 0, as_stream(stream), x, n,
                       (uint16_t*)y);
    return launch_status("sp_cast_f32_bf16");
}

/*
 * stenpy_hip.h -- C ABI of libstenpy_hip.so, the MI355X (gfx950) execution backend of Stenpy.
 *
 * The reference's plugin boundary is Python-level: ``register_operator(name, func, ...)``
 * stores a callable that the eager executor invokes as ``func(*tensors, **params)``
 * (stenpy_engine.py:1184-1212, called at 1074-1081).  This library is what those callables
 * reach through ctypes: every entry point takes plain pointers, sizes and enums (no torch
 * types), works on caller-owned device memory, is ordered on the caller's HIP stream, never
 * throws or aborts across the ABI, and returns an ``int`` status (SP_OK = 0).  The message of
 * the last failure on the calling thread is ``sp_last_error()``.
 *
 * Layout contract: fields are C-order (n0, n1, n2) with n2 contiguous; vector and tensor
 * fields are channel-last, (n0, n1, n2, C) and (n0, n1, n2, 3, 3), exactly as the reference
 * stores them (stenpy_engine.py:1861, 2127).  Storage is FP32 (BF16 allowed for outputs);
 * reductions return FP64.
 */
#ifndef STENPY_HIP_H
#define STENPY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SP_ABI_VERSION 5

/* status codes */
enum { SP_OK = 0, SP_EINVAL = 1, SP_EHIP = 2, SP_EUNSUPPORTED = 3, SP_ECOMPILE = 4 };

/* element types; SP_C64 is interleaved (re, im) FP32 pairs, torch.complex64 */
enum { SP_F32 = 0, SP_BF16 = 1, SP_F64 = 2, SP_C64 = 3 };

/* boundary handling of padded stencils (stenpy_engine.py:1374-1391): neumann -> replicate,
 * dirichlet -> zeros, periodic -> circular, reflect -> mirror; any other name pads zeros. */
enum { SP_PERIODIC = 0, SP_NEUMANN = 1, SP_DIRICHLET = 2, SP_REFLECT = 3, SP_ZEROPAD = 4 };

#define SP_MAX_DIMS 6

/* A dense tensor view.  ``stride`` is in elements; a stride of 0 broadcasts. */
typedef struct sp_tensor {
    void*   ptr;
    int32_t dtype;
    int32_t ndim;
    int64_t shape[SP_MAX_DIMS];
    int64_t stride[SP_MAX_DIMS];
} sp_tensor;

/* ---- library -------------------------------------------------------------------------- */
int         sp_abi_version(void);
const char* sp_last_error(void);
/* Name and compute-unit count of the current device. */
int         sp_device_info(char* name, int name_len, int* n_cu, int* n_xcd);

/* ---- point-wise (stenpy_engine.py:1218-1367) ------------------------------------------- */
/* unary ops */
enum { SP_NEG = 0, SP_EXP = 1, SP_LOG = 2, SP_SQRT = 3, SP_SIN = 4, SP_COS = 5, SP_TANH = 6,
       SP_CLAMP = 7, SP_SCALE = 8, SP_COPY = 9, SP_SQUARE = 10 };
/* binary ops */
enum { SP_ADD = 0, SP_SUB = 1, SP_MUL = 2, SP_DIV = 3 };

/* out = op(a; p0, p1).  LOG/SQRT clamp their argument below at p0 (eps), CLAMP uses
 * [p0, p1], SCALE multiplies by p0.  ``out`` is contiguous, a is broadcast to it. */
int sp_op_unary(int op, const sp_tensor* a, sp_tensor* out, double p0, double p1, void* stream);
/* out = a op b with broadcasting; DIV replaces |b| < p0 by sign-preserving +-p0 (b = 0 -> +p0),
 * stenpy_engine.py:1263-1279. */
int sp_op_binary(int op, const sp_tensor* a, const sp_tensor* b, sp_tensor* out, double p0,
                 void* stream);

/* Scalar programs: n <= SP_SCALAR_PROGRAM_MAX point-wise ops on single FP64 values, evaluated
 * in order by one launch (an op may read an earlier op's result), in FP64.  Each op is a unary
 * (binary = 0: SP_NEG .. SP_SQUARE, p0 / p1 as for sp_op_unary) or binary op (binary = 1:
 * SP_ADD .. SP_DIV, DIV's eps in p0) on device doubles ``a`` / ``b``; a NULL pointer takes the
 * constant av / bv instead.  The descriptors are copied into the launch's arguments, so the
 * launch can be captured.  Replaces one launch per op of the reference's scalar arithmetic on
 * reduction results (the point-wise ops of stenpy_engine.py:1218-1367 on 0-d tensors, run
 * between the levels of a fused or batched graph). */
typedef struct sp_scalar_op {
    const double* a;
    const double* b;
    double* out;
    double av, bv, p0, p1;
    int32_t binary;
    int32_t op;
} sp_scalar_op;
#define SP_SCALAR_PROGRAM_MAX 48
int sp_scalar_program(const sp_scalar_op* ops, int n, void* stream);

/* ---- channel-last helpers (stenpy_engine.py:2247-2313) --------------------------------- */
/* out[..., :] = sum_c a[..., c] * b[..., c] folded left (extension op ``dot``). */
int sp_op_dot(const sp_tensor* a, const sp_tensor* b, sp_tensor* out, void* stream);
/* out[...] = sum_c a[..., c]^2 folded left (extension op ``sqnorm``). */
int sp_op_sqnorm(const sp_tensor* a, sp_tensor* out, void* stream);
/* out = max(|a|_2 over the last axis, eps), keepdim (``norm_last``). */
int sp_op_norm_last(const sp_tensor* a, sp_tensor* out, double eps, void* stream);
/* out[..., c] = a[..., c] / max(|a|_2, eps) (normalisation used by mean_curvature). */
int sp_op_normalize_last(const sp_tensor* a, sp_tensor* out, double eps, void* stream);

/* ---- stencils ---------------------------------------------------------------------------
 * Axis pass: the building block for any rank.  The array is viewed as (outer, n, inner)
 * along the differentiated axis; ``x`` may select one channel of a channel-last field via
 * its strides.  Modes:
 *   SP_D1_PERIODIC  (x[i+1] - x[i-1]) / (2h) with wrap           (torch.roll path, 1829-1834)
 *   SP_D1_EDGE2     torch.gradient, 2nd-order one-sided edges      (torch.gradient edge_order=2)
 *   SP_D1_EDGE1     torch.gradient, 1st-order one-sided edges      (edge_order=1)
 *   SP_D2_PAD       (x[i+1] + x[i-1] - 2x) / h^2, ``pad`` = boundary enum (1501-1511)
 * out = scale * value, or out += scale * value when ``accumulate``. */
enum { SP_D1_PERIODIC = 0, SP_D1_EDGE2 = 1, SP_D1_EDGE1 = 2, SP_D2_PAD = 3 };
int sp_axis_pass(const sp_tensor* x, int axis, sp_tensor* out, int mode, int pad, double h,
                 double scale, int accumulate, void* stream);
/* Separable filter pass: out[i] = sum_k w[k] * x[i + k - ntaps/2] along ``axis``, reads past
 * the ends padded by ``pad`` (boundary enum, as the reference's _pad1d).  ``weights``: host
 * array of ``ntaps`` (odd, <= SP_FIR_MAX_TAPS).  Building block of moving_average (box taps,
 * stenpy_engine.py:2385-2408) and conv3d_spectral ((1, 2, 1) / 4 periodic taps: the
 * reference's FFT convolution with the 27-point Gaussian, 2800-2846). */
#define SP_FIR_MAX_TAPS 31
int sp_axis_fir(const sp_tensor* x, int axis, sp_tensor* out, int ntaps, const double* weights,
                int pad, void* stream);

/* ---- spectral (stenpy_engine.py:2871-3012) -------------------------------------------- */
/* torch.fft ``norm``: "backward" scales the inverse by 1/n, "forward" the forward transform,
 * "ortho" both by 1/sqrt(n). */
enum { SP_NORM_BACKWARD = 0, SP_NORM_ORTHO = 1, SP_NORM_FORWARD = 2 };
/* FFT over every axis of a contiguous 1-D, 2-D or 3-D field, computed in FP64 (rocFFT) and
 * rounded once.  x: FP32 (real) or C64.  Forward: out C64 -- fft / fft_complex, 2875-2885,
 * 2901-2908.  Inverse: out C64 (ifft_complex, 2910-2918) or FP32, the real part (ifft,
 * 2887-2896). */
int sp_fft(const sp_tensor* x, sp_tensor* out, int inverse, int norm, void* stream);
/* out = Re(ifftn(m(k) * fftn(x))) in FP64, rounded once, with k = 2 pi fftfreq(n_d, dx[d])
 * (fftfreq in FP32, as torch's default dtype computes it) and m = i k_dim
 * (SP_SPECTRAL_GRADIENT, spectral_gradient, 2979-2992) or -sum_d k_d^2 (SP_SPECTRAL_LAPLACIAN,
 * spectral_laplacian, 2997-3010).  x, out: FP32, contiguous, 1-D to 3-D; dx: one spacing per
 * axis. */
enum { SP_SPECTRAL_GRADIENT = 0, SP_SPECTRAL_LAPLACIAN = 1 };
int sp_spectral(const sp_tensor* x, sp_tensor* out, const double* dx, int kind, int dim, void* stream);

/* ---- statistics and integrals (CPU branches of stenpy_engine.py) ------------------------ */
/* Every sum is accumulated in FP64 from FP32 inputs; global results are written to ``result``,
 * one double in device memory, with no host synchronisation. */
/* cumulative_integral, 2462-2483: out[k] = sum_{j<k} (x[j] + x[j+1]) * dx / 2 along ``dim``,
 * out[0] = 0; x, out FP32 of one shape. */
int sp_cumulative_trapezoid(const sp_tensor* x, sp_tensor* out, int dim, double dx, void* stream);
/* surface_integral, 2486-2496: sum_i w[i] * sum_c t[i, c] * normals[i, c]; normals (..., C),
 * area_weights (...), t (..., C) or (..., 1). */
int sp_surface_integral(const sp_tensor* t, const sp_tensor* normals, const sp_tensor* weights,
                        double* result, void* stream);
/* covariance (2575-2592): dot(a - mean a, b - mean b) / max(n - 1, 1);
 * correlation (2595-2615): dot(a', b') / (|a'| |b'| + 1e-15). */
enum { SP_COVARIANCE = 0, SP_CORRELATION = 1 };
int sp_pair_stat(int kind, const sp_tensor* a, const sp_tensor* b, double* result, void* stream);
/* entropy, 2638-2669: -sum p log(p + eps), p = |x| / (sum |x| + eps); over everything into
 * ``result`` (dim < 0), or along ``dim`` into ``out`` (FP32, the input's shape without dim). */
int sp_entropy(const sp_tensor* x, int dim, double eps, double* result, sp_tensor* out, void* stream);

/* ---- geometry (stenpy_engine.py:3115-3156) ------------------------------------------- */
/* Exact Euclidean distance transform with per-axis sampling dx[ndim]: the CPU branch,
 * scipy.ndimage.distance_transform_edt(t == 0, sampling=dx).  Voxels with t != 0 get 0, every
 * voxel with t == 0 the distance to the nearest voxel with t != 0 (+inf when there is none).
 * FP64 internally, FP32 out; x, out contiguous 1-D to 3-D. */
int sp_distance_transform(const sp_tensor* x, sp_tensor* out, const double* dx, void* stream);

/* 3-D single-pass kernels for scalar (n0,n1,n2) / vector (n0,n1,n2,3) fields.
 * Each replaces the CPU branch of the cited reference op (the parity target). */
/* laplacian, stenpy_engine.py:1478-1512 */
int sp_laplacian3d(const sp_tensor* x, sp_tensor* out, const double dx[3], int boundary,
                   void* stream);
/* gradient (dim=None), 1780-1865: periodic -> roll, else torch.gradient per axis */
int sp_gradient3d(const sp_tensor* x, sp_tensor* out, double dx, int boundary, void* stream);
/* divergence of (n0,n1,n2,3), 1870-1931 */
int sp_divergence3d(const sp_tensor* v, sp_tensor* out, const double dx[3], int boundary,
                    void* stream);
/* curl of (n0,n1,n2,3), boundary ignored (torch.gradient edge_order=2), 1936-2032 */
int sp_curl3d(const sp_tensor* v, sp_tensor* out, const double dx[3], void* stream);
/* gradient_nd of a 3-D field without singleton axes, 2037-2081 */
int sp_gradient_nd3d(const sp_tensor* x, sp_tensor* out, const double dx[3], void* stream);
/* hessian H[i][j] = D_j D_i f with torch.gradient edges, 2086-2128 */
int sp_hessian3d(const sp_tensor* x, sp_tensor* out, const double dx[3], void* stream);
/* jacobian J[c][d] = D_d v_c, 2133-2175 (CPU branch) */
int sp_jacobian3d(const sp_tensor* v, sp_tensor* out, const double dx[3], void* stream);
/* mean curvature div(g/max(|g|,eps)), g = gradient_nd f, 2180-2195 */
int sp_mean_curvature3d(const sp_tensor* x, sp_tensor* out, const double dx[3], int boundary,
                        double eps, void* stream);
/* material derivative sum_d v_d (gradient_nd f)_d, 2223-2243 */
int sp_material_derivative3d(const sp_tensor* x, const sp_tensor* v, sp_tensor* out,
                             const double dx[3], void* stream);

/* ---- reductions (stenpy_engine.py:2340-2572, 2436-2457) -------------------------------
 * Block partials are compensated FP32 sums (x, and exactly split x^2 for variance); one
 * final block combines them in FP64, deterministically.  ``result`` is device memory:
 *   SP_R_SUM, SP_R_MEAN, SP_R_NORM2, SP_R_VAR (unbiased), SP_R_VAR_BIASED: 1 double
 *   SP_R_MINMAX: 2 doubles;  SP_R_INTEGRATE: 1 double (Simpson/trapezoid weights per axis,
 *   ``wparams`` = {dx0, dx1, dx2, method}, method 0 = simpson, 1 = trapezoid; x is 3-D)
 *   SP_R_ANYSMALLDIAG: 1 double = 1.0 if any |H[..., i, i]| < 1e-12 else 0.0 (x is (...,3,3))
 * ``workspace`` must hold sp_reduce_workspace_bytes(numel) bytes of device memory. */
enum { SP_R_SUM = 0, SP_R_MEAN = 1, SP_R_NORM2 = 2, SP_R_VAR = 3, SP_R_VAR_BIASED = 4,
       SP_R_MINMAX = 5, SP_R_INTEGRATE = 6, SP_R_ANYSMALLDIAG = 7 };
size_t sp_reduce_workspace_bytes(int64_t numel);
/* Final FP64 combine of ``nrecords`` block records (4 doubles each: {s, c} for sums,
 * {s1, c1, s2, c2} for variance, {lo, hi} for min/max) written by the fused kernels; ``count``
 * is the number of reduced elements (used by mean and variance). */
int sp_reduce_combine(int kind, const double* records, int nrecords, int64_t count,
                      double* result, void* stream);
int sp_reduce(int kind, const sp_tensor* x, double* result, void* workspace,
              const double* wparams, void* stream);
/* sp_reduce without a workspace and without any allocation: the block records live in a
 * file-scope device array of sp_reduce_noalloc_scratch_bytes() = 64 KiB per GPU (2048 records,
 * the most the clamped stage-1 grid writes, so every input size fits; an input that needed
 * more would be rejected with SP_EINVAL).  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops).
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU race. */
int    sp_reduce_noalloc(int kind, const sp_tensor* x, double* result, const double* wparams,
                         void* stream);
size_t sp_reduce_noalloc_scratch_bytes(void);
/* Many final combines in one launch (one block each): ``descs`` is a device array of n
 * descriptors, each as for sp_reduce_combine.  Used after a batched launch of many grids.
 * No counterpart in the reference (it reduces one tensor per call, engine 2340-2382). */
typedef struct sp_combine_desc {
    const double* records;
    double* result;
    int64_t count;
    int32_t kind;
    int32_t nrecords;
} sp_combine_desc;
int sp_reduce_combine_batch(const sp_combine_desc* descs, int n, void* stream);
/* The same for at most SP_COMBINE_MULTI_MAX descriptors given in host memory (copied into the
 * launch's arguments): one launch for all reductions of one fused group. */
#define SP_COMBINE_MULTI_MAX 16
int sp_reduce_combine_multi(const sp_combine_desc* descs, int n, void* stream);

/* ---- per-voxel symmetric 3x3 tensor ops (stenpy_engine.py:2674-2730, 2852-2867) -------
 *   SP_T_TRACE -> (...), SP_T_DET -> (...) [``flag`` = device double from SP_R_ANYSMALLDIAG:
 *   when nonzero, det(H + 1e-12 I)], SP_T_EIG -> (..., 3) ascending (lower triangle, as
 *   eigvalsh), SP_T_DEV -> (..., 3, 3), SP_T_INV -> (..., 3, 3). */
enum { SP_T_TRACE = 0, SP_T_DET = 1, SP_T_EIG = 2, SP_T_DEV = 3, SP_T_INV = 4 };
int sp_tensor3(int kind, const sp_tensor* h, sp_tensor* out, const double* flag, void* stream);

/* ---- synthetic input ---------------------------------------------------------------------
 * Planes [z0, z0 + out->shape[0]) of the seeded Fourier field of full shape n (FP64 formula,
 * FP32 store); k: 8x3 wavenumbers, a: 8 amplitudes, phi: 8 phases (host arrays). */
int sp_gen_fourier(sp_tensor* out, const int64_t n[3], int64_t z0, const int32_t* k,
                   const double* a, const double* phi, uint64_t noise_seed, void* stream);

/* Host (CPU, OpenMP) twin of sp_gen_fourier for out-of-core slabs: writes planes
 * [z0, z0 + nz) into host memory ``out`` (pinned for fast copies) with ``threads`` threads
 * (0 = all).  Separable FP64 evaluation; FP32 values agree with the device generator up to a
 * rare one-ulp difference. */
int sp_host_gen_fourier(float* out, const int64_t n[3], int64_t z0, int64_t nz,
                        const int32_t* k, const double* a, const double* phi, uint64_t noise_seed,
                        int threads);

/* ---- run-time compiled kernels (fused plans) ----------------------------------------------
 * The fused executor generates CDNA4 HIP source per graph and compiles it with hipRTC. */
typedef struct sp_module sp_module;
/* Compile ``source`` for the current device; options are '\n'-separated.  On success
 * *code/*code_size hold a malloc'ed code object (free with sp_free_code). */
int  sp_rtc_compile(const char* source, const char* name, const char* options, void** code,
                    size_t* code_size, char* log, size_t log_len);
void sp_free_code(void* code);
int  sp_module_load(const void* code, size_t code_size, const char* kernel_name,
                    sp_module** mod);
/* Launch the module's kernel with one by-value argument block of ``args_size`` bytes. */
int  sp_module_launch(sp_module* mod, const unsigned grid[3], const unsigned block[3],
                      unsigned shared_bytes, void* stream, const void* args, size_t args_size);
/* registers, LDS bytes and spill bytes of the loaded kernel */
int  sp_module_attributes(sp_module* mod, int* vgprs, int* lds_bytes, int* spill_bytes,
                          int* max_threads);
/* Blocks of ``block_threads`` threads with ``dyn_shared`` bytes of dynamic LDS that fit on
 * one CU at once (sizes the grid of a persistent batch launch). */
int  sp_module_occupancy(sp_module* mod, int block_threads, size_t dyn_shared, int* blocks_per_cu);
void sp_module_unload(sp_module* mod);

/* ---- schedule model (MLP trained on autotuner timings) --------------------------------
 * The one dense contraction of the backend, on MFMA: C[M,N] (+)= alpha * op(A) op(B) with
 * bf16 row-major A (M x K, or K x M when ta) and B (K x N, or N x K when tb), fp32 C.
 * No counterpart in the reference (its cost model is static, fused_compiler.py:20-43). */
int sp_gemm_bf16(int ta, int tb, int M, int N, int K, const void* A, int lda, const void* B,
                 int ldb, float* C, int ldc, float alpha, int accumulate, void* stream);
/* z = x + bias[col] (fp32, optional) and a = relu(z) or z as bf16 (optional) */
int sp_mlp_bias_act(const float* x, const float* bias, int rows, int cols, int relu, float* z,
                    void* a_bf16, void* stream);
/* ReLU backward: g_out = g * (z > 0) as fp32 and/or bf16 */
int sp_mlp_relu_grad(const float* g, const float* z, int64_t n, float* g_out, void* g_out_bf16,
                     void* stream);
/* out[c] = sum_r x[r, c] (compensated, fixed order) */
int sp_mlp_colsum(const float* x, int rows, int cols, float* out, void* stream);
/* loss = mean((y - t)^2) (device scalar), grad = 2 (y - t) / n (fp32 and/or bf16) */
int sp_mlp_mse(const float* y, const float* t, int n, float* grad, void* grad_bf16, float* loss,
               void* stream);
/* Adam step (bias-corrected, decoupled-free weight decay added to the gradient) on fp32
 * master weights; refreshes the bf16 copy when given */
int sp_mlp_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float beta1,
                float beta2, float eps, int step, float weight_decay, void* w_bf16, void* stream);
int sp_cast_f32_bf16(const float* x, int64_t n, void* y, void* stream);

/* ---- measurement probes (no reference counterpart) ------------------------------------- */
/* HBM streaming probes that measure what this GPU sustains for a read:write mix, reported
 * beside the 8 TB/s roofline.  COPY: out = a (1:1); 4R1W: reads a (n floats) and u (3n floats,
 * as three coalesced thirds), writes out (n floats), the stencil graph's 16 B read / 4 B written
 * per voxel; 4R1W_AOS: the same mix with u channel-last (3 floats per voxel, the graph's layout);
 * READ: one double per block (sum of its elements) into ``partials``; WRITE: fills ``a`` (n
 * floats) with values that depend on the index, writes only.  ``persistent`` = 0
 * launches one float4 group per thread; 1 launches 8 blocks per CU that stride over the range.
 * n must be a multiple of 4 and every pointer 16-byte aligned. */
enum { SP_PROBE_COPY = 0, SP_PROBE_4R1W = 1, SP_PROBE_READ = 2, SP_PROBE_WRITE = 3,
       SP_PROBE_4R1W_AOS = 4 };
int     sp_probe_stream(int mix, int persistent, const float* a, const float* u, float* out,
                        double* partials, int64_t n, int n_cu, void* stream);
/* Blocks (and so ``partials`` entries) that sp_probe_stream launches for these arguments. */
int64_t sp_probe_blocks(int persistent, int64_t n, int n_cu);

#ifdef __cplusplus
}
#endif
#endif /* STENPY_HIP_H */

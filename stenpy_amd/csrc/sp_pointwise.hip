// Point-wise operators (stenpy_engine.py:1218-1367) and channel-last helpers
// (stenpy_engine.py:2247-2313, plus the ``dot``/``sqnorm`` extensions).
//
// Arithmetic is FP32; outputs may be stored as BF16 (round to nearest even).  Same-shape
// contiguous operands take a 4-wide vector path; anything else (scalar or general
// broadcasting) takes the strided path.
#include "sp_common.h"
#include "sp_math32.h"

using namespace sp;

namespace {

struct Bcast {
    int nd;
    int64_t shape[SP_MAX_DIMS];
    int64_t sa[SP_MAX_DIMS];
    int64_t sb[SP_MAX_DIMS];
};

__device__ __forceinline__ float unary_f(int op, float a, float p0, float p1) {
    switch (op) {
        case SP_NEG: return -a;
        // transcendentals: an FP64 evaluation rounded once to FP32, 0.5 ulp + 2^-16 (the FP32
        // library versions are 1-2 ulp; the reference's CPU path is near correctly rounded)
        case SP_EXP: return spm_expf(a);
        case SP_LOG: return spm_logf(a < p0 ? p0 : a);
        case SP_SQRT: return sqrtf(a < p0 ? p0 : a);
        case SP_SIN: return spm_sinf(a);
        case SP_COS: return spm_cosf(a);
        case SP_TANH: return spm_tanhf(a);
        case SP_CLAMP: return a < p0 ? p0 : (a > p1 ? p1 : a);
        case SP_SCALE: return a * p0;
        case SP_SQUARE: return a * a;
        default: return a;
    }
}

__device__ __forceinline__ float binary_f(int op, float a, float b, float eps) {
    switch (op) {
        case SP_ADD: return a + b;
        case SP_SUB: return a - b;
        case SP_MUL: return a * b;
        default: {
            float bs = fabsf(b) < eps ? (b < 0.f ? -eps : eps) : b;
            return a / bs;
        }
    }
}

__device__ __forceinline__ void offsets(const Bcast& bc, int64_t i, int64_t& oa, int64_t& ob) {
    oa = 0;
    ob = 0;
    for (int d = bc.nd - 1; d >= 0; --d) {
        int64_t q = i / bc.shape[d];
        int64_t r = i - q * bc.shape[d];
        oa += r * bc.sa[d];
        ob += r * bc.sb[d];
        i = q;
    }
}

template <typename TO>
__global__ void __launch_bounds__(256) k_unary_contig(int op, const float* __restrict__ a,
                                                      TO* __restrict__ out, int64_t n, float p0,
                                                      float p1) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        Store<TO>::put(out, i, unary_f(op, a[i], p0, p1));
}

// Contiguous unary op on 16-byte aligned input, 4 elements per thread per step.
template <typename TO>
__global__ void __launch_bounds__(256) k_unary_vec4(int op, const float* __restrict__ a,
                                                    TO* __restrict__ out, int64_t n, float p0,
                                                    float p1) {
    const int64_t n4 = n / 4;
    const float4* a4 = reinterpret_cast<const float4*>(a);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float4 x = a4[i];
        Store<TO>::put(out, 4 * i + 0, unary_f(op, x.x, p0, p1));
        Store<TO>::put(out, 4 * i + 1, unary_f(op, x.y, p0, p1));
        Store<TO>::put(out, 4 * i + 2, unary_f(op, x.z, p0, p1));
        Store<TO>::put(out, 4 * i + 3, unary_f(op, x.w, p0, p1));
    }
    for (int64_t i = 4 * n4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        Store<TO>::put(out, i, unary_f(op, a[i], p0, p1));
}

template <typename TO>
__global__ void __launch_bounds__(256) k_unary_strided(int op, const float* __restrict__ a,
                                                       TO* __restrict__ out, int64_t n, Bcast bc,
                                                       float p0, float p1) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t oa, ob;
        offsets(bc, i, oa, ob);
        Store<TO>::put(out, i, unary_f(op, a[oa], p0, p1));
    }
}

// Same-shape contiguous binary op, 4 elements per thread per step.
template <typename TO>
__global__ void __launch_bounds__(256) k_binary_vec4(int op, const float* __restrict__ a,
                                                     const float* __restrict__ b,
                                                     TO* __restrict__ out, int64_t n, float eps) {
    const int64_t n4 = n / 4;
    const float4* a4 = reinterpret_cast<const float4*>(a);
    const float4* b4 = reinterpret_cast<const float4*>(b);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
         i += (int64_t)gridDim.x * blockDim.x) {
        float4 x = a4[i], y = b4[i];
        Store<TO>::put(out, 4 * i + 0, binary_f(op, x.x, y.x, eps));
        Store<TO>::put(out, 4 * i + 1, binary_f(op, x.y, y.y, eps));
        Store<TO>::put(out, 4 * i + 2, binary_f(op, x.z, y.z, eps));
        Store<TO>::put(out, 4 * i + 3, binary_f(op, x.w, y.w, eps));
    }
    for (int64_t i = 4 * n4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        Store<TO>::put(out, i, binary_f(op, a[i], b[i], eps));
}

template <typename TO>
__global__ void __launch_bounds__(256) k_binary_strided(int op, const float* __restrict__ a,
                                                        const float* __restrict__ b,
                                                        TO* __restrict__ out, int64_t n,
                                                        Bcast bc, float eps) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t oa, ob;
        offsets(bc, i, oa, ob);
        Store<TO>::put(out, i, binary_f(op, a[oa], b[ob], eps));
    }
}

unsigned grid_for(int64_t n, int per_thread = 1) {
    int64_t blocks = (n / per_thread + 255) / 256;
    if (blocks < 1) blocks = 1;
    if (blocks > 65536) blocks = 65536;  // grid-stride loops cover the rest
    return (unsigned)blocks;
}

int check_out(const sp_tensor* out) {
    SP_REQUIRE(out && (out->ptr || numel(out) == 0), "output tensor missing");
    SP_REQUIRE(out->dtype == SP_F32 || out->dtype == SP_BF16, "output must be f32 or bf16");
    SP_REQUIRE(is_contiguous(out), "output must be contiguous");
    return SP_OK;
}

// Broadcast strides of ``t`` against the output shape (right-aligned, like torch).
int bcast_strides(const sp_tensor* t, const sp_tensor* out, int64_t* s) {
    const int off = out->ndim - t->ndim;
    SP_REQUIRE(off >= 0, "operand has more dims than the output");
    for (int d = 0; d < out->ndim; ++d) {
        if (d < off) { s[d] = 0; continue; }
        int64_t n = t->shape[d - off];
        SP_REQUIRE(n == out->shape[d] || n == 1, "operand does not broadcast to the output");
        s[d] = (n == 1) ? 0 : t->stride[d - off];
    }
    return SP_OK;
}

template <typename TO>
int unary_impl(int op, const sp_tensor* a, sp_tensor* out, float p0, float p1, hipStream_t st) {
    const int64_t n = numel(out);
    if (n == 0) return SP_OK;
    if (is_contiguous(a) && numel(a) == n && ((uintptr_t)a->ptr & 15) == 0) {
        hipLaunchKernelGGL(k_unary_vec4<TO>, dim3(grid_for(n, 4)), dim3(256), 0, st, op,
                           (const float*)a->ptr, (TO*)out->ptr, n, p0, p1);
    } else if (is_contiguous(a) && numel(a) == n) {
        hipLaunchKernelGGL(k_unary_contig<TO>, dim3(grid_for(n)), dim3(256), 0, st, op,
                           (const float*)a->ptr, (TO*)out->ptr, n, p0, p1);
    } else {
        Bcast bc{};
        bc.nd = out->ndim;
        for (int d = 0; d < out->ndim; ++d) bc.shape[d] = out->shape[d];
        int rc = bcast_strides(a, out, bc.sa);
        if (rc) return rc;
        hipLaunchKernelGGL(k_unary_strided<TO>, dim3(grid_for(n)), dim3(256), 0, st, op,
                           (const float*)a->ptr, (TO*)out->ptr, n, bc, p0, p1);
    }
    return launch_status("sp_op_unary");
}

template <typename TO>
int binary_impl(int op, const sp_tensor* a, const sp_tensor* b, sp_tensor* out, float eps,
                hipStream_t st) {
    const int64_t n = numel(out);
    if (n == 0) return SP_OK;
    const bool fast = is_contiguous(a) && is_contiguous(b) && numel(a) == n && numel(b) == n &&
                      (((uintptr_t)a->ptr | (uintptr_t)b->ptr) & 15) == 0;
    if (fast) {
        hipLaunchKernelGGL(k_binary_vec4<TO>, dim3(grid_for(n, 4)), dim3(256), 0, st, op,
                           (const float*)a->ptr, (const float*)b->ptr, (TO*)out->ptr, n, eps);
    } else {
        Bcast bc{};
        bc.nd = out->ndim;
        for (int d = 0; d < out->ndim; ++d) bc.shape[d] = out->shape[d];
        int rc = bcast_strides(a, out, bc.sa);
        if (rc) return rc;
        rc = bcast_strides(b, out, bc.sb);
        if (rc) return rc;
        hipLaunchKernelGGL(k_binary_strided<TO>, dim3(grid_for(n)), dim3(256), 0, st, op,
                           (const float*)a->ptr, (const float*)b->ptr, (TO*)out->ptr, n, bc, eps);
    }
    return launch_status("sp_op_binary");
}

// ---- channel-last helpers ----------------------------------------------------------------
template <int MODE>  // 0 dot, 1 sqnorm, 2 norm_last, 3 normalize_last
__global__ void __launch_bounds__(256) k_channel(const float* __restrict__ a,
                                                 const float* __restrict__ b,
                                                 float* __restrict__ out, int64_t nvox, int C,
                                                 float eps) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvox;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float* pa = a + i * C;
        if (MODE == 0) {
            const float* pb = b + i * C;
            float acc = pa[0] * pb[0];
            for (int c = 1; c < C; ++c) acc = acc + pa[c] * pb[c];
            out[i] = acc;
        } else if (MODE == 1) {
            float acc = pa[0] * pa[0];
            for (int c = 1; c < C; ++c) acc = acc + pa[c] * pa[c];
            out[i] = acc;
        } else {
            float ss = 0.f;
            for (int c = 0; c < C; ++c) ss = ss + pa[c] * pa[c];
            float nrm = sqrtf(ss);
            nrm = nrm < eps ? eps : nrm;
            if (MODE == 2) {
                out[i] = nrm;
            } else {
                for (int c = 0; c < C; ++c) out[i * C + c] = pa[c] / nrm;
            }
        }
    }
}

template <int MODE>
int channel_impl(const sp_tensor* a, const sp_tensor* b, sp_tensor* out, float eps, void* stream,
                 const char* what) {
    SP_REQUIRE(a && a->ptr && a->dtype == SP_F32 && a->ndim >= 1, "channel op: bad input");
    SP_REQUIRE(is_contiguous(a), "channel op: input must be contiguous");
    SP_REQUIRE(out && out->dtype == SP_F32 && is_contiguous(out), "channel op: bad output");
    const int C = (int)a->shape[a->ndim - 1];
    SP_REQUIRE(C >= 1, "channel op: empty channel axis");
    const int64_t nvox = numel(a) / C;
    if (MODE == 0) {
        SP_REQUIRE(b && b->ptr && b->dtype == SP_F32 && is_contiguous(b) && numel(b) == numel(a),
                   "dot: operands must have the same contiguous shape");
    }
    SP_REQUIRE(numel(out) == (MODE == 3 ? numel(a) : nvox), "channel op: output size mismatch");
    if (nvox == 0) return SP_OK;
    hipLaunchKernelGGL(k_channel<MODE>, dim3(grid_for(nvox)), dim3(256), 0, as_stream(stream),
                       (const float*)a->ptr, b ? (const float*)b->ptr : nullptr, (float*)out->ptr,
                       nvox, C, eps);
    return launch_status(what);
}

}  // namespace

extern "C" int sp_op_unary(int op, const sp_tensor* a, sp_tensor* out, double p0, double p1,
                           void* stream) {
    SP_REQUIRE(a && (a->ptr || numel(a) == 0) && a->dtype == SP_F32, "unary: input must be f32");
    SP_REQUIRE(op >= SP_NEG && op <= SP_SQUARE, "unary: unknown op");
    int rc = check_out(out);
    if (rc) return rc;
    if (out->dtype == SP_F32)
        return unary_impl<float>(op, a, out, (float)p0, (float)p1, as_stream(stream));
    return unary_impl<uint16_t>(op, a, out, (float)p0, (float)p1, as_stream(stream));
}

extern "C" int sp_op_binary(int op, const sp_tensor* a, const sp_tensor* b, sp_tensor* out,
                            double p0, void* stream) {
    SP_REQUIRE(a && b && a->dtype == SP_F32 && b->dtype == SP_F32, "binary: inputs must be f32");
    SP_REQUIRE(op >= SP_ADD && op <= SP_DIV, "binary: unknown op");
    int rc = check_out(out);
    if (rc) return rc;
    if (out->dtype == SP_F32)
        return binary_impl<float>(op, a, b, out, (float)p0, as_stream(stream));
    return binary_impl<uint16_t>(op, a, b, out, (float)p0, as_stream(stream));
}

extern "C" int sp_op_dot(const sp_tensor* a, const sp_tensor* b, sp_tensor* out, void* stream) {
    return channel_impl<0>(a, b, out, 0.f, stream, "sp_op_dot");
}

extern "C" int sp_op_sqnorm(const sp_tensor* a, sp_tensor* out, void* stream) {
    return channel_impl<1>(a, nullptr, out, 0.f, stream, "sp_op_sqnorm");
}

extern "C" int sp_op_norm_last(const sp_tensor* a, sp_tensor* out, double eps, void* stream) {
    return channel_impl<2>(a, nullptr, out, (float)eps, stream, "sp_op_norm_last");
}

extern "C" int sp_op_normalize_last(const sp_tensor* a, sp_tensor* out, double eps,
                                    void* stream) {
    return channel_impl<3>(a, nullptr, out, (float)eps, stream, "sp_op_normalize_last");
}

// ---- scalar programs: the scalar steps between fused levels (ops on reduction results) ----
namespace {

struct ScalarProgram { sp_scalar_op op[SP_SCALAR_PROGRAM_MAX]; int n; };

__device__ double scalar_unary(int op, double a, double p0, double p1) {
    switch (op) {
        case SP_NEG: return -a;
        case SP_EXP: return exp(a);
        case SP_LOG: return log(a < p0 ? p0 : a);
        case SP_SQRT: return sqrt(a < p0 ? p0 : a);
        case SP_SIN: return sin(a);
        case SP_COS: return cos(a);
        case SP_TANH: return tanh(a);
        case SP_CLAMP: return a < p0 ? p0 : (a > p1 ? p1 : a);
        case SP_SCALE: return a * p0;
        case SP_SQUARE: return a * a;
        default: return a;
    }
}

__device__ double scalar_binary(int op, double a, double b, double eps) {
    switch (op) {
        case SP_ADD: return a + b;
        case SP_SUB: return a - b;
        case SP_MUL: return a * b;
        default: return a / (fabs(b) < eps ? (b < 0.0 ? -eps : eps) : b);
    }
}

// An operand that an earlier op of the program wrote is taken from that op's result as held
// in registers, never re-read from memory (so no store-to-load ordering is needed).
__device__ double scalar_operand(const ScalarProgram& prog, const double* vals, int i,
                                 const double* ptr, double constant) {
    if (!ptr) return constant;
    for (int j = i - 1; j >= 0; --j)
        if (prog.op[j].out == ptr) return vals[j];
    return *ptr;
}

// One lane runs the ops in order, so an op may read the result of an earlier one.
__global__ void k_scalar_program(const ScalarProgram prog) {
    if (threadIdx.x != 0) return;
    double vals[SP_SCALAR_PROGRAM_MAX];
    for (int i = 0; i < prog.n; ++i) {
        const sp_scalar_op& o = prog.op[i];
        const double a = scalar_operand(prog, vals, i, o.a, o.av);
        const double b = scalar_operand(prog, vals, i, o.b, o.bv);
        vals[i] = o.binary ? scalar_binary(o.op, a, b, o.p0) : scalar_unary(o.op, a, o.p0, o.p1);
        *o.out = vals[i];
    }
}

}  // namespace

extern "C" int sp_scalar_program(const sp_scalar_op* ops, int n, void* stream) {
    SP_REQUIRE(ops && n > 0 && n <= SP_SCALAR_PROGRAM_MAX, "scalar_program: bad arguments");
    ScalarProgram prog;
    for (int i = 0; i < n; ++i) {
        const sp_scalar_op& o = ops[i];
        SP_REQUIRE(o.out, "scalar_program: missing output");
        SP_REQUIRE(o.binary ? (o.op >= SP_ADD && o.op <= SP_DIV) : (o.op >= SP_NEG && o.op <= SP_SQUARE),
                   "scalar_program: unknown op");
        prog.op[i] = o;
    }
    prog.n = n;
    hipLaunchKernelGGL(k_scalar_program, dim3(1), dim3(64), 0, as_stream(stream), prog);
    return launch_status("sp_scalar_program");
}

// Per-voxel 3x3 tensor operators (stenpy_engine.py:2674-2730, 2852-2867) on channel-last
// (..., 3, 3) FP32 fields.  Shared device routines live in sp_sym3.h so the fused templates
// compiled at run time use the same arithmetic.
#include "sp_common.h"
#include "sp_sym3.h"

using namespace sp;

namespace {

__global__ void __launch_bounds__(256) k_tensor3(int kind, const float* __restrict__ h,
                                                 float* __restrict__ out, int64_t nvox,
                                                 const double* __restrict__ flag) {
    const float reg = (flag != nullptr && flag[0] != 0.0) ? 1e-12f : 0.0f;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nvox;
         v += (int64_t)gridDim.x * blockDim.x) {
        float a[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) a[i] = h[v * 9 + i];
        switch (kind) {
            case SP_T_TRACE: out[v] = sym3_trace(a); break;
            case SP_T_DET: out[v] = sym3_det_regularised(a, reg); break;
            case SP_T_EIG: {
                float e[3];
                sym3_eigvalsh(a, e);
                out[v * 3 + 0] = e[0];
                out[v * 3 + 1] = e[1];
                out[v * 3 + 2] = e[2];
                break;
            }
            case SP_T_DEV: {
                float d[9];
                sym3_deviatoric(a, d);
#pragma unroll
                for (int i = 0; i < 9; ++i) out[v * 9 + i] = d[i];
                break;
            }
            default: {
                float inv[9];
                mat3_inverse(a, inv);
#pragma unroll
                for (int i = 0; i < 9; ++i) out[v * 9 + i] = inv[i];
            }
        }
    }
}

}  // namespace

extern "C" int sp_tensor3(int kind, const sp_tensor* h, sp_tensor* out, const double* flag,
                          void* stream) {
    SP_REQUIRE(h && h->ptr && h->dtype == SP_F32 && is_contiguous(h), "tensor op: bad input");
    SP_REQUIRE(h->ndim >= 2 && h->shape[h->ndim - 1] == 3 && h->shape[h->ndim - 2] == 3,
               "tensor op: need (*spatial, 3, 3)");
    SP_REQUIRE(kind >= SP_T_TRACE && kind <= SP_T_INV, "tensor op: unknown kind");
    SP_REQUIRE(out && out->ptr && out->dtype == SP_F32 && is_contiguous(out), "tensor op: bad output");
    const int64_t nvox = numel(h) / 9;
    const int64_t per = (kind == SP_T_EIG) ? 3 : (kind == SP_T_DEV || kind == SP_T_INV) ? 9 : 1;
    SP_REQUIRE(numel(out) == nvox * per, "tensor op: output size mismatch");
    if (nvox == 0) return SP_OK;
    int64_t blocks = (nvox + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_tensor3, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), kind,
                       (const float*)h->ptr, (float*)out->ptr, nvox, flag);
    return launch_status("sp_tensor3");
}

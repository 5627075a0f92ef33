// Plain per-operator stencil kernels (one kernel per graph node).
//
// Semantics are the CPU branches of the reference operators (the parity target):
//   * periodic first derivatives use the torch.roll form (x[i+1] - x[i-1]) / (2h);
//   * non-periodic first derivatives use torch.gradient: interior ((x[i+1]-x[i-1])/h)/2,
//     edges (-1.5 x0 + 2 x1 - 0.5 x2)/h and (0.5 x[n-3] - 2 x[n-2] + 1.5 x[n-1])/h
//     (edge_order 2) or one-sided first differences (edge_order 1);
//   * curl, gradient_nd, hessian, jacobian and the inner gradient of mean_curvature ignore
//     ``boundary`` and always use torch.gradient (stenpy_engine.py:1994-2003, 2037-2128);
//   * the Laplacian pads per ``boundary`` (replicate / zeros / circular / mirror).
// Each 3-D kernel maps one thread to one voxel, x (the contiguous axis) fastest, grid
// (ceil(n2/256), n1, n0); neighbour gathers are served by the caches.
#include "sp_common.h"

using namespace sp;

namespace {

// Value of x at coordinate q along an axis of length n after the Laplacian's padding; the
// caller multiplies the returned index by the axis stride.  Returns false for a zero pad.
__device__ __forceinline__ bool pad_index(int64_t q, int64_t n, int mode, int64_t& src) {
    if (q >= 0 && q < n) { src = q; return true; }
    switch (mode) {
        case SP_PERIODIC: src = q < 0 ? q + n : q - n; return true;
        case SP_NEUMANN: src = q < 0 ? 0 : n - 1; return true;
        case SP_REFLECT:
            if (n == 1) return false;
            // radius-1 mirror without repeating the edge
            src = q < 0 ? (-q < n - 1 ? -q : n - 1) : (2 * (n - 1) - q > 0 ? 2 * (n - 1) - q : 0);
            return true;
        default: return false;  // dirichlet and unknown names pad zeros
    }
}

// torch.gradient first derivative at coordinate p of an axis of length n; fetch(d) returns
// the value at offset d along that axis.
template <class Fetch>
__device__ __forceinline__ float tg_d1(const Fetch& fetch, int64_t p, int64_t n, float h,
                                       int order) {
    if (order == 2) {
        if (p == 0) return ((-1.5f * fetch(0) + 2.0f * fetch(1)) - 0.5f * fetch(2)) / h;
        if (p == n - 1) return ((0.5f * fetch(-2) - 2.0f * fetch(-1)) + 1.5f * fetch(0)) / h;
    } else {
        if (p == 0) return (fetch(1) - fetch(0)) / h;
        if (p == n - 1) return (fetch(0) - fetch(-1)) / h;
    }
    return ((fetch(1) - fetch(-1)) / h) * 0.5f;
}

__device__ __forceinline__ int64_t wrap(int64_t q, int64_t n) {
    return q < 0 ? q + n : (q >= n ? q - n : q);
}

struct Geom {
    int64_t n[3];
    int64_t s[3];  // element strides of the scalar (or per-channel) grid
};

// ------------------------------------------------------------------------------ axis pass
__global__ void __launch_bounds__(256) k_axis_pass(const float* __restrict__ x, int64_t outer,
                                                   int64_t n, int64_t inner, int64_t xo,
                                                   int64_t xn, int64_t xi, float* __restrict__ y,
                                                   int64_t yo, int64_t yn, int64_t yi, int mode,
                                                   int pad, float h, float scale, int accumulate) {
    const int64_t total = outer * n * inner;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = t % inner;
        const int64_t r = t / inner;
        const int64_t i = r % n;
        const int64_t o = r / n;
        const float* base = x + o * xo + j * xi;
        auto fetch = [&](int d) { return base[(i + d) * xn]; };
        float v;
        if (mode == SP_D1_PERIODIC) {
            v = (base[wrap(i + 1, n) * xn] - base[wrap(i - 1, n) * xn]) / (2.0f * h);
        } else if (mode == SP_D1_EDGE2 || mode == SP_D1_EDGE1) {
            v = tg_d1(fetch, i, n, h, mode == SP_D1_EDGE2 ? 2 : 1);
        } else {
            int64_t sp_, sm_;
            const float c = base[i * xn];
            const float xp = pad_index(i + 1, n, pad, sp_) ? base[sp_ * xn] : 0.0f;
            const float xm = pad_index(i - 1, n, pad, sm_) ? base[sm_ * xn] : 0.0f;
            v = ((xp + xm) - 2.0f * c) / (h * h);
        }
        float* dst = y + o * yo + i * yn + j * yi;
        *dst = accumulate ? (*dst + scale * v) : (scale == 1.0f ? v : scale * v);
    }
}

// ------------------------------------------------------------------------------ FIR pass
struct FirTaps { float w[SP_FIR_MAX_TAPS]; };

// Source index of padded coordinate q for a pad of radius r, as the reference's _pad1d fallback
// loops fill it: the left side wraps / mirrors / replicates, while on the right side cell
// n-1+k is filled from loop index r-k, so for r >= 2 a circular pad comes out mirrored
// (src = r-k) and a reflect pad reads n-2-r+k (stenpy_engine.py:1395-1413).  false: zero.
__device__ __forceinline__ bool fir_pad_index(int64_t q, int64_t n, int mode, int r, int64_t& src) {
    if (q >= 0 && q < n) { src = q; return true; }
    if (mode == SP_NEUMANN) { src = q < 0 ? 0 : n - 1; return true; }
    if (mode == SP_PERIODIC) { src = q < 0 ? q + n : r - (q - (n - 1)); return src >= 0 && src < n; }
    if (mode == SP_REFLECT && n > 1) {
        if (q < 0) { src = -q < n - 1 ? -q : n - 1; return true; }
        const int64_t v = n - 2 - r + (q - (n - 1));
        src = v > 0 ? v : 0;
        return true;
    }
    return false;
}

// out[i] = sum_k w[k] * x_pad[i + k - ntaps/2] along one axis, taps summed in order (the
// order of the reference's conv1d cross-correlation); x_pad follows the padding mode.
__global__ void __launch_bounds__(256) k_axis_fir(const float* __restrict__ x, int64_t outer,
                                                  int64_t n, int64_t inner, int64_t xo, int64_t xn,
                                                  int64_t xi, float* __restrict__ y, int64_t yo,
                                                  int64_t yn, int64_t yi, int ntaps, int pad,
                                                  const FirTaps taps) {
    const int64_t total = outer * n * inner;
    const int h = ntaps / 2;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = t % inner;
        const int64_t r = t / inner;
        const int64_t i = r % n;
        const int64_t o = r / n;
        const float* base = x + o * xo + j * xi;
        float acc = 0.0f;
        for (int k = 0; k < ntaps; ++k) {
            int64_t src;
            const float v = fir_pad_index(i + k - h, n, pad, h, src) ? base[src * xn] : 0.0f;
            acc += v * taps.w[k];
        }
        y[o * yo + i * yn + j * yi] = acc;
    }
}

// ------------------------------------------------------------------------------ 3-D kernels
#define SP_VOXEL_PROLOGUE(G)                                               \
    const int64_t px = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;    \
    const int64_t py = blockIdx.y;                                         \
    const int64_t pz = blockIdx.z;                                         \
    if (px >= (G).n[2]) return;                                            \
    const int64_t vox = (pz * (G).n[1] + py) * (G).n[2] + px;              \
    const int64_t pc[3] = {pz, py, px};                                    \
    (void)pc;

__global__ void __launch_bounds__(256) k_laplacian3d(const float* __restrict__ x,
                                                     float* __restrict__ out, Geom g, float h0,
                                                     float h1, float h2, int mode) {
    SP_VOXEL_PROLOGUE(g)
    const float c = x[vox];
    const float hs[3] = {h0, h1, h2};
    float acc = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int64_t base = vox - pc[d] * g.s[d];
        int64_t sp_, sm_;
        const float xp = pad_index(pc[d] + 1, g.n[d], mode, sp_) ? x[base + sp_ * g.s[d]] : 0.0f;
        const float xm = pad_index(pc[d] - 1, g.n[d], mode, sm_) ? x[base + sm_ * g.s[d]] : 0.0f;
        acc = acc + ((xp + xm) - 2.0f * c) / (hs[d] * hs[d]);
    }
    out[vox] = acc;
}

__global__ void __launch_bounds__(256) k_gradient3d(const float* __restrict__ x,
                                                    float* __restrict__ out, Geom g, float h,
                                                    int periodic) {
    SP_VOXEL_PROLOGUE(g)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int64_t base = vox - pc[d] * g.s[d];
        float v;
        if (periodic) {
            v = (x[base + wrap(pc[d] + 1, g.n[d]) * g.s[d]] -
                 x[base + wrap(pc[d] - 1, g.n[d]) * g.s[d]]) / (2.0f * h);
        } else {
            auto fetch = [&](int o) { return x[vox + o * g.s[d]]; };
            v = tg_d1(fetch, pc[d], g.n[d], h, g.n[d] >= 3 ? 2 : 1);
        }
        out[vox * 3 + d] = v;
    }
}

__global__ void __launch_bounds__(256) k_divergence3d(const float* __restrict__ v,
                                                      float* __restrict__ out, Geom g, float h0,
                                                      float h1, float h2, int periodic) {
    SP_VOXEL_PROLOGUE(g)
    const float hs[3] = {h0, h1, h2};
    float acc = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int64_t base = vox - pc[d] * g.s[d];
        float gd;
        if (periodic) {
            gd = (v[(base + wrap(pc[d] + 1, g.n[d]) * g.s[d]) * 3 + d] -
                  v[(base + wrap(pc[d] - 1, g.n[d]) * g.s[d]) * 3 + d]) / (2.0f * hs[d]);
        } else {
            auto fetch = [&](int o) { return v[(vox + o * g.s[d]) * 3 + d]; };
            gd = tg_d1(fetch, pc[d], g.n[d], hs[d], 2);
        }
        acc = acc + gd;
    }
    out[vox] = acc;
}

// d/d(axis a) of channel c of a 3-channel field with torch.gradient edges of ``order``.
__device__ __forceinline__ float dchan(const float* __restrict__ v, const Geom& g, int64_t vox,
                                       const int64_t* pc, int a, int c, float h, int order) {
    auto fetch = [&](int o) { return v[(vox + o * g.s[a]) * 3 + c]; };
    return tg_d1(fetch, pc[a], g.n[a], h, order);
}

__global__ void __launch_bounds__(256) k_curl3d(const float* __restrict__ v,
                                                float* __restrict__ out, Geom g, float h0,
                                                float h1, float h2) {
    SP_VOXEL_PROLOGUE(g)
    const float d1v2 = dchan(v, g, vox, pc, 1, 2, h1, 2);
    const float d2v1 = dchan(v, g, vox, pc, 2, 1, h2, 2);
    const float d2v0 = dchan(v, g, vox, pc, 2, 0, h2, 2);
    const float d0v2 = dchan(v, g, vox, pc, 0, 2, h0, 2);
    const float d0v1 = dchan(v, g, vox, pc, 0, 1, h0, 2);
    const float d1v0 = dchan(v, g, vox, pc, 1, 0, h1, 2);
    out[vox * 3 + 0] = d1v2 - d2v1;
    out[vox * 3 + 1] = d2v0 - d0v2;
    out[vox * 3 + 2] = d0v1 - d1v0;
}

__device__ __forceinline__ int nd_order(const Geom& g) {
    return (g.n[0] >= 3 && g.n[1] >= 3 && g.n[2] >= 3) ? 2 : 1;
}

// torch.gradient of a scalar field along axis a at voxel ``q`` (linear index) with coords qc.
__device__ __forceinline__ float dscal(const float* __restrict__ x, const Geom& g, int64_t q,
                                       const int64_t* qc, int a, float h, int order) {
    auto fetch = [&](int o) { return x[q + o * g.s[a]]; };
    return tg_d1(fetch, qc[a], g.n[a], h, order);
}

__global__ void __launch_bounds__(256) k_gradient_nd3d(const float* __restrict__ x,
                                                       float* __restrict__ out, Geom g, float h0,
                                                       float h1, float h2) {
    SP_VOXEL_PROLOGUE(g)
    const int order = nd_order(g);
    const float hs[3] = {h0, h1, h2};
#pragma unroll
    for (int d = 0; d < 3; ++d) out[vox * 3 + d] = dscal(x, g, vox, pc, d, hs[d], order);
}

__global__ void __launch_bounds__(256) k_hessian3d(const float* __restrict__ x,
                                                   float* __restrict__ out, Geom g, float h0,
                                                   float h1, float h2) {
    SP_VOXEL_PROLOGUE(g)
    const int order = nd_order(g);
    const float hs[3] = {h0, h1, h2};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            // H[i][j] = D_j (D_i f): the outer derivative along j evaluates g_i at p + o e_j.
            auto outer = [&](int o) {
                int64_t qc[3] = {pc[0], pc[1], pc[2]};
                qc[j] += o;
                const int64_t q = vox + o * g.s[j];
                return dscal(x, g, q, qc, i, hs[i], order);
            };
            out[vox * 9 + i * 3 + j] = tg_d1(outer, pc[j], g.n[j], hs[j], order);
        }
    }
}

__global__ void __launch_bounds__(256) k_jacobian3d(const float* __restrict__ v,
                                                    float* __restrict__ out, Geom g, float h0,
                                                    float h1, float h2) {
    SP_VOXEL_PROLOGUE(g)
    const int order = nd_order(g);
    const float hs[3] = {h0, h1, h2};
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d) out[vox * 9 + c * 3 + d] = dchan(v, g, vox, pc, d, c, hs[d], order);
}

__global__ void __launch_bounds__(256) k_mean_curvature3d(const float* __restrict__ x,
                                                          float* __restrict__ out, Geom g,
                                                          float h0, float h1, float h2,
                                                          int periodic, float eps) {
    SP_VOXEL_PROLOGUE(g)
    const int order = nd_order(g);
    const float hs[3] = {h0, h1, h2};
    // component d of the unit normal at the voxel p + o e_d (o in {-2..2}, inside the grid)
    auto nrm = [&](int d, int64_t o) {
        int64_t qc[3] = {pc[0], pc[1], pc[2]};
        qc[d] += o;
        const int64_t q = vox + o * g.s[d];
        const float g0 = dscal(x, g, q, qc, 0, hs[0], order);
        const float g1 = dscal(x, g, q, qc, 1, hs[1], order);
        const float g2 = dscal(x, g, q, qc, 2, hs[2], order);
        float m = sqrtf(g0 * g0 + g1 * g1 + g2 * g2);
        m = m < eps ? eps : m;
        const float gd = d == 0 ? g0 : (d == 1 ? g1 : g2);
        return gd / m;
    };
    float acc = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        float gd;
        if (periodic) {
            const int64_t n = g.n[d];
            const int64_t op = wrap(pc[d] + 1, n) - pc[d];
            const int64_t om = wrap(pc[d] - 1, n) - pc[d];
            gd = (nrm(d, op) - nrm(d, om)) / (2.0f * hs[d]);
        } else {
            auto fetch = [&](int o) { return nrm(d, o); };
            gd = tg_d1(fetch, pc[d], g.n[d], hs[d], 2);
        }
        acc = acc + gd;
    }
    out[vox] = acc;
}

__global__ void __launch_bounds__(256) k_material_derivative3d(const float* __restrict__ x,
                                                               const float* __restrict__ v,
                                                               float* __restrict__ out, Geom g,
                                                               float h0, float h1, float h2) {
    SP_VOXEL_PROLOGUE(g)
    const int order = nd_order(g);
    const float hs[3] = {h0, h1, h2};
    float acc = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) acc = acc + v[vox * 3 + d] * dscal(x, g, vox, pc, d, hs[d], order);
    out[vox] = acc;
}

// ------------------------------------------------------------------------------ host side
int geom3(const sp_tensor* t, int channels, Geom& g, const char* what) {
    SP_REQUIRE(t && t->ptr && t->dtype == SP_F32, std::string(what) + ": input must be f32");
    SP_REQUIRE(is_contiguous(t), std::string(what) + ": input must be contiguous");
    if (channels == 0) {
        SP_REQUIRE(t->ndim == 3, std::string(what) + ": expected a 3-D scalar field");
    } else {
        SP_REQUIRE(t->ndim == 4 && t->shape[3] == channels,
                   std::string(what) + ": expected a (n0,n1,n2," + std::to_string(channels) + ") field");
    }
    for (int d = 0; d < 3; ++d) g.n[d] = t->shape[d];
    g.s[2] = 1;
    g.s[1] = g.n[2];
    g.s[0] = g.n[1] * g.n[2];
    SP_REQUIRE(g.n[0] <= 65535 && g.n[1] <= 65535, std::string(what) + ": grid too large");
    return SP_OK;
}

int check_out3(const sp_tensor* out, int64_t expect, const char* what) {
    SP_REQUIRE(out && out->ptr && out->dtype == SP_F32 && is_contiguous(out),
               std::string(what) + ": output must be contiguous f32");
    SP_REQUIRE(numel(out) == expect, std::string(what) + ": output size mismatch");
    return SP_OK;
}

inline dim3 grid3(const Geom& g) { return dim3(cdiv(g.n[2], 256), (unsigned)g.n[1], (unsigned)g.n[0]); }

inline int64_t vox(const Geom& g) { return g.n[0] * g.n[1] * g.n[2]; }

int require_min(const Geom& g, int m, const char* what) {
    for (int d = 0; d < 3; ++d)
        SP_REQUIRE(g.n[d] >= m, std::string(what) + ": torch.gradient needs every axis of size >= " +
                                    std::to_string(m));
    return SP_OK;
}

// Collapse the dims of a view around ``axis`` into (outer, n, inner) with element strides.
int collapse(const sp_tensor* t, int axis, int64_t& outer, int64_t& n, int64_t& inner,
             int64_t& so, int64_t& sn, int64_t& si) {
    SP_REQUIRE(axis >= 0 && axis < t->ndim, "axis pass: axis out of range");
    n = t->shape[axis];
    sn = t->stride[axis];
    outer = 1;
    so = 0;
    {
        int64_t expect = -1;
        for (int d = axis - 1; d >= 0; --d) {
            if (t->shape[d] == 1) continue;
            if (expect < 0) {
                so = t->stride[d];
                expect = t->stride[d] * t->shape[d];
            } else {
                SP_REQUIRE(t->stride[d] == expect, "axis pass: outer dims not collapsible");
                expect *= t->shape[d];
            }
            outer *= t->shape[d];
        }
    }
    inner = 1;
    si = 1;
    {
        int64_t expect = -1;
        for (int d = t->ndim - 1; d > axis; --d) {
            if (t->shape[d] == 1) continue;
            if (expect < 0) {
                si = t->stride[d];
                expect = t->stride[d] * t->shape[d];
            } else {
                SP_REQUIRE(t->stride[d] == expect, "axis pass: inner dims not collapsible");
                expect *= t->shape[d];
            }
            inner *= t->shape[d];
        }
    }
    return SP_OK;
}

}  // namespace

extern "C" int sp_axis_pass(const sp_tensor* x, int axis, sp_tensor* out, int mode, int pad,
                            double h, double scale, int accumulate, void* stream) {
    SP_REQUIRE(x && x->ptr && x->dtype == SP_F32, "axis pass: input must be f32");
    SP_REQUIRE(out && out->ptr && out->dtype == SP_F32, "axis pass: output must be f32");
    SP_REQUIRE(x->ndim == out->ndim, "axis pass: rank mismatch");
    for (int d = 0; d < x->ndim; ++d)
        SP_REQUIRE(x->shape[d] == out->shape[d], "axis pass: shape mismatch");
    SP_REQUIRE(mode >= SP_D1_PERIODIC && mode <= SP_D2_PAD, "axis pass: unknown mode");
    int64_t xo, xn, xi, yo, yn, yi, outer, n, inner, o2, n2, i2;
    int rc = collapse(x, axis, outer, n, inner, xo, xn, xi);
    if (rc) return rc;
    rc = collapse(out, axis, o2, n2, i2, yo, yn, yi);
    if (rc) return rc;
    if (mode == SP_D1_EDGE2) SP_REQUIRE(n >= 3, "torch.gradient edge_order=2 needs an axis of size >= 3");
    if (mode == SP_D1_EDGE1) SP_REQUIRE(n >= 2, "torch.gradient needs an axis of size >= 2");
    const int64_t total = outer * n * inner;
    if (total == 0) return SP_OK;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_axis_pass, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream),
                       (const float*)x->ptr, outer, n, inner, xo, xn, xi, (float*)out->ptr, yo,
                       yn, yi, mode, pad, (float)h, (float)scale, accumulate);
    return launch_status("sp_axis_pass");
}

extern "C" int sp_laplacian3d(const sp_tensor* x, sp_tensor* out, const double dx[3],
                              int boundary, void* stream) {
    Geom g;
    int rc = geom3(x, 0, g, "laplacian");
    if (rc) return rc;
    rc = check_out3(out, vox(g), "laplacian");
    if (rc) return rc;
    if (vox(g) == 0) return SP_OK;
    hipLaunchKernelGGL(k_laplacian3d, grid3(g), dim3(256), 0, as_stream(stream),
                       (const float*)x->ptr, (float*)out->ptr, g, (float)dx[0], (float)dx[1],
                       (float)dx[2], boundary);
    return launch_status("sp_laplacian3d");
}

extern "C" int sp_gradient3d(const sp_tensor* x, sp_tensor* out, double dx, int boundary,
                             void* stream) {
    Geom g;
    int rc = geom3(x, 0, g, "gradient");
    if (rc) return rc;
    rc = check_out3(out, vox(g) * 3, "gradient");
    if (rc) return rc;
    if (boundary != SP_PERIODIC) {
        rc = require_min(g, 2, "gradient");
        if (rc) return rc;
    }
    if (vox(g) == 0) return SP_OK;
    hipLaunchKernelGGL(k_gradient3d, grid3(g), dim3(256), 0, as_stream(stream),
                       (const float*)x->ptr, (float*)out->ptr, g, (float)dx,
                       boundary == SP_PERIODIC ? 1 : 0);
    return launch_status("sp_gradient3d");
}

extern "C" int sp_divergence3d(const sp_tensor* v, sp_tensor* out, const double dx[3],
                               int boundary, void* stream) {
    Geom g;
    int rc = geom3(v, 3, g, "divergence");
    if (rc) return rc;
    rc = check_out3(out, vox(g), "divergence");
    if (rc) return rc;
    if (boundary != SP_PERIODIC) {
        rc = require_min(g, 3, "divergence");
        if (rc) return rc;
    }
    if (vox(g) == 0) return SP_OK;
    hipLaunchKernelGGL(k_divergence3d, grid3(g), dim3(256), 0, as_stream(stream),
                       (const float*)v->ptr, (float*)out->ptr, g, (float)dx[0], (float)dx[1],
                       (float)dx[2], boundary == SP_PERIODIC ? 1 : 0);
    return launch_status("sp_divergence3d");
}

extern "C" int sp_curl3d(const sp_tensor* v, sp_tensor* out, const double dx[3], void* stream) {
    Geom g;
    int rc = geom3(v, 3, g, "curl");
    if (rc) return rc;
    rc = check_out3(out, vox(g) * 3, "curl");
    if (rc) return rc;
    rc = require_min(g, 3, "curl");
    if (rc) return rc;
    hipLaunchKernelGGL(k_curl3d, grid3(g), dim3(256), 0, as_stream(stream), (const float*)v->ptr,
                       (float*)out->ptr, g, (float)dx[0], (float)dx[1], (float)dx[2]);
    return launch_status("sp_curl3d");
}

extern "C" int sp_gradient_nd3d(const sp_tensor* x, sp_tensor* out, const double dx[3],
                                void* stream) {
    Geom g;
    int rc = geom3(x, 0, g, "gradient_nd");
    if (rc) return rc;
    rc = check_out3(out, vox(g) * 3, "gradient_nd");
    if (rc) return rc;
    rc = require_min(g, 2, "gradient_nd");
    if (rc) return rc;
    hipLaunchKernelGGL(k_gradient_nd3d, grid3(g), dim3(256), 0, as_stream(stream),
                       (const float*)x->ptr, (float*)out->ptr, g, (float)dx[0], (float)dx[1],
                       (float)dx[2]);
    return launch_status("sp_gradient_nd3d");
}

extern "C" int sp_hessian3d(const sp_tensor* x, sp_tensor* out, const double dx[3],
                            void* stream) {
    Geom g;
    int rc = geom3(x, 0, g, "hessian");
    if (rc) return rc;
    rc = check_out3(out, vox(g) * 9, "hessian");
    if (rc) return rc;
    rc = require_min(g, 2, "hessian");
    if (rc) return rc;
    hipLaunchKernelGGL(k_hessian3d, grid3(g), dim3(256), 0, as_stream(stream),
                       (const float*)x->ptr, (float*)out->ptr, g, (float)dx[0], (float)dx[1],
                       (float)dx[2]);
    return launch_status("sp_hessian3d");
}

extern "C" int sp_jacobian3d(const sp_tensor* v, sp_tensor* out, const double dx[3],
                             void* stream) {
    Geom g;
    int rc = geom3(v, 3, g, "jacobian");
    if (rc) return rc;
    rc = check_out3(out, vox(g) * 9, "jacobian");
    if (rc) return rc;
    rc = require_min(g, 2, "jacobian");
    if (rc) return rc;
    hipLaunchKernelGGL(k_jacobian3d, grid3(g), dim3(256), 0, as_stream(stream),
                       (const float*)v->ptr, (float*)out->ptr, g, (float)dx[0], (float)dx[1],
                       (float)dx[2]);
    return launch_status("sp_jacobian3d");
}

extern "C" int sp_mean_curvature3d(const sp_tensor* x, sp_tensor* out, const double dx[3],
                                   int boundary, double eps, void* stream) {
    Geom g;
    int rc = geom3(x, 0, g, "mean_curvature");
    if (rc) return rc;
    rc = check_out3(out, vox(g), "mean_curvature");
    if (rc) return rc;
    rc = require_min(g, 3, "mean_curvature");
    if (rc) return rc;
    hipLaunchKernelGGL(k_mean_curvature3d, grid3(g), dim3(256), 0, as_stream(stream),
                       (const float*)x->ptr, (float*)out->ptr, g, (float)dx[0], (float)dx[1],
                       (float)dx[2], boundary == SP_PERIODIC ? 1 : 0, (float)eps);
    return launch_status("sp_mean_curvature3d");
}

extern "C" int sp_material_derivative3d(const sp_tensor* x, const sp_tensor* v, sp_tensor* out,
                                        const double dx[3], void* stream) {
    Geom g, gv;
    int rc = geom3(x, 0, g, "material_derivative");
    if (rc) return rc;
    rc = geom3(v, 3, gv, "material_derivative(velocity)");
    if (rc) return rc;
    for (int d = 0; d < 3; ++d) SP_REQUIRE(g.n[d] == gv.n[d], "material_derivative: shape mismatch");
    rc = check_out3(out, vox(g), "material_derivative");
    if (rc) return rc;
    rc = require_min(g, 2, "material_derivative");
    if (rc) return rc;
    hipLaunchKernelGGL(k_material_derivative3d, grid3(g), dim3(256), 0, as_stream(stream),
                       (const float*)x->ptr, (const float*)v->ptr, (float*)out->ptr, g,
                       (float)dx[0], (float)dx[1], (float)dx[2]);
    return launch_status("sp_material_derivative3d");
}

extern "C" int sp_axis_fir(const sp_tensor* x, int axis, sp_tensor* out, int ntaps,
                           const double* weights, int pad, void* stream) {
    SP_REQUIRE(x && x->ptr && x->dtype == SP_F32, "axis fir: input must be f32");
    SP_REQUIRE(out && out->ptr && out->dtype == SP_F32, "axis fir: output must be f32");
    SP_REQUIRE(x->ndim == out->ndim, "axis fir: rank mismatch");
    for (int d = 0; d < x->ndim; ++d)
        SP_REQUIRE(x->shape[d] == out->shape[d], "axis fir: shape mismatch");
    SP_REQUIRE(weights && ntaps >= 1 && ntaps <= SP_FIR_MAX_TAPS && (ntaps & 1),
               "axis fir: odd tap count 1..SP_FIR_MAX_TAPS required");
    SP_REQUIRE(pad >= SP_PERIODIC && pad <= SP_ZEROPAD, "axis fir: unknown padding");
    int64_t xo, xn, xi, yo, yn, yi, outer, n, inner, o2, n2, i2;
    int rc = collapse(x, axis, outer, n, inner, xo, xn, xi);
    if (rc) return rc;
    rc = collapse(out, axis, o2, n2, i2, yo, yn, yi);
    if (rc) return rc;
    SP_REQUIRE(pad != SP_PERIODIC || n >= ntaps / 2, "axis fir: periodic pad wider than the axis");
    FirTaps taps;
    for (int k = 0; k < SP_FIR_MAX_TAPS; ++k) taps.w[k] = k < ntaps ? (float)weights[k] : 0.0f;
    const int64_t total = outer * n * inner;
    if (total == 0) return SP_OK;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_axis_fir, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream),
                       (const float*)x->ptr, outer, n, inner, xo, xn, xi, (float*)out->ptr, yo, yn,
                       yi, ntaps, pad, taps);
    return launch_status("sp_axis_fir");
}

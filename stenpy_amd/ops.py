"""Operators registered under the reference's names, each backed by libstenpy_hip.

Every function keeps the reference signature (stenpy_engine.py:1218-2869) and the CPU-branch
semantics, and computes on the GPU through the C ABI.  Deliberate contract changes:

* storage and arithmetic are FP32 (FP64 / BF16 inputs are converted on entry; outputs are
  FP32, reductions return FP64 0-d tensors); the reference asserts FP64 everywhere;
* inputs must be on a HIP device -- there is no CPU path;
* ``dot`` and ``sqnorm`` are added (u . grad f and |curl u|^2 need them).

Pure data movement (stacking, slicing, reshaping) uses torch's device copies; every
arithmetic step is a hand-written kernel.
"""
from __future__ import annotations

import math
from typing import Any, Optional, Sequence, Tuple, Union

import torch

from . import _lib as L
from .registry import register_operator

F32 = torch.float32


def _f32(t: Any, device: Optional[torch.device] = None) -> torch.Tensor:
    if not isinstance(t, torch.Tensor):
        t = torch.tensor(float(t), dtype=F32)
    if t.dtype != F32:
        t = t.to(F32)
    if device is not None and t.device != device:
        t = t.to(device)
    return t


def _device_of(*ts) -> torch.device:
    for t in ts:
        if isinstance(t, torch.Tensor) and t.is_cuda:
            return t.device
    raise RuntimeError("stenpy_amd executes on the GPU only: no operand lives on a HIP device")


def _contig(t: torch.Tensor) -> torch.Tensor:
    return t if t.is_contiguous() else t.contiguous()


def _empty(shape, device, dtype=F32) -> torch.Tensor:
    return torch.empty(tuple(shape), dtype=dtype, device=device)


def _dxs(dx: Any, D: int) -> Tuple[float, ...]:
    if isinstance(dx, (int, float)):
        return (float(dx),) * D
    if isinstance(dx, torch.Tensor):
        dx = dx.tolist()
    dx = tuple(float(v) for v in dx)
    if len(dx) != D:
        raise ValueError(f"dx must have length {D}, got {len(dx)}")
    return dx


# ----------------------------------------------------------------------------- constants
def _constant(*, value) -> Any:
    """Source node: returns the held tensor/number unchanged (stenpy_engine.py:1218-1223)."""
    if isinstance(value, (int, float)):
        return torch.tensor(float(value), dtype=torch.float64)
    return value


register_operator("_constant", _constant, cost="low", bytes_out=4, kind="source")


# ----------------------------------------------------------------------------- point-wise
def _binary(op: str, a, b, eps: float = 0.0, out_dtype=F32) -> torch.Tensor:
    dev = _device_of(a, b)
    a, b = _f32(a, dev), _f32(b, dev)
    out = _empty(torch.broadcast_shapes(a.shape, b.shape), dev, out_dtype)
    return L.binary(op, a, b, out, eps)


def _unary(op: str, a, p0: float = 0.0, p1: float = 0.0, out_dtype=F32) -> torch.Tensor:
    dev = _device_of(a)
    a = _f32(a, dev)
    return L.unary(op, a, _empty(a.shape, dev, out_dtype), p0, p1)


def add(a, b):
    return _binary("add", a, b)


def sub(a, b):
    return _binary("sub", a, b)


def mul(a, b=None, scalar: float = None):
    if scalar is not None and b is None:
        return _unary("scale", a, float(scalar))
    if b is None:
        raise ValueError("mul: must provide either a second tensor 'b' or 'scalar'")
    return _binary("mul", a, b)


def div(a, b, eps: float = 1e-15):
    return _binary("div", a, b, eps)


def neg(a):
    return _unary("neg", a)


def clamp(a, lo: float = 0.0, hi: float = 1.0):
    return _unary("clamp", a, float(lo), float(hi))


def exp(a):
    return _unary("exp", a)


def log(a, eps: float = 1e-15):
    return _unary("log", a, float(eps))


def sqrt(a, eps: float = 0.0):
    return _unary("sqrt", a, float(eps))


def sin(a):
    return _unary("sin", a)


def cos(a):
    return _unary("cos", a)


def tanh(a):
    return _unary("tanh", a)


def cast_bf16(a):
    """Store a field as BF16 (round to nearest even)."""
    return _unary("copy", a, out_dtype=torch.bfloat16)


for _name, _fn, _fl in [("add", add, 1), ("sub", sub, 1), ("mul", mul, 1), ("div", div, 1),
                        ("neg", neg, 1), ("clamp", clamp, 2), ("exp", exp, 8), ("log", log, 8),
                        ("sqrt", sqrt, 8), ("sin", sin, 8), ("cos", cos, 8), ("tanh", tanh, 8),
                        ("cast_bf16", cast_bf16, 0)]:
    register_operator(_name, _fn, cost="low" if _fl <= 2 else "medium", flops=_fl,
                      bytes_in=8 if _name in ("add", "sub", "mul", "div") else 4, bytes_out=4,
                      kind="pointwise")


# ----------------------------------------------------------------------------- stencils
def _tg_mode(order: int) -> int:
    return L.D1_EDGE2 if order == 2 else L.D1_EDGE1


def laplacian(x, dx: Any = 1.0, boundary: str = "periodic"):
    """Compact Laplacian over every axis of ``x`` (stenpy_engine.py:1466-1516)."""
    dev = _device_of(x)
    x = _contig(_f32(x, dev))
    D = x.ndim
    dv = _dxs(dx, D)
    out = _empty(x.shape, dev)
    if D == 3:
        return L.stencil3d("laplacian", x, out, dv, L.boundary_code(boundary))
    for d in range(D):
        L.axis_pass(x, d, out, L.D2_PAD, L.boundary_code(boundary), dv[d], 1.0, accumulate=d > 0)
    return out


def gradient(t, dx: float = 1.0, boundary: str = "neumann", dim: Optional[int] = None):
    """Per-axis first derivatives, stacked channel-last (stenpy_engine.py:1780-1865)."""
    dev = _device_of(t)
    t = _contig(_f32(t, dev))
    periodic = boundary == "periodic"
    h = float(dx)

    def mode_for(d):
        if periodic:
            return L.D1_PERIODIC
        return _tg_mode(2 if t.shape[d] >= 3 else 1)

    if dim is not None:
        d = dim % t.ndim
        if t.shape[d] == 1:
            out = _empty(t.shape, dev)
            return L.unary("scale", t, out, 0.0)
        return L.axis_pass(t, d, _empty(t.shape, dev), mode_for(d), 0, h)
    out = _empty(tuple(t.shape) + (t.ndim,), dev)
    if t.ndim == 3:
        return L.stencil3d("gradient", t, out, h, 0 if periodic else 1)
    for d in range(t.ndim):
        L.axis_pass(t, d, out[..., d], mode_for(d), 0, h)
    return out


def divergence(t, dx: Any = 1.0, boundary: str = "neumann"):
    """Sum_d d_d t[..., d] (stenpy_engine.py:1870-1931)."""
    dev = _device_of(t)
    t = _contig(_f32(t, dev))
    D = t.shape[-1]
    dv = _dxs(dx, D)
    out = _empty(t.shape[:-1], dev)
    if t.ndim == 4 and D == 3:
        return L.stencil3d("divergence", t, out, dv, 0 if boundary == "periodic" else 1)
    mode = L.D1_PERIODIC if boundary == "periodic" else L.D1_EDGE2
    for d in range(D):
        L.axis_pass(t[..., d], d, out, mode, 0, dv[d], 1.0, accumulate=d > 0)
    return out


def curl(t, dx: Any = 1.0, boundary: str = "neumann"):
    """Curl with torch.gradient edges; ``boundary`` ignored (stenpy_engine.py:1936-2032)."""
    dev = _device_of(t)
    t = _contig(_f32(t, dev))
    if t.ndim < 2 or t.shape[-1] < 2:
        raise ValueError(f"curl: need (*spatial, components>=2), got {tuple(t.shape)}")
    D = t.ndim - 1
    if D not in (2, 3):
        raise ValueError(f"curl: only 2D or 3D vector fields supported, got spatial rank {D}")
    dv = _dxs(dx, D)
    if D == 3:
        if t.shape[-1] != 3:
            raise ValueError("curl: 3D vector field must have exactly 3 components")
        return L.stencil3d("curl", t, _empty(t.shape, dev), dv)
    out = _empty(t.shape[:-1], dev)
    L.axis_pass(t[..., 1], 0, out, L.D1_EDGE2, 0, dv[0])
    L.axis_pass(t[..., 0], 1, out, L.D1_EDGE2, 0, dv[1], -1.0, accumulate=True)
    return out


def _gradient_nd_generic(t: torch.Tensor, dv: Sequence[float]) -> torch.Tensor:
    """torch.gradient over every non-singleton axis (stenpy_engine.py:2037-2081)."""
    original = tuple(t.shape)
    ts = t.squeeze() if (1 in t.shape and t.squeeze().ndim > 0) else t
    if ts.ndim == 0:
        return torch.zeros(original + (0,), dtype=F32, device=t.device)
    ts = _contig(ts)
    order = 2 if min(ts.shape) >= 3 else 1
    if ts.ndim != len(dv):
        raise ValueError("gradient_nd: spacing does not match the squeezed rank")
    out = _empty(tuple(ts.shape) + (ts.ndim,), t.device)
    for d in range(ts.ndim):
        L.axis_pass(ts, d, out[..., d], _tg_mode(order), 0, dv[d])
    if out.shape[:-1] != original:
        out = out.reshape(list(original) + [out.shape[-1]])
    return out


def _plain3d(t: torch.Tensor) -> bool:
    return t.ndim == 3 and min(t.shape) >= 2


def gradient_nd(t, dx: Any = 1.0, boundary: str = "neumann", dim: Optional[int] = None):
    dev = _device_of(t)
    t = _contig(_f32(t, dev))
    if _plain3d(t):
        return L.stencil3d("gradient_nd", t, _empty(tuple(t.shape) + (3,), dev), _dxs(dx, 3))
    squeezed = sum(1 for s in t.shape if s != 1)
    return _gradient_nd_generic(t, _dxs(dx, squeezed if isinstance(dx, (int, float)) else len(dx)))


def hessian(t, dx: Any = 1.0, boundary: str = "neumann"):
    """H[..., i, j] = D_j D_i t with torch.gradient edges (stenpy_engine.py:2086-2128)."""
    dev = _device_of(t)
    t = _contig(_f32(t, dev))
    D = t.ndim
    dv = _dxs(dx, D)
    if _plain3d(t):
        return L.stencil3d("hessian", t, _empty(tuple(t.shape) + (3, 3), dev), dv)
    g = _gradient_nd_generic(t, dv)
    out = _empty(tuple(t.shape) + (D, D), dev)
    for i in range(D):
        out[..., i, :] = _gradient_nd_generic(_contig(g[..., i]), dv)
    return out


def jacobian(t, dx: Any = 1.0, boundary: str = "neumann"):
    """J[..., c, d] = D_d t[..., c] (stenpy_engine.py:2133-2175, CPU branch)."""
    dev = _device_of(t)
    t = _contig(_f32(t, dev))
    if t.ndim != 4 or t.shape[-1] != 3:
        raise ValueError(f"jacobian: need (nx,ny,nz,3), got {tuple(t.shape)}")
    dv = _dxs(dx, 3)
    if min(t.shape[:3]) >= 2:
        return L.stencil3d("jacobian", t, _empty(tuple(t.shape) + (3,), dev), dv)
    rows = [_gradient_nd_generic(_contig(t[..., i]), dv) for i in range(3)]
    return torch.stack(rows, dim=-2)


def mean_curvature(t, dx: Any = 1.0, boundary: str = "neumann", eps: float = 1e-15):
    """div(grad t / max(|grad t|, eps)) (stenpy_engine.py:2180-2195)."""
    dev = _device_of(t)
    t = _contig(_f32(t, dev))
    dv = _dxs(dx, t.ndim)
    if t.ndim == 3 and min(t.shape) >= 3:
        return L.stencil3d("mean_curvature", t, _empty(t.shape, dev), dv,
                           0 if boundary == "periodic" else 1, eps=eps)
    g = _gradient_nd_generic(t, dv)
    n = L.channel("normalize", _contig(g), _empty(g.shape, dev), eps=eps)
    return divergence(n, dv, boundary)


def surface_normals(t, dx: Any = 1.0, boundary: str = "neumann"):
    """Unit normals (-gx, -gy, 1)/|.| of a 2-D height map (stenpy_engine.py:2199-2219)."""
    dev = _device_of(t)
    t = _contig(_f32(t, dev))
    if t.ndim != 2:
        raise ValueError(f"surface_normals: expected 2D field, got {t.ndim}D")
    dv = _dxs(dx, 2)
    n = _empty(tuple(t.shape) + (3,), dev)
    L.axis_pass(t, 0, n[..., 0], L.D1_EDGE2, 0, dv[0], -1.0)
    L.axis_pass(t, 1, n[..., 1], L.D1_EDGE2, 0, dv[1], -1.0)
    n[..., 2].fill_(1.0)
    return L.channel("normalize", n, _empty(n.shape, dev), eps=1e-15)


def material_derivative(t, velocity, dx: Any = 1.0, boundary: str = "neumann"):
    """Sum_d v_d (gradient_nd t)_d (stenpy_engine.py:2223-2243)."""
    dev = _device_of(t, velocity)
    t = _contig(_f32(t, dev))
    v = _contig(_f32(velocity, dev))
    if _plain3d(t) and v.shape == tuple(t.shape) + (3,):
        return L.stencil3d("material_derivative", t, _empty(t.shape, dev), _dxs(dx, 3), v=v)
    g = _gradient_nd_generic(t, _dxs(dx, t.ndim))
    return dot(v, g)


for _name, _fn, _r, _cost in [("laplacian", laplacian, 1, "high"), ("gradient", gradient, 1, "medium"),
                              ("divergence", divergence, 1, "medium"), ("curl", curl, 1, "high"),
                              ("gradient_nd", gradient_nd, 1, "high"), ("hessian", hessian, 2, "very_high"),
                              ("jacobian", jacobian, 1, "high"),
                              ("mean_curvature", mean_curvature, 2, "very_high"),
                              ("surface_normals", surface_normals, 1, "high"),
                              ("material_derivative", material_derivative, 1, "high")]:
    register_operator(_name, _fn, radius=_r, cost=_cost,
                      exchange_dims=[0] if _name in ("laplacian", "gradient", "divergence") else None,
                      kind="stencil")


# ----------------------------------------------------------------------------- separable filters
def moving_average(t, window: int = 3, boundary: str = "neumann"):
    """Box filter of width ``window`` along every axis, axis 0 first, each pass padded by
    ``boundary`` (stenpy_engine.py:2385-2408).  An even window changes the length in the
    reference's conv1d formulation and fails there; it is rejected here."""
    dev = _device_of(t)
    t = _contig(_f32(t, dev))
    if t.ndim > 3:
        raise NotImplementedError("moving_average: only 1-D/2-D/3-D supported")
    window = int(window)
    if window < 1 or window % 2 == 0:
        raise ValueError(f"moving_average: window must be odd, got {window}")
    w = [1.0 / window] * window
    out = t
    for d in range(t.ndim):
        out = L.axis_fir(out, d, _empty(out.shape, dev), w, L.boundary_code(boundary))
    return out


def conv3d_spectral(a, boundary: str = "periodic", dx: Any = None):
    """27-point Gaussian blur, (1, 2, 1)/4 along each axis with periodic wrap: the reference's
    CPU path convolves through the FFT, which wraps whatever ``boundary`` says
    (stenpy_engine.py:2800-2846).  ``dx`` (added by the compiler for stencil ops) is ignored."""
    dev = _device_of(a)
    a = _contig(_f32(a, dev))
    if a.ndim != 3 or min(a.shape) < 3:
        raise ValueError(f"conv3d_spectral: expected a 3-D field with every axis >= 3, got {tuple(a.shape)}")
    out = a
    for d in range(3):
        out = L.axis_fir(out, d, _empty(out.shape, dev), [0.25, 0.5, 0.25], L.BOUNDARY["periodic"])
    return out


register_operator("moving_average", moving_average, radius=1, halo_l=1, halo_r=1, cost="medium",
                  kind="stencil")
register_operator("conv3d_spectral", conv3d_spectral, radius=1, halo_l=1, halo_r=1, cost="very_high",
                  flops=2, bytes_in=8, bytes_out=8, kind="stencil")


# ----------------------------------------------------------------------------- spectral
def _c64_or_f32(t, dev) -> torch.Tensor:
    if not isinstance(t, torch.Tensor):
        t = torch.tensor(float(t))
    if torch.is_complex(t):
        t = t.to(torch.complex64)
    elif t.dtype != F32:
        t = t.to(F32)
    if t.device != dev:
        t = t.to(dev)
    return _contig(t)


def _fft_norm(norm: str) -> str:
    return "backward" if norm == "numpy_compatible" else norm


def fft(t, norm: str = "backward", alloc=None):
    """n-D FFT over every axis, complex64 result computed in FP64 (stenpy_engine.py:2875-2885;
    the reference returns complex128)."""
    dev = _device_of(t)
    t = _c64_or_f32(t, dev)
    return L.fft(t, _empty(t.shape, dev, torch.complex64), False, _fft_norm(norm))


def ifft(t, norm: str = "backward", alloc=None):
    """Real part of the n-D inverse FFT, FP32 (stenpy_engine.py:2887-2896)."""
    dev = _device_of(t)
    t = _c64_or_f32(t, dev)
    return L.fft(t, _empty(t.shape, dev), True, _fft_norm(norm))


def fft_complex(t, norm: str = "backward", alloc=None):
    """n-D FFT, complex64; ``norm="numpy_compatible"`` means "backward"
    (stenpy_engine.py:2901-2908)."""
    return fft(t, norm)


def ifft_complex(t, norm: str = "backward", alloc=None):
    """n-D inverse FFT keeping the complex result (stenpy_engine.py:2910-2918)."""
    dev = _device_of(t)
    t = _c64_or_f32(t, dev)
    return L.fft(t, _empty(t.shape, dev, torch.complex64), True, _fft_norm(norm))


def spectral_gradient(t, dx: Any = 1.0, dim: int = 0, norm: str = "backward", alloc=None):
    """Derivative along ``dim`` through the FFT: Re(ifftn(i k_dim fftn(t))), k = 2 pi fftfreq
    (stenpy_engine.py:2979-2992).  The transforms' norm cancels, as in the reference."""
    dev = _device_of(t)
    t = _contig(_f32(t, dev))
    dim = int(dim)
    if not -t.ndim <= dim < t.ndim:
        raise IndexError(f"spectral_gradient: dim {dim} out of range for a {t.ndim}-D field")
    return L.spectral(t, _empty(t.shape, dev), _dxs(dx, t.ndim), L.SPECTRAL_GRADIENT, dim % t.ndim)


def spectral_laplacian(t, dx: Any = 1.0, norm: str = "backward", alloc=None):
    """Laplacian through the FFT: Re(ifftn(-|k|^2 fftn(t))) (stenpy_engine.py:2997-3010)."""
    dev = _device_of(t)
    t = _contig(_f32(t, dev))
    return L.spectral(t, _empty(t.shape, dev), _dxs(dx, t.ndim), L.SPECTRAL_LAPLACIAN)


register_operator("fft", fft, cost="high")
register_operator("ifft", ifft, cost="high")
register_operator("fft_complex", fft_complex, cost="high")
register_operator("ifft_complex", ifft_complex, cost="high")
register_operator("spectral_gradient", spectral_gradient, cost="high")
register_operator("spectral_laplacian", spectral_laplacian, cost="high")


# ----------------------------------------------------------------------------- statistics and integrals
def cumulative_integral(t, dx: float = 1.0, dim: int = 0, alloc=None):
    """Cumulative trapezoid along ``dim`` with a leading 0, accumulated in FP64
    (stenpy_engine.py:2462-2483)."""
    dev = _device_of(t)
    t = _contig(_f32(t, dev))
    dim = int(dim)
    if not -t.ndim <= dim < t.ndim:
        raise IndexError(f"cumulative_integral: dim {dim} out of range for a {t.ndim}-D field")
    return L.cumulative_trapezoid(t, _empty(t.shape, dev), dim % t.ndim, float(dx))


def surface_integral(t, normals, area_weights, alloc=None):
    """sum_i w_i sum_c t[i, c] normals[i, c] in FP64 (stenpy_engine.py:2486-2496).  ``t`` has the
    normals' shape or a last axis of length 1; ``area_weights`` the normals' shape without the
    last axis."""
    dev = _device_of(t, normals, area_weights)
    t, n, w = (_contig(_f32(v, dev)) for v in (t, normals, area_weights))
    if n.ndim < 1 or w.shape != n.shape[:-1] or t.shape not in (n.shape, n.shape[:-1] + (1,)):
        raise ValueError(f"surface_integral: shapes t {tuple(t.shape)}, normals {tuple(n.shape)}, "
                         f"area_weights {tuple(w.shape)} are not supported")
    return L.surface_integral(t, n, w)


def _pair(kind: str, a, b):
    dev = _device_of(a, b)
    a, b = _contig(_f32(a, dev)), _contig(_f32(b, dev))
    if a.shape != b.shape:
        raise ValueError(f"{kind}: shape mismatch {tuple(a.shape)} vs {tuple(b.shape)}")
    return L.pair_stat(kind, a, b)


def covariance(a, b, alloc=None):
    """dot(a - mean a, b - mean b) / max(n - 1, 1), FP64 (stenpy_engine.py:2575-2592)."""
    return _pair("covariance", a, b)


def correlation(a, b, alloc=None):
    """dot(a', b') / (|a'| |b'| + 1e-15) of the centred fields, FP64 (stenpy_engine.py:2595-2615)."""
    return _pair("correlation", a, b)


def entropy(t, dim=None, eps: float = 1e-15, alloc=None):
    """-sum p log(p + eps), p = |t| / (sum |t| + eps): over everything (FP64 scalar) or along
    ``dim`` (stenpy_engine.py:2638-2669)."""
    dev = _device_of(t)
    t = _contig(_f32(t, dev))
    if dim is not None:
        dim = int(dim)
        if not -t.ndim <= dim < t.ndim:
            raise IndexError(f"entropy: dim {dim} out of range for a {t.ndim}-D field")
        dim %= t.ndim
    return L.entropy(t, dim, float(eps))


register_operator("cumulative_integral", cumulative_integral, cost="low")
register_operator("surface_integral", surface_integral, cost="medium")
register_operator("covariance", covariance, cost="medium")
register_operator("correlation", correlation, cost="medium")
register_operator("entropy", entropy, cost="medium")


# ----------------------------------------------------------------------------- geometry
def distance_transform(t, dx: Any = 1.0, alloc=None, method: str = "auto"):
    """Exact Euclidean distance transform of (t == 0) with per-axis sampling ``dx``: 0 where
    t != 0, elsewhere the distance to the nearest voxel with t != 0 -- the reference's CPU
    branch, scipy's distance_transform_edt (stenpy_engine.py:3115-3156; its Triton JFA branch
    seeds the opposite set and is approximate).  A field without a nonzero voxel gives +inf."""
    dev = _device_of(t)
    t = _contig(_f32(t, dev))
    if t.ndim > 3:
        raise NotImplementedError("distance_transform: only 1-D/2-D/3-D supported")
    return L.distance_transform(t, _empty(t.shape, dev), _dxs(dx, t.ndim))


register_operator("distance_transform", distance_transform, cost="very_high")


# ----------------------------------------------------------------------------- channel-last helpers
def stack_components(*tensors, dim: int = -1):
    dev = _device_of(*tensors)
    return torch.stack([_f32(t, dev) for t in tensors], dim=dim)


def select_component(t, index: int):
    dev = _device_of(t)
    return _f32(t, dev)[..., index].contiguous()


def norm_last(t, eps: float = 1e-15):
    dev = _device_of(t)
    t = _contig(_f32(t, dev))
    return L.channel("norm_last", t, _empty(tuple(t.shape[:-1]) + (1,), dev), eps=eps)


def div_last(a, b):
    dev = _device_of(a, b)
    a, b = _f32(a, dev), _f32(b, dev)
    if b.ndim == a.ndim - 1:
        b = b.unsqueeze(-1)
    return _binary("div", a, b, 0.0)


def neg_stack3(gx, gy):
    dev = _device_of(gx, gy)
    gx, gy = _f32(gx, dev), _f32(gy, dev)
    n = torch.stack([neg(gx), neg(gy), torch.ones_like(gx)], dim=-1)
    return L.channel("normalize", n, _empty(n.shape, dev), eps=1e-15)


def scale_eye(scalar_field, D: int):
    dev = _device_of(scalar_field)
    s = _f32(scalar_field, dev)
    eye = torch.eye(int(D), dtype=F32, device=dev)
    return _binary("mul", s.unsqueeze(-1).unsqueeze(-1), eye)


def velocity_dot_grad(velocity_component, grad_component):
    return _binary("mul", velocity_component, grad_component)


def dot(u, g):
    """Extension: sum_c u[..., c] * g[..., c] folded left."""
    dev = _device_of(u, g)
    u, g = _contig(_f32(u, dev)), _contig(_f32(g, dev))
    if u.shape != g.shape:
        raise ValueError(f"dot: shape mismatch {tuple(u.shape)} vs {tuple(g.shape)}")
    return L.channel("dot", u, _empty(u.shape[:-1], dev), b=g)


def sqnorm(v):
    """Extension: sum_c v[..., c]^2 folded left."""
    dev = _device_of(v)
    v = _contig(_f32(v, dev))
    return L.channel("sqnorm", v, _empty(v.shape[:-1], dev))


for _name, _fn in [("stack_components", stack_components), ("select_component", select_component),
                   ("norm_last", norm_last), ("div_last", div_last), ("neg_stack3", neg_stack3),
                   ("scale_eye", scale_eye), ("velocity_dot_grad", velocity_dot_grad),
                   ("dot", dot), ("sqnorm", sqnorm)]:
    register_operator(_name, _fn, radius=0, cost="low",
                      kind="pointwise" if _name in ("dot", "sqnorm", "velocity_dot_grad") else "opaque")


# ----------------------------------------------------------------------------- reductions
def _no_dim(name, dim, keepdim=False):
    if dim is not None or keepdim:
        raise NotImplementedError(f"{name}: only full reductions (dim=None) run on the GPU backend")


def op_sum(t, dim=None, keepdim: bool = False):
    _no_dim("sum", dim, keepdim)
    return L.reduce("sum", _f32(t, _device_of(t)))


def mean(t, dim=None, keepdim: bool = False):
    _no_dim("mean", dim, keepdim)
    return L.reduce("mean", _f32(t, _device_of(t)))


def norm_l2(t):
    return L.reduce("norm_l2", _f32(t, _device_of(t)))


def min_max(t):
    return L.reduce("min_max", _f32(t, _device_of(t)))


def variance(t, dim=None, unbiased: bool = True, keepdim: bool = False):
    _no_dim("variance", dim, keepdim)
    return L.reduce("variance" if unbiased else "variance_biased", _f32(t, _device_of(t)))


def integrate(t, dx: Any = 1.0, dims=None, method: str = "simpson"):
    """Composite Simpson/trapezoid over every axis of a 3-D field (stenpy_engine.py:2436-2457)."""
    dev = _device_of(t)
    t = _f32(t, dev)
    if dims is not None and tuple(sorted((dims,) if isinstance(dims, int) else dims)) != (0, 1, 2):
        raise NotImplementedError("integrate: only the full 3-D integral runs on the GPU backend")
    if t.ndim != 3:
        raise NotImplementedError("integrate: the GPU backend integrates 3-D fields")
    dv = _dxs(dx, 3)
    return L.reduce("integrate", t, (dv[0], dv[1], dv[2], 0.0 if method == "simpson" else 1.0))


for _name, _fn in [("sum", op_sum), ("mean", mean), ("norm_l2", norm_l2), ("min_max", min_max),
                   ("variance", variance), ("integrate", integrate)]:
    register_operator(_name, _fn, cost="low" if _name != "integrate" else "medium", kind="reduction")


# ----------------------------------------------------------------------------- tensor ops
def _mat3(t, name: str) -> torch.Tensor:
    dev = _device_of(t)
    t = _contig(_f32(t, dev))
    if t.ndim < 2 or t.shape[-1] != t.shape[-2]:
        raise ValueError(f"{name}: need (*spatial, D, D), got {tuple(t.shape)}")
    if t.shape[-1] != 3:
        raise NotImplementedError(f"{name}: the GPU backend handles 3x3 tensors")
    return t


def trace(t):
    t = _mat3(t, "trace")
    return L.tensor3("trace", t, _empty(t.shape[:-2], t.device))


def determinant(t):
    """det with the reference's global +1e-12 I regularisation (stenpy_engine.py:2683-2712)."""
    t = _mat3(t, "determinant")
    flag = L.reduce("any_small_diag", t).reshape(1)
    return L.tensor3("determinant", t, _empty(t.shape[:-2], t.device), flag)


def eigenvalues(t):
    t = _mat3(t, "eigenvalues")
    return L.tensor3("eigenvalues", t, _empty(tuple(t.shape[:-1]), t.device))


def inverse(t):
    t = _mat3(t, "inverse")
    return L.tensor3("inverse", t, _empty(t.shape, t.device))


def deviatoric(t):
    t = _mat3(t, "deviatoric")
    return L.tensor3("deviatoric", t, _empty(t.shape, t.device))


for _name, _fn, _cost in [("trace", trace, "low"), ("determinant", determinant, "high"),
                          ("eigenvalues", eigenvalues, "very_high"), ("inverse", inverse, "very_high"),
                          ("deviatoric", deviatoric, "medium")]:
    register_operator(_name, _fn, cost=_cost, kind="tensor")

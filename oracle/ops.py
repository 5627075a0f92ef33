"""CPU restatement of the reference's torch eager operator semantics.

TEST INFRASTRUCTURE ONLY.  Nothing in the ``stenpy_amd`` package may import this module:
only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` use it,
and only as the checker (or the timed CPU baseline), never as the thing measured.

Every function takes a ``dtype``-carrying tensor and computes in that dtype, so the same code
is the FP64 oracle and the FP32 twin ("ref32") used by the 2x error rule.  Each function
cites the reference lines (stenpy_engine.py) whose CPU branch it restates; the reference's
Triton/GPU branches are not the target.  ``dot`` and ``sqnorm`` have no reference op: they are
defined as the reference-vocabulary compositions ``sum_c select_component(u,c) *
select_component(G,c)`` (left fold) and ``sum_c select_component(C,c)**2``.

How each function is pinned (tests/test_oracle.py): periodic stencils by the discrete Fourier
symbols of single modes, one-sided edges by exact derivatives of polynomials of degree <= 2,
hessian/tensor ops by the quadratic i^2+j^2+k^2 (H = 2I) and by numpy.linalg, reductions by
math.fsum / numpy, Simpson by exact polynomial integrals, the FFT ops by numpy.fft and the exact
spectral derivatives of Fourier modes, and everything else (the distance transform included) by
brute-force loops on tiny grids written from the formulas.
"""
from __future__ import annotations

import math
from typing import Any, Optional, Sequence, Tuple, Union

import torch
import torch.nn.functional as F

Tensor = torch.Tensor


# ----------------------------------------------------------------------------- point-wise
def add(a: Tensor, b: Tensor) -> Tensor:          # stenpy_engine.py:1228-1235
    return torch.add(a, b)


def sub(a: Tensor, b: Tensor) -> Tensor:          # stenpy_engine.py:1238-1245
    return torch.sub(a, b)


def mul(a: Tensor, b: Optional[Tensor] = None, scalar: Optional[float] = None) -> Tensor:
    # stenpy_engine.py:1248-1260: ``scalar`` builds a full tensor like ``a``
    if scalar is not None and b is None:
        b = torch.full_like(a, scalar)
    return torch.mul(a, b)


def div(a: Tensor, b: Tensor, eps: float = 1e-15) -> Tensor:
    # stenpy_engine.py:1263-1279: |b| < eps -> sign-preserving +-eps (b == 0 -> +eps)
    e = torch.as_tensor(eps, dtype=b.dtype)
    b_safe = torch.where(b.abs() < e, torch.where(b < 0, -e, e), b)
    return torch.div(a, b_safe)


def neg(a: Tensor) -> Tensor:                     # stenpy_engine.py:1282-1289
    return torch.neg(a)


def clamp(a: Tensor, lo: float = 0.0, hi: float = 1.0) -> Tensor:   # 1292-1305
    return torch.clamp(a, lo, hi)


def exp(a: Tensor) -> Tensor:                     # 1311-1318
    return torch.exp(a)


def log(a: Tensor, eps: float = 1e-15) -> Tensor:  # 1321-1328
    return torch.log(a.clamp(min=eps))


def sqrt(a: Tensor, eps: float = 0.0) -> Tensor:   # 1331-1338
    return torch.sqrt(a.clamp(min=eps))


def sin(a: Tensor) -> Tensor:                     # 1341-1347
    return torch.sin(a)


def cos(a: Tensor) -> Tensor:                     # 1350-1356
    return torch.cos(a)


def tanh(a: Tensor) -> Tensor:                    # 1360-1366
    return torch.tanh(a)


# ----------------------------------------------------------------------------- padding
_PAD_MODE = {"neumann": "replicate", "dirichlet": "constant", "periodic": "circular",
             "reflect": "reflect"}


def pad1d(t: Tensor, dim: int, r: int, bc: str) -> Tensor:
    """Pad ``r`` cells on both sides of ``dim`` (stenpy_engine.py:1374-1416).

    neumann -> replicate, dirichlet -> zeros, periodic -> circular, reflect -> mirror without
    repeating the edge; any other name pads zeros.  F.pad does not support every
    mode/rank pair, so the padding is built by explicit index maps (same values).
    """
    mode = _PAD_MODE.get(bc, "constant")
    n = t.shape[dim]
    idx = torch.arange(-r, n + r)
    if mode == "reflect" and n == 1:
        mode = "constant"          # the fallback leaves a size-1 axis zero-padded
    if mode == "constant":
        inner = t
        shape = list(t.shape)
        shape[dim] = n + 2 * r
        out = torch.zeros(shape, dtype=t.dtype)
        sl = [slice(None)] * t.ndim
        sl[dim] = slice(r, r + n)
        out[tuple(sl)] = inner
        return out
    # Right-hand pad cell n-1+k (k = 1..r) is filled from loop index i = r-k of the
    # reference fallback, so for r >= 2 the right side comes out mirrored (quirk kept).
    k_right = idx - (n - 1)
    if mode == "replicate":
        src = idx.clamp(0, n - 1)
    elif mode == "circular":
        src = torch.where(idx < 0, idx + n, torch.where(idx > n - 1, r - k_right, idx))
    else:  # reflect: left offset -k reads min(k, n-1)
        src = torch.where(idx < 0, (-idx).clamp(max=n - 1),
                          torch.where(idx > n - 1, (n - 2 - r + k_right).clamp(min=0), idx))
    return t.index_select(dim, src)


# ----------------------------------------------------------------------------- stencils
def _dx_list(dx: Any, D: int) -> Tuple[float, ...]:
    if isinstance(dx, (int, float)):
        return (float(dx),) * D
    dx = tuple(float(v) for v in dx)
    if len(dx) != D:
        raise ValueError(f"dx must have length {D}, got {len(dx)}")
    return dx


def laplacian(x: Tensor, dx: Any = 1.0, boundary: str = "periodic") -> Tensor:
    """Compact 7-point Laplacian over *all* dims of x (stenpy_engine.py:1478-1512)."""
    D = x.ndim
    dv = torch.tensor(_dx_list(dx, D), dtype=x.dtype)
    out = torch.zeros_like(x)
    for d in range(D):
        if boundary == "periodic":
            xp = torch.roll(x, -1, d)
            xm = torch.roll(x, 1, d)
        else:
            p = pad1d(x, d, 1, boundary)
            n = x.shape[d]
            xp = p.narrow(d, 2, n)
            xm = p.narrow(d, 0, n)
        out.add_((xp + xm - 2.0 * x) / (dv[d] ** 2))
    return out


def gradient(t: Tensor, dx: float = 1.0, boundary: str = "neumann",
             dim: Optional[int] = None) -> Tensor:
    """Central differences, periodic roll or torch.gradient edges (stenpy_engine.py:1780-1865)."""
    periodic = boundary == "periodic"

    def one(d: int) -> Tensor:
        if periodic:
            return (torch.roll(t, -1, d) - torch.roll(t, 1, d)) / (2.0 * dx)
        order = 2 if t.shape[d] >= 3 else 1
        return torch.gradient(t, spacing=dx, dim=d, edge_order=order)[0]

    if dim is not None:
        if t.shape[dim] == 1:
            return torch.zeros_like(t)
        return one(dim)
    return torch.stack([one(d) for d in range(t.ndim)], dim=-1)


def divergence(t: Tensor, dx: Any = 1.0, boundary: str = "neumann") -> Tensor:
    """Sum_d d_d t[...,d] of a channel-last field (stenpy_engine.py:1870-1931)."""
    D = t.shape[-1]
    dxs = _dx_list(dx, D)
    out = torch.zeros_like(t[..., 0])
    for d in range(D):
        comp = t[..., d]
        if boundary == "periodic":
            g = (torch.roll(comp, -1, d) - torch.roll(comp, 1, d)) / (2.0 * dxs[d])
        else:
            g = torch.gradient(comp, spacing=dxs[d], dim=d, edge_order=2)[0]
        out += g
    return out


def curl(t: Tensor, dx: Any = 1.0, boundary: str = "neumann") -> Tensor:
    """Curl with torch.gradient one-sided edges; ``boundary`` is ignored (stenpy_engine.py:1936-2032)."""
    if t.ndim < 2 or t.shape[-1] < 2:
        raise ValueError(f"curl: need (*spatial, components>=2), got {tuple(t.shape)}")
    D = t.ndim - 1
    if D not in (2, 3):
        raise ValueError(f"curl: only 2D or 3D vector fields supported, got spatial rank {D}")
    dxs = _dx_list(dx, D)
    if D == 3:
        if t.shape[-1] != 3:
            raise ValueError("curl: 3D vector field must have exactly 3 components")
        g0 = torch.gradient(t[..., 0], spacing=dxs, edge_order=2)
        g1 = torch.gradient(t[..., 1], spacing=dxs, edge_order=2)
        g2 = torch.gradient(t[..., 2], spacing=dxs, edge_order=2)
        return torch.stack([g2[1] - g1[2], g0[2] - g2[0], g1[0] - g0[1]], dim=-1)
    g0 = torch.gradient(t[..., 0], spacing=dxs, edge_order=2)
    g1 = torch.gradient(t[..., 1], spacing=dxs, edge_order=2)
    return g1[0] - g0[1]


def gradient_nd(t: Tensor, dx: Any = 1.0, boundary: str = "neumann") -> Tensor:
    """torch.gradient over every non-singleton axis, stacked last (stenpy_engine.py:2037-2081)."""
    original = t.shape
    ts = t.squeeze() if (1 in t.shape and t.squeeze().ndim > 0) else t
    if ts.ndim == 0:
        return torch.zeros(tuple(original) + (0,), dtype=t.dtype)
    order = 2 if min(ts.shape) >= 3 else 1
    sp = dx if isinstance(dx, (int, float)) else tuple(dx)
    try:
        grads = torch.gradient(ts, spacing=sp, edge_order=order)
    except RuntimeError:
        grads = torch.gradient(ts, spacing=sp, edge_order=1)
    out = torch.stack(grads, dim=-1)
    if out.shape[:-1] != original:
        out = out.reshape(list(original) + [out.shape[-1]])
    return out


def hessian(t: Tensor, dx: Any = 1.0, boundary: str = "neumann") -> Tensor:
    """H[..., i, :] = gradient_nd(gradient_nd(t)[..., i]) (stenpy_engine.py:2086-2128)."""
    D = t.ndim
    dxs = _dx_list(dx, D)
    g = gradient_nd(t, dxs, boundary)
    out = torch.empty(tuple(t.shape) + (D, D), dtype=t.dtype)
    for i in range(D):
        out[..., i, :] = gradient_nd(g[..., i], dxs, boundary)
    return out


def jacobian(t: Tensor, dx: Any = 1.0, boundary: str = "neumann") -> Tensor:
    """J[..., c, d] = d_d t[..., c] (stenpy_engine.py:2133-2175, CPU branch)."""
    if t.ndim != 4 or t.shape[-1] != 3:
        raise ValueError(f"jacobian: need (nx,ny,nz,3), got {tuple(t.shape)}")
    dxs = _dx_list(dx, 3)
    return torch.stack([gradient_nd(t[..., i], dxs, boundary) for i in range(3)], dim=-2)


def mean_curvature(t: Tensor, dx: Any = 1.0, boundary: str = "neumann",
                   eps: float = 1e-15) -> Tensor:
    """divergence(grad/|grad|) with a one-sided gradient_nd (stenpy_engine.py:2180-2195)."""
    dxs = _dx_list(dx, t.ndim)
    g = gradient_nd(t, dxs, boundary)
    mag = torch.linalg.norm(g, dim=-1, keepdim=True).clamp(min=eps)
    return divergence(g / mag, dxs, boundary)


def surface_normals(t: Tensor, dx: Any = 1.0, boundary: str = "neumann") -> Tensor:
    """Unit normals of a 2-D height map (stenpy_engine.py:2199-2219)."""
    if t.ndim != 2:
        raise ValueError(f"surface_normals: expected 2D field, got {t.ndim}D")
    gx, gy = torch.gradient(t, spacing=_dx_list(dx, 2), edge_order=2)
    n = torch.stack([-gx, -gy, torch.ones_like(gx)], dim=-1)
    return n / torch.linalg.norm(n, dim=-1, keepdim=True).clamp(min=1e-15)


def material_derivative(t: Tensor, velocity: Tensor, dx: Any = 1.0,
                        boundary: str = "neumann") -> Tensor:
    """Sum_d v[...,d] * gradient_nd(t)[...,d] (stenpy_engine.py:2223-2243)."""
    g = gradient_nd(t, _dx_list(dx, t.ndim), boundary)
    out = torch.zeros_like(t)
    for d in range(t.ndim):
        out.add_(velocity[..., d] * g[..., d])
    return out


# ----------------------------------------------------------------------------- channel-last helpers
def stack_components(*ts: Tensor, dim: int = -1) -> Tensor:   # 2247-2252
    return torch.stack(list(ts), dim=dim)


def select_component(t: Tensor, index: int) -> Tensor:        # 2255-2259
    return t[..., index].contiguous()


def norm_last(t: Tensor, eps: float = 1e-15) -> Tensor:       # 2262-2266
    return torch.linalg.norm(t, dim=-1, keepdim=True).clamp(min=eps)


def div_last(a: Tensor, b: Tensor) -> Tensor:                 # 2269-2275
    if b.ndim == a.ndim - 1:
        b = b.unsqueeze(-1)
    return a / b


def neg_stack3(gx: Tensor, gy: Tensor) -> Tensor:             # 2278-2289
    n = torch.stack([-gx, -gy, torch.ones_like(gx)], dim=-1)
    return n / torch.linalg.norm(n, dim=-1, keepdim=True).clamp(min=1e-15)


def scale_eye(s: Tensor, D: int) -> Tensor:                   # 2292-2300
    return s.unsqueeze(-1).unsqueeze(-1) * torch.eye(D, dtype=s.dtype)


def velocity_dot_grad(v: Tensor, g: Tensor) -> Tensor:        # 2305-2311
    return torch.mul(v, g)


def dot(u: Tensor, g: Tensor) -> Tensor:
    """Extension: left fold of select_component products (no reference op)."""
    acc = select_component(u, 0) * select_component(g, 0)
    for c in range(1, u.shape[-1]):
        acc = acc + select_component(u, c) * select_component(g, c)
    return acc


def sqnorm(v: Tensor) -> Tensor:
    """Extension: left fold of squared components (no reference op)."""
    c0 = select_component(v, 0)
    acc = c0 * c0
    for c in range(1, v.shape[-1]):
        cc = select_component(v, c)
        acc = acc + cc * cc
    return acc


# ----------------------------------------------------------------------------- reductions
def op_sum(t: Tensor, dim=None, keepdim: bool = False) -> Tensor:   # 2340-2350
    return t.sum() if dim is None else t.sum(dim=dim, keepdim=keepdim)


def mean(t: Tensor, dim=None, keepdim: bool = False) -> Tensor:     # 2353-2366
    return t.mean() if dim is None else t.mean(dim=dim, keepdim=keepdim)


def norm_l2(t: Tensor) -> Tensor:                                   # 2369-2375
    return t.norm()


def min_max(t: Tensor) -> Tensor:                                   # 2378-2382
    return torch.stack([t.min(), t.max()])


def variance(t: Tensor, dim=None, unbiased: bool = True, keepdim: bool = False) -> Tensor:
    # CPU branch, stenpy_engine.py:2570
    if dim is None:
        return torch.var(t, unbiased=unbiased)
    return torch.var(t, dim=dim, unbiased=unbiased, keepdim=keepdim)


def simpson_1d(y: Tensor, dx: float, dim: int) -> Tensor:
    """Composite Simpson; even n: Simpson on n-1 points + last trapezoid (stenpy_engine.py:2415-2433)."""
    n = y.shape[dim]
    if n < 3:
        return torch.trapezoid(y, dx=dx, dim=dim)
    if n % 2 == 0:
        fst, mid, lst = y.select(dim, 0), y.select(dim, n - 2), y.select(dim, n - 1)
        odd_i = torch.arange(1, n - 2, 2)
        evn_i = torch.arange(2, n - 2, 2)
        odd = y.index_select(dim, odd_i).sum(dim) if odd_i.numel() else torch.zeros_like(fst)
        evn = y.index_select(dim, evn_i).sum(dim) if evn_i.numel() else torch.zeros_like(fst)
        return (dx / 3.0) * (fst + mid + 4.0 * odd + 2.0 * evn) + (dx * 0.5) * (mid + lst)
    odd_i = torch.arange(1, n - 1, 2)
    evn_i = torch.arange(2, n - 1, 2)
    fst, lst = y.select(dim, 0), y.select(dim, n - 1)
    odd = y.index_select(dim, odd_i).sum(dim)
    evn = y.index_select(dim, evn_i).sum(dim) if evn_i.numel() else torch.zeros_like(fst)
    return (dx / 3.0) * (fst + lst + 4.0 * odd + 2.0 * evn)


def integrate(t: Tensor, dx: Any = 1.0, dims=None, method: str = "simpson") -> Tensor:
    """Axis-by-axis Simpson/trapezoid, last axis first (stenpy_engine.py:2436-2457)."""
    nd = t.ndim
    dxs = (float(dx),) * nd if isinstance(dx, (int, float)) else tuple(dx)
    seq = tuple(range(nd)) if dims is None else ((dims,) if isinstance(dims, int) else tuple(dims))
    r = t
    for d in sorted(seq, reverse=True):
        r = simpson_1d(r, dxs[d], d) if method == "simpson" else torch.trapezoid(r, dx=dxs[d], dim=d)
    return r


def moving_average(t: Tensor, window: int = 3, boundary: str = "neumann") -> Tensor:
    """Separable box filter of width ``window`` per axis (stenpy_engine.py:2385-2408)."""
    if t.ndim > 3:
        raise NotImplementedError("moving_average: only 1-D/2-D/3-D supported")
    pad = window // 2
    w = torch.full((1, 1, window), 1.0 / window, dtype=t.dtype)
    out = t
    for d in range(t.ndim):
        p = pad1d(out, d, pad, boundary)
        perm = list(range(out.ndim))
        perm[d], perm[-1] = perm[-1], perm[d]
        x = p.permute(perm).contiguous()
        B = math.prod(x.shape[:-1]) if x.ndim > 1 else 1
        y = F.conv1d(x.reshape(B, 1, x.shape[-1]), w)
        inv = [0] * out.ndim
        for s, dd in enumerate(perm):
            inv[dd] = s
        out = y.reshape(*x.shape[:-1], out.shape[d]).permute(inv).contiguous()
    return out


def conv3d_spectral(a: Tensor, boundary: str = "periodic", dx: Any = None) -> Tensor:
    """FFT circular convolution with the 27-point (1, 2, 1)^3 / 64 kernel centred at the
    origin (stenpy_engine.py:2823-2846, the CPU path: the boundary argument has no effect)."""
    k = torch.tensor([[[1, 2, 1], [2, 4, 2], [1, 2, 1]],
                      [[2, 4, 2], [4, 8, 4], [2, 4, 2]],
                      [[1, 2, 1], [2, 4, 2], [1, 2, 1]]], dtype=a.dtype) / 64.0
    k_full = torch.zeros(a.shape, dtype=a.dtype)
    k_full[:3, :3, :3] = k
    k_full = torch.roll(k_full, shifts=(-1, -1, -1), dims=(0, 1, 2))
    c = torch.fft.ifftn(torch.fft.fftn(a) * torch.fft.fftn(k_full))
    return c.real.to(a.dtype)


def cumulative_integral(t: Tensor, dx: float = 1.0, dim: int = 0) -> Tensor:   # 2462-2483
    n = t.shape[dim]
    if n < 2:
        return torch.zeros_like(t)
    a = t.narrow(dim, 0, n - 1)
    b = t.narrow(dim, 1, n - 1)
    trapz = (a + b) * (dx * 0.5)
    zshape = list(t.shape)
    zshape[dim] = 1
    return torch.cat([torch.zeros(zshape, dtype=t.dtype), torch.cumsum(trapz, dim=dim)], dim=dim)


# ----------------------------------------------------------------------------- tensor ops
def trace(t: Tensor) -> Tensor:                                   # 2674-2680
    return torch.diagonal(t, dim1=-2, dim2=-1).sum(-1)


def determinant(t: Tensor) -> Tensor:
    """Global +1e-12 I regularisation when any |diag| < 1e-12 (stenpy_engine.py:2683-2712)."""
    D = t.shape[-1]
    eye = torch.eye(D, dtype=t.dtype)
    diag = torch.diagonal(t, dim1=-2, dim2=-1)
    t_reg = t + 1e-12 * eye if bool((diag.abs() < 1e-12).any()) else t
    d = torch.linalg.det(t_reg)
    d = torch.nan_to_num(d, nan=0.0, posinf=1e30, neginf=-1e30)
    return torch.clamp(d, min=-1e30, max=1e30)


def eigenvalues(t: Tensor) -> Tensor:                             # 2715-2721
    return torch.linalg.eigvalsh(t)


def inverse(t: Tensor) -> Tensor:                                 # 2724-2730
    return torch.linalg.inv(t)


def deviatoric(t: Tensor) -> Tensor:                              # 2852-2867
    D = t.shape[-1]
    tr = torch.diagonal(t, dim1=-2, dim2=-1).sum(-1, keepdim=True).unsqueeze(-1)
    r = t - (tr / max(D, 1e-15)) * torch.eye(D, dtype=t.dtype)
    return torch.nan_to_num(r, nan=0.0, posinf=0.0, neginf=0.0)


# ----------------------------------------------------------------------------- spectral
# The reference works in FP64 / complex128; with FP32 input the same code is the FP32 twin
# (complex64).  Wave numbers: torch.fft.fftfreq in torch's default dtype (FP32), widened to the
# working dtype and scaled by 2 pi (stenpy_engine.py:2931-2964).
def _complex_of(t: Tensor) -> torch.dtype:
    if torch.is_complex(t):
        return t.dtype
    return torch.complex128 if t.dtype == torch.float64 else torch.complex64


def _real_of(t: Tensor) -> torch.dtype:
    return torch.float64 if t.dtype in (torch.float64, torch.complex128) else torch.float32


def fft(t: Tensor, norm: str = "backward") -> Tensor:              # 2875-2885
    if not torch.is_complex(t):
        t = t.to(_complex_of(t))
    return torch.fft.fftn(t, norm=norm)


def ifft(t: Tensor, norm: str = "backward") -> Tensor:             # 2887-2896
    if not torch.is_complex(t):
        t = torch.complex(t, torch.zeros_like(t))
    return torch.fft.ifftn(t, norm=norm).real.to(_real_of(t))


def fft_complex(t: Tensor, norm: str = "backward") -> Tensor:      # 2901-2908
    if norm == "numpy_compatible":
        norm = "backward"
    return torch.fft.fftn(t, norm=norm)


def ifft_complex(t: Tensor, norm: str = "backward") -> Tensor:     # 2910-2918
    if not torch.is_complex(t):
        t = torch.complex(t, torch.zeros_like(t))
    if norm == "numpy_compatible":
        norm = "backward"
    return torch.fft.ifftn(t, norm=norm)


def _k_grids(shape, spacing, dtype):                                # 2931-2964
    grids = []
    for d, (n, dx) in enumerate(zip(shape, spacing)):
        k = torch.fft.fftfreq(n, d=dx).to(dtype) * (2 * math.pi)
        sh = [1] * len(shape)
        sh[d] = -1
        grids.append(k.view(sh))
    return grids, sum(kg ** 2 for kg in grids)


def spectral_gradient(t: Tensor, dx: Any = 1.0, dim: int = 0, norm: str = "backward") -> Tensor:
    # 2979-2992
    dxs = (float(dx),) * t.ndim if isinstance(dx, (int, float)) else tuple(dx)
    grids, _ = _k_grids(t.shape, dxs, t.dtype)
    T_hat = torch.fft.fftn(t, norm=norm)
    return torch.fft.ifftn(1j * grids[dim] * T_hat, norm=norm).real.to(t.dtype)


def spectral_laplacian(t: Tensor, dx: Any = 1.0, norm: str = "backward") -> Tensor:   # 2997-3010
    dxs = (float(dx),) * t.ndim if isinstance(dx, (int, float)) else tuple(dx)
    _, k2 = _k_grids(t.shape, dxs, t.dtype)
    T_hat = torch.fft.fftn(t, norm=norm)
    return torch.fft.ifftn(-k2 * T_hat, norm=norm).real.to(t.dtype)


# ----------------------------------------------------------------------------- statistics
def surface_integral(t: Tensor, normals: Tensor, area_weights: Tensor) -> Tensor:   # 2486-2496
    return ((t * normals).sum(dim=-1) * area_weights).sum()


def covariance(a: Tensor, b: Tensor) -> Tensor:                    # 2575-2592, CPU branch
    af = (a - a.mean()).flatten()
    bf = (b - b.mean()).flatten()
    return torch.dot(af, bf) / max(af.numel() - 1, 1)


def correlation(a: Tensor, b: Tensor) -> Tensor:                   # 2595-2615, CPU branch
    af = (a - a.mean()).flatten()
    bf = (b - b.mean()).flatten()
    return torch.dot(af, bf) / (af.norm() * bf.norm() + 1e-15)


def entropy(t: Tensor, dim=None, eps: float = 1e-15) -> Tensor:   # 2638-2669, CPU branch
    p = t.abs()
    if dim is None:
        p = p / (p.sum() + eps)
        return -(p * (p + eps).log()).sum()
    p = p / (p.sum(dim=dim, keepdim=True) + eps)
    return -(p * (p + eps).log()).sum(dim=dim)


# ----------------------------------------------------------------------------- geometry
def distance_transform(t: Tensor, dx: Any = 1.0) -> Tensor:           # 3115-3156, CPU branch
    from scipy.ndimage import distance_transform_edt
    dxs = (float(dx),) * t.ndim if isinstance(dx, (int, float)) else tuple(dx)
    out = distance_transform_edt((t == 0).numpy(), sampling=list(dxs))
    return torch.from_numpy(out.astype("float64")).to(t.dtype)


OPS = {
    "add": add, "sub": sub, "mul": mul, "div": div, "neg": neg, "clamp": clamp, "exp": exp,
    "log": log, "sqrt": sqrt, "sin": sin, "cos": cos, "tanh": tanh,
    "laplacian": laplacian, "gradient": gradient, "divergence": divergence, "curl": curl,
    "gradient_nd": gradient_nd, "hessian": hessian, "jacobian": jacobian,
    "mean_curvature": mean_curvature, "surface_normals": surface_normals,
    "material_derivative": material_derivative,
    "stack_components": stack_components, "select_component": select_component,
    "norm_last": norm_last, "div_last": div_last, "neg_stack3": neg_stack3,
    "scale_eye": scale_eye, "velocity_dot_grad": velocity_dot_grad,
    "dot": dot, "sqnorm": sqnorm,
    "sum": op_sum, "mean": mean, "norm_l2": norm_l2, "min_max": min_max,
    "variance": variance, "integrate": integrate, "moving_average": moving_average, "conv3d_spectral": conv3d_spectral,
    "cumulative_integral": cumulative_integral,
    "trace": trace, "determinant": determinant, "eigenvalues": eigenvalues,
    "inverse": inverse, "deviatoric": deviatoric,
    "fft": fft, "ifft": ifft, "fft_complex": fft_complex, "ifft_complex": ifft_complex,
    "spectral_gradient": spectral_gradient, "spectral_laplacian": spectral_laplacian,
    "distance_transform": distance_transform,
    "surface_integral": surface_integral, "covariance": covariance, "correlation": correlation,
    "entropy": entropy,
}

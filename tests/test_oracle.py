"""Pin the CPU oracle (oracle/ops.py) against closed forms, library routines and brute force.

Every check here is independent of the oracle's own code: discrete Fourier symbols of single
modes, exact derivatives/integrals of low-order polynomials, numpy.linalg / math.fsum, and
triple loops written from the reference formulas on tiny grids.  Each check is chosen so that
a plausible slip (wrong sign, index or coefficient, swapped axis) fails it.
"""
import math

import numpy as np
import pytest
import torch

from oracle import ops as O

D64 = torch.float64
RNG = np.random.default_rng(1234)


def mode(shape, k, phi=0.3):
    """cos(2 pi sum k_d i_d / n_d + phi) on the index grid."""
    idx = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    ph = sum(2 * math.pi * kd * i / n for kd, i, n in zip(k, idx, shape)) + phi
    return torch.from_numpy(np.cos(ph)), torch.from_numpy(np.sin(ph))


def rand(shape):
    return torch.from_numpy(RNG.standard_normal(shape))


# ------------------------------------------------------------------ brute-force formulas
def bf_tg(a, axis, i, h, order):
    """torch.gradient at index tuple i along axis (formulas of the documented edge cases)."""
    n = a.shape[axis]
    def at(d):
        j = list(i); j[axis] += d
        return float(a[tuple(j)])
    p = i[axis]
    if order == 2 and p == 0:
        return (-1.5 * at(0) + 2 * at(1) - 0.5 * at(2)) / h
    if order == 2 and p == n - 1:
        return (0.5 * at(-2) - 2 * at(-1) + 1.5 * at(0)) / h
    if order == 1 and p == 0:
        return (at(1) - at(0)) / h
    if order == 1 and p == n - 1:
        return (at(0) - at(-1)) / h
    return (at(1) - at(-1)) / (2 * h)


def bf_grid(fn, shape):
    out = np.empty(shape)
    for i in np.ndindex(*shape):
        out[i] = fn(i)
    return torch.from_numpy(out)


# ------------------------------------------------------------------ periodic symbols
@pytest.mark.parametrize("k", [(1, 2, 3), (3, 0, 5)])
def test_periodic_laplacian_symbol(k):
    shape, dx = (8, 10, 12), (0.5, 1.0, 2.0)
    c, _ = mode(shape, k)
    sym = -sum((2 - 2 * math.cos(2 * math.pi * kd / n)) / h ** 2 for kd, n, h in zip(k, shape, dx))
    torch.testing.assert_close(O.laplacian(c, dx, "periodic"), sym * c, rtol=0, atol=1e-12)


def test_periodic_gradient_divergence_symbol():
    shape, k, h = (8, 10, 12), (1, 2, 3), 0.7
    c, s = mode(shape, k)
    g = O.gradient(c, h, "periodic")
    for d in range(3):
        exp = -math.sin(2 * math.pi * k[d] / shape[d]) / h * s
        torch.testing.assert_close(g[..., d], exp, rtol=0, atol=1e-12)
    div = O.divergence(torch.stack([c, c, c], -1), (h, h, h), "periodic")
    exp = -sum(math.sin(2 * math.pi * k[d] / shape[d]) for d in range(3)) / h * s
    torch.testing.assert_close(div, exp, rtol=0, atol=1e-12)


def test_wide_hessian_interior_symbol():
    """Interior diagonal of the CPU hessian is the wide (f[+2]-2f+f[-2])/(4h^2) stencil."""
    shape, k = (12, 12, 12), (1, 2, 1)
    c, _ = mode(shape, k)
    H = O.hessian(c, 1.0)
    inner = (slice(2, -2),) * 3
    for d in range(3):
        sym = -math.sin(2 * math.pi * k[d] / shape[d]) ** 2
        torch.testing.assert_close(H[..., d, d][inner], (sym * c)[inner], rtol=0, atol=1e-12)


# ------------------------------------------------------------------ exact polynomials
def quad_field(shape, dx):
    idx = np.meshgrid(*[np.arange(n) * h for n, h in zip(shape, dx)], indexing="ij")
    x, y, z = idx
    f = 1.5 - 0.5 * x + 2 * y + 0.25 * z + 0.3 * x * x - 0.2 * y * z + 0.7 * z * z
    gx, gy, gz = -0.5 + 0.6 * x, 2 - 0.2 * z, 0.25 - 0.2 * y + 1.4 * z
    return [torch.from_numpy(a) for a in (f, gx, gy, gz)]


def test_one_sided_gradients_exact_on_quadratics():
    shape, dx = (5, 6, 7), (0.5, 1.0, 2.0)
    f, gx, gy, gz = quad_field(shape, dx)
    g = O.gradient_nd(f, dx)
    for d, ex in enumerate((gx, gy, gz)):
        torch.testing.assert_close(g[..., d], ex, rtol=0, atol=1e-12)
    for d, ex in enumerate((gx, gy, gz)):
        torch.testing.assert_close(O.gradient(f, dx[d], "neumann", dim=d), ex, rtol=0, atol=1e-12)


def test_curl_exact_on_linear_field():
    shape = (5, 6, 7)
    x, y, z = [torch.from_numpy(a) for a in np.meshgrid(*[np.arange(n) * 1.0 for n in shape], indexing="ij")]
    v = torch.stack([2 * y - z, 3 * z + x, -x + 0.5 * y], -1)
    c = O.curl(v, 1.0, "periodic")  # boundary is ignored: one-sided edges stay exact
    exp = torch.stack([torch.full_like(x, 0.5 - 3), torch.full_like(x, -1 + 1.0), torch.full_like(x, 1 - 2.0)], -1)
    torch.testing.assert_close(c, exp, rtol=0, atol=1e-12)


def test_hessian_and_tensor_ops_on_integer_quadratic():
    shape = (6, 7, 8)
    x, y, z = np.meshgrid(*[np.arange(n) * 1.0 for n in shape], indexing="ij")
    f = torch.from_numpy(x * x + y * y + z * z)
    H = O.hessian(f, 1.0)
    eye = torch.eye(3, dtype=D64).expand_as(H)
    assert torch.equal(H, 2 * eye)
    assert torch.equal(O.trace(H), torch.full(shape, 6.0, dtype=D64))
    torch.testing.assert_close(O.determinant(H), torch.full(shape, 8.0, dtype=D64))
    torch.testing.assert_close(O.eigenvalues(H), torch.full(shape + (3,), 2.0, dtype=D64))
    assert torch.equal(O.deviatoric(H), torch.zeros_like(H))


# ------------------------------------------------------------------ brute force
@pytest.mark.parametrize("bc", ["periodic", "neumann", "dirichlet", "reflect", "other"])
def test_laplacian_brute_force(bc):
    shape, dx = (4, 5, 6), (1.0, 0.5, 2.0)
    a = rand(shape)

    def val(i, d, s):
        n = shape[d]
        j = list(i); q = i[d] + s
        if 0 <= q < n:
            j[d] = q
        elif bc == "periodic":
            j[d] = q % n
        elif bc == "neumann":
            j[d] = min(max(q, 0), n - 1)
        elif bc == "reflect":
            j[d] = 1 if q < 0 else n - 2
        else:
            return 0.0
        return float(a[tuple(j)])

    exp = bf_grid(lambda i: sum((val(i, d, 1) + val(i, d, -1) - 2 * float(a[i])) / dx[d] ** 2
                                for d in range(3)), shape)
    torch.testing.assert_close(O.laplacian(a, dx, bc), exp, rtol=1e-13, atol=1e-12)


def test_stencils_brute_force():
    shape, h = (4, 5, 6), 0.5
    a, v = rand(shape), rand(shape + (3,))
    order = 2
    g = O.gradient_nd(a, h)
    exp = torch.stack([bf_grid(lambda i, d=d: bf_tg(a, d, i, h, order), shape) for d in range(3)], -1)
    torch.testing.assert_close(g, exp, rtol=1e-12, atol=1e-12)
    # hessian H[i][j] = D_j (D_i f)
    gi = [bf_grid(lambda i, d=d: bf_tg(a, d, i, h, order), shape) for d in range(3)]
    H = O.hessian(a, h)
    for r in range(3):
        for c in range(3):
            ex = bf_grid(lambda i: bf_tg(gi[r], c, i, h, order), shape)
            torch.testing.assert_close(H[..., r, c], ex, rtol=1e-12, atol=1e-11)
    # curl with one-sided edges
    dv = [[bf_grid(lambda i, a_=ax, c_=comp: bf_tg(v[..., c_], a_, i, h, 2), shape) for ax in range(3)]
          for comp in range(3)]
    cu = O.curl(v, h)
    torch.testing.assert_close(cu[..., 0], dv[2][1] - dv[1][2], rtol=1e-12, atol=1e-11)
    torch.testing.assert_close(cu[..., 1], dv[0][2] - dv[2][0], rtol=1e-12, atol=1e-11)
    torch.testing.assert_close(cu[..., 2], dv[1][0] - dv[0][1], rtol=1e-12, atol=1e-11)
    # divergence, non-periodic and periodic
    ex = dv[0][0] + dv[1][1] + dv[2][2]
    torch.testing.assert_close(O.divergence(v, h, "neumann"), ex, rtol=1e-12, atol=1e-11)
    per = bf_grid(lambda i: sum((float(v[tuple((i[k] + (1 if k == d else 0)) % shape[k] for k in range(3)) + (d,)])
                                 - float(v[tuple((i[k] - (1 if k == d else 0)) % shape[k] for k in range(3)) + (d,)]))
                                / (2 * h) for d in range(3)), shape)
    torch.testing.assert_close(O.divergence(v, h, "periodic"), per, rtol=1e-12, atol=1e-11)
    # jacobian J[c][d] = D_d v_c
    J = O.jacobian(v, h)
    for comp in range(3):
        for ax in range(3):
            torch.testing.assert_close(J[..., comp, ax], dv[comp][ax], rtol=1e-12, atol=1e-11)
    # material derivative
    md = sum(v[..., d] * gi[d] for d in range(3))
    torch.testing.assert_close(O.material_derivative(a, v, h), md, rtol=1e-12, atol=1e-11)


@pytest.mark.parametrize("bc", ["periodic", "neumann"])
def test_mean_curvature_brute_force(bc):
    shape, h, eps = (5, 6, 7), 1.0, 1e-15
    a = rand(shape)
    g = [bf_grid(lambda i, d=d: bf_tg(a, d, i, h, 2), shape) for d in range(3)]
    m = torch.sqrt(g[0] ** 2 + g[1] ** 2 + g[2] ** 2).clamp(min=eps)
    nv = [gd / m for gd in g]
    if bc == "periodic":
        ex = sum((torch.roll(nv[d], -1, d) - torch.roll(nv[d], 1, d)) / (2 * h) for d in range(3))
    else:
        ex = sum(bf_grid(lambda i, d=d: bf_tg(nv[d], d, i, h, 2), shape) for d in range(3))
    torch.testing.assert_close(O.mean_curvature(a, h, bc), ex, rtol=1e-12, atol=1e-11)


def test_gradient_periodic_and_short_axes():
    a = rand((2, 5, 3))
    g = O.gradient(a, 1.0, "neumann")
    # axis of size 2 uses first-order one-sided differences on both cells
    torch.testing.assert_close(g[0, ..., 0], a[1] - a[0])
    torch.testing.assert_close(g[1, ..., 0], a[1] - a[0])
    assert torch.equal(O.gradient(a[:1], 1.0, "neumann", dim=0), torch.zeros_like(a[:1]))


# ------------------------------------------------------------------ reductions
def test_reductions_against_fsum():
    a = rand((7, 8, 9)) + 3.0
    v = a.flatten().tolist()
    n = len(v)
    mu = math.fsum(v) / n
    assert abs(O.op_sum(a).item() - math.fsum(v)) < 1e-10
    assert abs(O.mean(a).item() - mu) < 1e-13
    assert abs(O.norm_l2(a).item() - math.sqrt(math.fsum(x * x for x in v))) < 1e-11
    var = math.fsum((x - mu) ** 2 for x in v) / (n - 1)
    assert abs(O.variance(a).item() - var) < 1e-12
    assert O.min_max(a).tolist() == [min(v), max(v)]


@pytest.mark.parametrize("n", [2, 3, 4, 7, 8])
def test_simpson_weights_and_exactness(n):
    dx = 0.3
    # exactness: composite Simpson integrates cubics exactly for odd n; the even-n tail is a
    # trapezoid, exact for linear functions
    xs = np.arange(n) * dx
    if n % 2 == 1 and n >= 3:
        p = lambda x: 1 - 2 * x + 0.5 * x ** 2 + 0.25 * x ** 3
        P = lambda x: x - x ** 2 + x ** 3 / 6 + x ** 4 / 16
    else:
        p = lambda x: 2 - 3 * x
        P = lambda x: 2 * x - 1.5 * x ** 2
    y = torch.from_numpy(p(xs))
    exact = P(xs[-1]) - P(xs[0])
    assert abs(O.simpson_1d(y, dx, 0).item() - exact) < 1e-12
    # brute-force weights for random data
    r = rand((n,))
    if n < 3:
        w = np.array([0.5, 0.5])[:n] * dx
    else:
        m = n if n % 2 else n - 1
        w = np.array([1] + [4 if i % 2 else 2 for i in range(1, m - 1)] + [1], float) * dx / 3
        if n % 2 == 0:
            w = np.append(w, 0.0)
            w[-2:] += 0.5 * dx
    assert abs(O.simpson_1d(r, dx, 0).item() - float((r.numpy() * w).sum())) < 1e-12


def test_integrate_separable():
    a = rand((5, 6, 7))
    dx = (0.5, 1.0, 0.25)

    def w(n, h):
        if n < 3:
            return np.array([0.5, 0.5])[:n] * h
        m = n if n % 2 else n - 1
        ww = np.array([1] + [4 if i % 2 else 2 for i in range(1, m - 1)] + [1], float) * h / 3
        if n % 2 == 0:
            ww = np.append(ww, 0.0)
            ww[-2:] += 0.5 * h
        return ww

    W = np.einsum("i,j,k->ijk", w(5, dx[0]), w(6, dx[1]), w(7, dx[2]))
    assert abs(O.integrate(a, dx).item() - float((a.numpy() * W).sum())) < 1e-12
    tr = np.einsum("i,j,k->ijk", *[np.r_[0.5, np.ones(n - 2), 0.5] * h for n, h in zip((5, 6, 7), dx)])
    assert abs(O.integrate(a, dx, method="trapezoid").item() - float((a.numpy() * tr).sum())) < 1e-12


# ------------------------------------------------------------------ tensor ops
def test_tensor_ops_against_numpy():
    M = rand((50, 3, 3))
    S = 0.5 * (M + M.transpose(1, 2))
    np.testing.assert_allclose(O.eigenvalues(S).numpy(), np.linalg.eigvalsh(S.numpy()), atol=1e-12)
    np.testing.assert_allclose(O.determinant(M).numpy(), np.linalg.det(M.numpy()), atol=1e-12)
    np.testing.assert_allclose(O.inverse(M).numpy(), np.linalg.inv(M.numpy()), atol=1e-9)
    np.testing.assert_allclose(O.trace(M).numpy(), np.trace(M.numpy(), axis1=1, axis2=2))
    tr = np.trace(M.numpy(), axis1=1, axis2=2)[:, None, None]
    np.testing.assert_allclose(O.deviatoric(M).numpy(), M.numpy() - tr / 3 * np.eye(3), atol=1e-14)


def test_determinant_global_regularisation():
    M = rand((4, 3, 3))
    M[2, 1, 1] = 0.0
    reg = np.linalg.det(M.numpy() + 1e-12 * np.eye(3))
    np.testing.assert_allclose(O.determinant(M).numpy(), reg, rtol=0, atol=1e-15)


# ------------------------------------------------------------------ point-wise guards
def test_pointwise_guards():
    a = torch.tensor([1.0, -2.0, 3.0, 4.0], dtype=D64)
    b = torch.tensor([0.0, -1e-16, 1e-16, 2.0], dtype=D64)
    assert O.div(a, b).tolist() == [1.0 / 1e-15, -2.0 / -1e-15, 3.0 / 1e-15, 2.0]
    x = torch.tensor([-1.0, 0.0, 1e-20, 4.0], dtype=D64)
    assert O.log(x).tolist()[:3] == [math.log(1e-15)] * 3
    assert O.sqrt(x).tolist() == [0.0, 0.0, 1e-10, 2.0]
    assert O.clamp(torch.tensor([-1.0, 0.5, 2.0], dtype=D64)).tolist() == [0.0, 0.5, 1.0]


def test_pad_matches_reference_fallback_loops():
    """pad1d equals the reference's manual padding loops (stenpy_engine.py:1393-1416)."""
    def fallback(t, dim, r, mode):
        shape = list(t.shape); shape[dim] += 2 * r
        p = torch.zeros(shape, dtype=t.dtype)
        sl = [slice(None)] * t.ndim; sl[dim] = slice(r, r + t.shape[dim]); p[tuple(sl)] = t
        n = t.shape[dim]
        def put(pos, src):
            s = [slice(None)] * t.ndim; s[dim] = pos; p[tuple(s)] = t.select(dim, src)
        if mode == "replicate":
            for i in range(r): put(i, 0); put(-(i + 1), n - 1)
        elif mode == "circular":
            for i in range(r): put(i, n - (r - i)); put(-(i + 1), i)
        elif mode == "reflect" and n > 1:
            for i in range(r): put(r - 1 - i, min(i + 1, n - 1)); put(-(i + 1), max(n - 2 - i, 0))
        return p
    modes = {"neumann": "replicate", "periodic": "circular", "reflect": "reflect", "dirichlet": "constant"}
    for shape in [(5, 6, 7), (1, 2, 3), (4, 5, 6, 3)]:
        t = rand(shape)
        for bc, m in modes.items():
            for d in range(t.ndim):
                for r in (1, 2):
                    if m == "circular" and t.shape[d] < r:
                        continue
                    assert torch.equal(O.pad1d(t, d, r, bc), fallback(t, d, r, m)), (shape, bc, d, r)


def test_fp32_twin_tracks_fp64():
    """The FP32 twin computes the same formulas: its error is at FP32 rounding level."""
    a = rand((8, 9, 10))
    for fn in (lambda t: O.laplacian(t, 1.0, "periodic"), lambda t: O.curl(torch.stack([t, 2 * t, -t], -1)),
               lambda t: O.hessian(t, 0.5)):
        r64, r32 = fn(a), fn(a.float())
        assert r32.dtype == torch.float32
        err = (r32.double() - r64).abs().max().item()
        assert 0 < err < 1e-5 * max(1.0, r64.abs().max().item())


# ------------------------------------------------------------------ separable filters
@pytest.mark.parametrize("bc", ["periodic", "neumann", "dirichlet"])
@pytest.mark.parametrize("window", [1, 3, 5])
def test_moving_average_brute_force(bc, window):
    """Separable box passes (axis 0, then 1, then 2) with _pad1d's fallback-loop padding
    (whose circular right side is mirrored for radius >= 2), as nested loops."""
    shape = (5, 6, 7)
    a = rand(shape)
    h = window // 2

    def pad_map(n, r):
        """Padded position -> source index, filled exactly as the reference's fallback loops
        do (stenpy_engine.py:1395-1413); None = zero."""
        m = [None] * r + list(range(n)) + [None] * r
        for i in range(r):
            if bc == "neumann":
                m[i], m[-(i + 1)] = 0, n - 1
            elif bc == "periodic":
                m[i], m[-(i + 1)] = n - (r - i), i
        return m

    def padded(arr, d, i):
        n = arr.shape[d]
        m = pad_map(n, h)
        src = m[i[d] + h]
        if src is None:
            return 0.0
        j = list(i)
        j[d] = src
        return float(arr[tuple(j)])

    cur = a.clone()
    for d in range(3):
        nxt = torch.zeros_like(cur)
        for i in np.ndindex(*shape):
            acc = 0.0
            for k in range(-h, h + 1):
                j = list(i)
                j[d] += k
                acc += padded(cur, d, tuple(j)) / window
            nxt[i] = acc
        cur = nxt
    torch.testing.assert_close(O.moving_average(a, window, bc), cur, rtol=1e-13, atol=1e-13)
    # a constant is a fixed point except where zero padding enters
    if bc != "dirichlet":
        c = torch.full(shape, 2.5, dtype=torch.float64)
        torch.testing.assert_close(O.moving_average(c, window, bc), c, rtol=1e-14, atol=1e-14)


def test_conv3d_spectral_is_the_periodic_27_point_blur():
    shape = (4, 5, 6)
    a = rand(shape)
    w = {-1: 0.25, 0: 0.5, 1: 0.25}
    exp = bf_grid(lambda i: sum(w[p] * w[q] * w[r] * float(a[(i[0] + p) % 4, (i[1] + q) % 5, (i[2] + r) % 6])
                                for p in (-1, 0, 1) for q in (-1, 0, 1) for r in (-1, 0, 1)), shape)
    torch.testing.assert_close(O.conv3d_spectral(a, boundary="neumann"), exp, rtol=1e-12, atol=1e-12)
    # Fourier symbol of a single mode: prod_d (1 + cos theta_d) / 2
    n = (8, 10, 12)
    k = (1, 3, 2)
    z, y, x = torch.meshgrid(*[torch.arange(m, dtype=torch.float64) for m in n], indexing="ij")
    th = [2 * math.pi * k[d] / n[d] for d in range(3)]
    f = torch.cos(th[0] * z + th[1] * y + th[2] * x)
    sym = math.prod((1 + math.cos(t)) / 2 for t in th)
    torch.testing.assert_close(O.conv3d_spectral(f), sym * f, rtol=1e-12, atol=1e-12)


# ----------------------------------------------------------------------------- spectral
def test_fft_oracle_matches_numpy_fft_for_every_norm():
    rng = np.random.default_rng(5)
    t = torch.from_numpy(rng.standard_normal((6, 7, 10)))
    for norm in ("backward", "ortho", "forward"):
        ref = np.fft.fftn(t.numpy(), norm=norm)
        assert np.allclose(O.fft(t, norm=norm).numpy(), ref, rtol=0, atol=1e-12)
        back = O.ifft(torch.from_numpy(ref), norm=norm)
        assert back.dtype == torch.float64 and torch.allclose(back, t, atol=1e-12)
        assert np.allclose(O.ifft_complex(torch.from_numpy(ref), norm=norm).numpy(),
                           np.fft.ifftn(ref, norm=norm), atol=1e-12)
    assert torch.equal(O.fft_complex(t.to(torch.complex128), norm="numpy_compatible"), O.fft(t))


def test_wave_numbers_are_fp32_fftfreq_widened():
    # the device kernel forms k = 2 pi * float(j) * float(1 / (n dx)); the oracle uses
    # torch.fft.fftfreq itself: both must agree bit for bit
    for n in (1, 2, 7, 8, 33, 64, 255, 256):
        for dx in (1.0, 0.5, 0.3, 2.0 / 3):
            grids, _ = O._k_grids((n,), (dx,), torch.float64)
            j = np.arange(n)
            j = np.where(j < (n + 1) // 2, j, j - n).astype(np.float32)
            f = (j * np.float32(1.0 / (n * dx))).astype(np.float32)
            k = f.astype(np.float64) * (2 * math.pi)
            assert np.array_equal(grids[0].numpy(), k), (n, dx)


def test_spectral_derivatives_of_fourier_modes_are_exact():
    # below the Nyquist frequency spectral differentiation is exact: d/dx cos(2 pi m x / L)
    shape, dxs = (8, 12, 10), (0.5, 1.0, 0.25)
    idx = torch.meshgrid(*[torch.arange(n, dtype=torch.float64) for n in shape], indexing="ij")
    modes = (1, 3, 2)
    ph = [2 * math.pi * m * i / n for m, i, n in zip(modes, idx, shape)]
    f = torch.cos(ph[0]) + torch.cos(ph[1] + 0.3) + torch.sin(ph[2])
    om = [2 * math.pi * m / (n * d) for m, n, d in zip(modes, shape, dxs)]
    exact = [-om[0] * torch.sin(ph[0]), -om[1] * torch.sin(ph[1] + 0.3), om[2] * torch.cos(ph[2])]
    for d in range(3):
        assert torch.allclose(O.spectral_gradient(f, dx=dxs, dim=d), exact[d], atol=1e-12), d
    lap = -(om[0] ** 2) * torch.cos(ph[0]) - om[1] ** 2 * torch.cos(ph[1] + 0.3) - om[2] ** 2 * torch.sin(ph[2])
    assert torch.allclose(O.spectral_laplacian(f, dx=dxs), lap, atol=1e-11)
    # the Nyquist mode cos(pi i) has a purely imaginary derivative spectrum: its real part is 0
    g = torch.cos(math.pi * idx[0])
    assert O.spectral_gradient(g, dim=0).abs().max().item() < 1e-12


# ----------------------------------------------------------------------------- geometry
def test_distance_transform_oracle_matches_brute_force():
    rng = np.random.default_rng(11)
    for trial in range(12):
        shape = tuple(int(v) for v in rng.integers(1, 7, 3))
        t = torch.from_numpy((rng.random(shape) < 0.3) * rng.standard_normal(shape))
        if not (t != 0).any():
            t.view(-1)[0] = 1.0
        s = tuple(float(v) for v in rng.choice([0.5, 1.0, 2.0, 0.3], 3))
        got = O.distance_transform(t, dx=s)
        bg = np.argwhere(t.numpy() != 0)
        want = np.zeros(shape)
        for idx in np.ndindex(*shape):
            if t[idx] == 0:
                want[idx] = np.sqrt((((bg - np.array(idx)) * np.array(s)) ** 2).sum(1)).min()
        assert np.allclose(got.numpy(), want, rtol=0, atol=1e-12), (shape, s)
    # one background voxel: the exact Euclidean distance to it, and 0 on it
    t = torch.zeros((5, 6, 7), dtype=torch.float64)
    t[2, 3, 4] = -1.5
    i, j, k = torch.meshgrid(*[torch.arange(n, dtype=torch.float64) for n in t.shape], indexing="ij")
    want = torch.sqrt(((i - 2) * 0.5) ** 2 + (j - 3) ** 2 + ((k - 4) * 2.0) ** 2)
    assert torch.allclose(O.distance_transform(t, dx=(0.5, 1.0, 2.0)), want, atol=1e-12)


# ----------------------------------------------------------------------------- statistics
def test_statistics_oracle_against_numpy_and_scipy():
    from scipy.stats import entropy as sp_entropy
    rng = np.random.default_rng(12)
    a = torch.from_numpy(rng.standard_normal((5, 6, 7)))
    b = torch.from_numpy(0.3 * a.numpy() + rng.standard_normal((5, 6, 7)))
    assert math.isclose(O.covariance(a, b).item(), np.cov(a.numpy().ravel(), b.numpy().ravel())[0, 1], rel_tol=1e-12)
    assert math.isclose(O.correlation(a, b).item(), np.corrcoef(a.numpy().ravel(), b.numpy().ravel())[0, 1],
                        rel_tol=1e-12)
    assert math.isclose(O.entropy(a).item(), sp_entropy(np.abs(a.numpy().ravel())), rel_tol=1e-12)
    per = O.entropy(a, dim=1).numpy()
    want = np.apply_along_axis(lambda v: sp_entropy(np.abs(v)), 1, a.numpy())
    assert np.allclose(per, want, rtol=1e-12)
    t, n, w = (torch.from_numpy(rng.standard_normal(s)) for s in ((4, 5, 3), (4, 5, 3), (4, 5)))
    brute = sum(w[i, j].item() * sum(t[i, j, c].item() * n[i, j, c].item() for c in range(3))
                for i in range(4) for j in range(5))
    assert math.isclose(O.surface_integral(t, n, w).item(), brute, rel_tol=1e-12)
    # cumulative trapezoid: its last entry is numpy's trapezoid rule over the axis
    assert np.allclose(O.cumulative_integral(a, dx=0.5, dim=2).numpy()[..., -1],
                       np.trapezoid(a.numpy(), dx=0.5, axis=2), rtol=1e-12)


# ------------------------------------------------------------------ channel-last helpers
# stenpy_engine.py:2199-2313.  Each check is a closed form or a per-voxel loop written from
# the formula, with distinct per-channel values so that a swapped channel, sign or axis fails.
def _lin_grid(shape):
    return [torch.from_numpy(a) for a in np.meshgrid(*[np.arange(n) * 1.0 for n in shape],
                                                     indexing="ij")]


def test_dot_of_gradient_of_linear_field_with_constant_vector():
    shape, dx = (5, 6, 7), (0.5, 1.0, 2.0)
    x, y, z = _lin_grid(shape)
    a, b, c = 3.0, -5.0, 7.0
    f = a * x * dx[0] + b * y * dx[1] + c * z * dx[2]     # df/dx_d = a, b, c exactly
    g = O.gradient_nd(f, dx)
    w = torch.stack([torch.full(shape, v, dtype=D64) for v in (2.0, 11.0, -13.0)], -1)
    assert torch.equal(O.dot(w, g), torch.full(shape, 2 * a + 11 * b - 13 * c, dtype=D64))
    assert torch.equal(O.dot(g, w), O.dot(w, g))
    u, v = rand((3, 4, 5, 3)), rand((3, 4, 5, 3))
    ex = bf_grid(lambda i: float(u[i][0] * v[i][0] + u[i][1] * v[i][1] + u[i][2] * v[i][2]),
                 (3, 4, 5))
    torch.testing.assert_close(O.dot(u, v), ex, rtol=0, atol=1e-14)


def test_sqnorm_of_curl_of_linear_field_is_constant():
    shape = (5, 6, 7)
    x, y, z = _lin_grid(shape)
    M = np.array([[1.0, 2.0, -3.0], [4.0, -1.0, 5.0], [-6.0, 7.0, 2.0]])
    X = (x, y, z)
    u = torch.stack([sum(M[cc, d] * X[d] for d in range(3)) for cc in range(3)], -1)
    # curl(Mx) = (M21 - M12, M02 - M20, M10 - M01), exact at one-sided edges too
    cv = (M[2, 1] - M[1, 2], M[0, 2] - M[2, 0], M[1, 0] - M[0, 1])
    assert torch.equal(O.curl(u, 1.0, "neumann"),
                       torch.stack([torch.full(shape, v, dtype=D64) for v in cv], -1))
    assert torch.equal(O.sqnorm(O.curl(u, 1.0, "neumann")),
                       torch.full(shape, sum(v * v for v in cv), dtype=D64))
    r = rand((3, 4, 5, 3))
    ex = bf_grid(lambda i: float(r[i][0] ** 2 + r[i][1] ** 2 + r[i][2] ** 2), (3, 4, 5))
    torch.testing.assert_close(O.sqnorm(r), ex, rtol=0, atol=1e-14)


def test_surface_normals_of_a_plane_exact_everywhere():
    shape, dx = (6, 9), (0.5, 2.0)
    i, j = np.meshgrid(np.arange(shape[0]) * dx[0], np.arange(shape[1]) * dx[1], indexing="ij")
    a, b = 3.0, -0.25
    n = O.surface_normals(torch.from_numpy(a * i + b * j), dx)
    m = math.sqrt(a * a + b * b + 1.0)
    ex = torch.tensor([-a / m, -b / m, 1.0 / m], dtype=D64).expand(shape + (3,))
    torch.testing.assert_close(n, ex, rtol=0, atol=1e-15)   # edge rows included
    with pytest.raises(ValueError):
        O.surface_normals(torch.zeros(3, 3, 3, dtype=D64))


def test_norm_last_and_neg_stack3_brute_force():
    t = rand((3, 4, 3))
    t[1, 2] = 0.0                                   # clamp: |0| -> eps
    out = O.norm_last(t)
    assert out.shape == (3, 4, 1)
    ex = bf_grid(lambda i: max(math.sqrt(sum(float(t[i][c]) ** 2 for c in range(3))), 1e-15),
                 (3, 4))
    torch.testing.assert_close(out[..., 0], ex, rtol=1e-15, atol=0)
    assert float(out[1, 2, 0]) == 1e-15
    g = np.random.default_rng(556)
    gx, gy = torch.from_numpy(g.standard_normal((3, 4))), torch.from_numpy(g.standard_normal((3, 4)))
    ns = O.neg_stack3(gx, gy)
    for i in np.ndindex(3, 4):
        m = math.sqrt(float(gx[i]) ** 2 + float(gy[i]) ** 2 + 1.0)
        ex3 = torch.tensor([-float(gx[i]) / m, -float(gy[i]) / m, 1.0 / m], dtype=D64)
        # the norm may round one ulp apart (torch's reduction order), then one division: 2 ulp
        torch.testing.assert_close(ns[i], ex3, rtol=4.5e-16, atol=0)


def test_div_last_scale_eye_stack_select_brute_force():
    a, b, b_full = rand((3, 4, 3)), rand((3, 4)), rand((3, 4, 3))
    q = O.div_last(a, b)                              # b gains a trailing axis
    for i in np.ndindex(3, 4, 3):
        assert float(q[i]) == float(a[i]) / float(b[i[:2]])
    assert torch.equal(O.div_last(a, b_full), a / b_full)
    s = rand((3, 4))
    e = O.scale_eye(s, 3)
    assert e.shape == (3, 4, 3, 3)
    for i in np.ndindex(3, 4, 3, 3):
        assert float(e[i]) == (float(s[i[:2]]) if i[2] == i[3] else 0.0)
    p, q2, r = rand((3, 4)), rand((3, 4)), rand((3, 4))
    st = O.stack_components(p, q2, r)
    for i in np.ndindex(3, 4):
        assert [float(st[i][c]) for c in range(3)] == [float(p[i]), float(q2[i]), float(r[i])]
    for c, src in enumerate((p, q2, r)):
        assert torch.equal(O.select_component(st, c), src)
    assert torch.equal(O.velocity_dot_grad(p, q2), bf_grid(lambda i: float(p[i]) * float(q2[i]), (3, 4)))


def test_material_derivative_of_linear_field_with_constant_velocity():
    shape, dx = (4, 5, 6), (0.5, 1.0, 2.0)
    x, y, z = _lin_grid(shape)
    f = 3.0 * x * dx[0] - 2.0 * y * dx[1] + 5.0 * z * dx[2]
    v = torch.stack([torch.full(shape, w, dtype=D64) for w in (1.5, -4.0, 0.25)], -1)
    assert torch.equal(O.material_derivative(f, v, dx),
                       torch.full(shape, 1.5 * 3 + 4.0 * 2 + 0.25 * 5, dtype=D64))

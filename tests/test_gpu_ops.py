"""Plain per-operator HIP kernels vs the CPU oracle (2x error rule, seeded Fourier inputs)."""
import pytest
import torch

import stenpy_amd as sp
from stenpy_amd import ops
from oracle import ops as O
from tests._util import assert_2x, field, vfield

pytestmark = pytest.mark.gpu
SHAPE = (20, 24, 28)


def ref(fn, *args, **kw):
    """(oracle64, ref32) of ``fn`` on the FP32 (or complex64) device inputs."""
    def up(a, real, cplx):
        if not isinstance(a, torch.Tensor):
            return a
        return a.detach().cpu().to(cplx if torch.is_complex(a) else real)
    a64 = [up(a, torch.float64, torch.complex128) for a in args]
    a32 = [up(a, torch.float32, torch.complex64) for a in args]
    return fn(*a64, **kw), fn(*a32, **kw)


@pytest.mark.parametrize("op", ["neg", "exp", "log", "sqrt", "sin", "cos", "tanh"])
def test_unary(op):
    f = field(SHAPE, 11)
    o64, o32 = ref(getattr(O, op), f)
    assert_2x(getattr(ops, op)(f), o64, o32, op)


@pytest.mark.parametrize("op", ["exp", "tanh", "sin"])
def test_unary_ragged_and_unaligned(op):
    """An element count that is not a multiple of 4 (the 4-wide kernel's scalar tail) and an
    input that starts 4 bytes past an aligned address (the scalar kernel)."""
    f = field((3, 5, 7), 15)
    o64, o32 = ref(getattr(O, op), f)
    assert_2x(getattr(ops, op)(f), o64, o32, op + " ragged")
    flat = field((4, 6, 9), 16).reshape(-1)
    g = flat[1:].reshape(5, 43)
    assert g.is_contiguous() and g.data_ptr() % 16 == 4
    o64, o32 = ref(getattr(O, op), g)
    assert_2x(getattr(ops, op)(g), o64, o32, op + " unaligned")


def test_clamp_and_scalar_mul():
    f = field(SHAPE, 12)
    o64, o32 = ref(O.clamp, f, lo=-0.2, hi=0.3)
    assert_2x(ops.clamp(f, lo=-0.2, hi=0.3), o64, o32, "clamp")
    o64, o32 = ref(O.mul, f, scalar=1.7)
    assert_2x(ops.mul(f, scalar=1.7), o64, o32, "mul scalar")


@pytest.mark.parametrize("op", ["add", "sub", "mul", "div"])
def test_binary(op):
    a, b = field(SHAPE, 13), field(SHAPE, 14)
    b[0, 0, :4] = torch.tensor([0.0, -0.0, 1e-16, -1e-17], device=b.device)
    o64, o32 = ref(getattr(O, op), a, b)
    assert_2x(getattr(ops, op)(a, b), o64, o32, op)


def test_binary_broadcast_scalar():
    a = field(SHAPE, 15)
    c = torch.tensor(2.5, dtype=torch.float64)
    o64, o32 = ref(O.mul, a, c)
    assert_2x(ops.mul(a, c), o64, o32, "mul 0-d")
    o64, o32 = ref(O.div, c, a)
    assert_2x(ops.div(c, a), o64, o32, "div 0-d")


@pytest.mark.parametrize("bc", ["periodic", "neumann", "dirichlet", "reflect", "other"])
@pytest.mark.parametrize("dx", [1.0, (0.5, 1.0, 2.0)])
def test_laplacian(bc, dx):
    f = field(SHAPE, 16)
    o64, o32 = ref(O.laplacian, f, dx=dx, boundary=bc)
    assert_2x(ops.laplacian(f, dx=dx, boundary=bc), o64, o32, f"laplacian {bc}")


@pytest.mark.parametrize("bc", ["periodic", "neumann"])
def test_laplacian_generic_rank(bc):
    f2 = field((1, 30, 26), 17)[0].contiguous()
    o64, o32 = ref(O.laplacian, f2, dx=0.5, boundary=bc)
    assert_2x(ops.laplacian(f2, dx=0.5, boundary=bc), o64, o32, "laplacian 2-D")
    v = vfield((10, 12, 14))
    o64, o32 = ref(O.laplacian, v, dx=1.0, boundary=bc)
    assert_2x(ops.laplacian(v, dx=1.0, boundary=bc), o64, o32, "laplacian 4-D quirk")


@pytest.mark.parametrize("bc", ["periodic", "neumann"])
@pytest.mark.parametrize("dim", [None, 0, 1, 2])
def test_gradient(bc, dim):
    f = field(SHAPE, 18)
    o64, o32 = ref(O.gradient, f, dx=0.7, boundary=bc, dim=dim)
    assert_2x(ops.gradient(f, dx=0.7, boundary=bc, dim=dim), o64, o32, f"gradient {bc} {dim}")


def test_gradient_short_axis():
    f = field((2, 9, 11), 19)
    o64, o32 = ref(O.gradient, f, dx=1.0, boundary="neumann")
    assert_2x(ops.gradient(f, dx=1.0, boundary="neumann"), o64, o32, "gradient n=2 axis")


@pytest.mark.parametrize("bc", ["periodic", "neumann"])
def test_divergence(bc):
    v = vfield(SHAPE)
    o64, o32 = ref(O.divergence, v, dx=(1.0, 0.5, 2.0), boundary=bc)
    assert_2x(ops.divergence(v, dx=(1.0, 0.5, 2.0), boundary=bc), o64, o32, f"divergence {bc}")


def test_curl_3d_and_2d():
    v = vfield(SHAPE)
    o64, o32 = ref(O.curl, v, dx=(0.5, 1.0, 1.5))
    assert_2x(ops.curl(v, dx=(0.5, 1.0, 1.5)), o64, o32, "curl 3-D")
    v2 = vfield((1, 18, 22), seeds=(6, 7))[0].contiguous()
    o64, o32 = ref(O.curl, v2, dx=1.0)
    assert_2x(ops.curl(v2, dx=1.0), o64, o32, "curl 2-D")


def test_gradient_nd_hessian_jacobian():
    f = field(SHAPE, 20)
    o64, o32 = ref(O.gradient_nd, f, dx=(1.0, 0.5, 0.25))
    assert_2x(ops.gradient_nd(f, dx=(1.0, 0.5, 0.25)), o64, o32, "gradient_nd")
    o64, o32 = ref(O.hessian, f, dx=0.5)
    assert_2x(ops.hessian(f, dx=0.5), o64, o32, "hessian")
    v = vfield(SHAPE)
    o64, o32 = ref(O.jacobian, v, dx=1.0)
    assert_2x(ops.jacobian(v, dx=1.0), o64, o32, "jacobian")


def test_gradient_nd_squeezed():
    f = field((1, 16, 18), 21)
    o64, o32 = ref(O.gradient_nd, f, dx=1.0)
    assert_2x(ops.gradient_nd(f, dx=1.0), o64, o32, "gradient_nd squeeze")


def test_hessian_quadratic_exact():
    from stenpy_amd.gen import quadratic_host
    q = torch.from_numpy(quadratic_host((12, 14, 16))).cuda()
    h = ops.hessian(q, dx=1.0).cpu()
    o64 = O.hessian(q.cpu().double(), dx=1.0)
    assert torch.equal(h.double(), o64), "integer quadratic hessian must be exact"


@pytest.mark.parametrize("bc", ["periodic", "neumann"])
def test_mean_curvature(bc):
    f = field(SHAPE, 22)
    o64, o32 = ref(O.mean_curvature, f, dx=1.0, boundary=bc)
    assert_2x(ops.mean_curvature(f, dx=1.0, boundary=bc), o64, o32, f"mean_curvature {bc}")


def test_material_derivative_surface_normals():
    f, v = field(SHAPE, 23), vfield(SHAPE)
    o64, o32 = ref(O.material_derivative, f, v, dx=1.0)
    assert_2x(ops.material_derivative(f, v, dx=1.0), o64, o32, "material_derivative")
    h = field((1, 20, 24), 24)[0].contiguous()
    o64, o32 = ref(O.surface_normals, h, dx=1.0)
    assert_2x(ops.surface_normals(h, dx=1.0), o64, o32, "surface_normals")


def test_channel_helpers():
    u, g = vfield(SHAPE), vfield(SHAPE, seeds=(8, 9, 10))
    o64, o32 = ref(O.dot, u, g)
    assert_2x(ops.dot(u, g), o64, o32, "dot")
    o64, o32 = ref(O.sqnorm, u)
    assert_2x(ops.sqnorm(u), o64, o32, "sqnorm")
    o64, o32 = ref(O.norm_last, u)
    assert_2x(ops.norm_last(u), o64, o32, "norm_last")
    o64, o32 = ref(O.select_component, u, 1)
    assert_2x(ops.select_component(u, 1), o64, o32, "select_component")


def test_remaining_channel_helpers():
    """stack_components, div_last (both divisor ranks), neg_stack3, scale_eye and
    velocity_dot_grad (stenpy_engine.py:2247-2313), directly and inside a graph."""
    a, b, c = field(SHAPE, 31), field(SHAPE, 32), field(SHAPE, 33)
    u = vfield(SHAPE, seeds=(34, 35, 36))
    o64, o32 = ref(O.stack_components, a, b, c)
    assert_2x(ops.stack_components(a, b, c), o64, o32, "stack_components")
    d = (b.abs() + 0.25).contiguous()
    o64, o32 = ref(O.div_last, u, d)
    assert_2x(ops.div_last(u, d), o64, o32, "div_last rank-1 divisor")
    w = (u.abs() + 0.5).contiguous()
    o64, o32 = ref(O.div_last, u, w)
    assert_2x(ops.div_last(u, w), o64, o32, "div_last same-rank divisor")
    o64, o32 = ref(O.neg_stack3, a * 40, b * 40)
    assert_2x(ops.neg_stack3(a * 40, b * 40), o64, o32, "neg_stack3")
    o64, o32 = ref(O.scale_eye, a, 3)
    got = ops.scale_eye(a, 3)
    assert got.shape == SHAPE + (3, 3)
    assert_2x(got, o64, o32, "scale_eye")
    o64, o32 = ref(O.velocity_dot_grad, a, c)
    assert_2x(ops.velocity_dot_grad(a, c), o64, o32, "velocity_dot_grad")
    # the same helpers reached from an expression through the eager executor
    from oracle import graph_eval
    g, sink, *_ = sp.compile_expression("norm_last(stack_components(f, g, h))",
                                        field_map={"f": a, "g": b, "h": c})
    res = sp.Runtime(device="cuda", fusion_enabled=False).run(g)[sink]
    assert_2x(res, graph_eval.evaluate(g, torch.float64)[sink],
              graph_eval.evaluate(g, torch.float32)[sink], "helpers in a graph")


@pytest.mark.parametrize("kind", ["sum", "mean", "norm_l2", "variance", "min_max"])
def test_reductions(kind):
    f = field((32, 33, 34), 25) + 0.5
    fn = {"sum": O.op_sum, "mean": O.mean, "norm_l2": O.norm_l2, "variance": O.variance,
          "min_max": O.min_max}[kind]
    o64, o32 = ref(fn, f)
    b = getattr(ops, {"sum": "op_sum"}.get(kind, kind))(f)
    assert b.dtype == torch.float64
    assert_2x(b, o64, o32, kind)


@pytest.mark.parametrize("method", ["simpson", "trapezoid"])
@pytest.mark.parametrize("shape", [(20, 24, 28), (21, 25, 2)])
def test_integrate(method, shape):
    f = field(shape, 26)
    o64, o32 = ref(O.integrate, f, dx=(0.5, 1.0, 0.25), method=method)
    assert_2x(ops.integrate(f, dx=(0.5, 1.0, 0.25), method=method), o64, o32, "integrate")


def test_tensor_ops():
    f = field(SHAPE, 27)
    h = ops.hessian(f, dx=1.0)
    for name in ["trace", "determinant", "eigenvalues", "deviatoric", "inverse"]:
        o64, o32 = ref(getattr(O, name), h)
        assert_2x(getattr(ops, name)(h), o64, o32, name)


def test_determinant_global_regularisation():
    f = field((10, 10, 10), 28)
    h = ops.hessian(f, dx=1.0)
    h[3, 3, 3, 1, 1] = 0.0                      # one tiny diagonal entry switches on +1e-12 I
    o64, o32 = ref(O.determinant, h)
    assert_2x(ops.determinant(h), o64, o32, "det regularised")


def test_eigenvalues_of_a_multiple_of_identity_are_exact():
    from stenpy_amd.gen import quadratic_host
    q = torch.from_numpy(quadratic_host((8, 9, 10))).cuda()
    e = ops.eigenvalues(ops.hessian(q, dx=1.0)).cpu()
    assert torch.equal(e, torch.full_like(e, 2.0)), "2I must give exactly (2, 2, 2)"


@pytest.mark.parametrize("spectrum", ["gap1e-1", "gap1e-4", "gap1e-7", "double", "triple",
                                      "offset1e4", "tiny", "huge"])
def test_eigenvalues_pernicious_spectra(spectrum):
    # rotated diagonals: near-double, exactly double and near-triple eigenvalues, a large common
    # offset, and magnitudes whose sixth powers leave the FP32 range
    g = torch.Generator().manual_seed(0)
    n = 4096
    Q, _ = torch.linalg.qr(torch.randn(n, 3, 3, generator=g, dtype=torch.float64))
    one, r = torch.ones(n, dtype=torch.float64), lambda: torch.randn(n, generator=g, dtype=torch.float64)
    scale = 1.0
    if spectrum.startswith("gap"):
        lam = [one, one + float(spectrum[3:]) * r(), 3 * r()]
    elif spectrum == "double":
        lam = [one, one, 3 * r()]
    elif spectrum == "triple":
        lam = [2 * one, 2 * one + 1e-6 * r(), 2 * one + 1e-6 * r()]
    elif spectrum == "offset1e4":
        lam = [1e4 * one, 1e4 * one + 1e-2 * r(), r()]
    else:
        lam, scale = [r(), r(), r()], (1e-25 if spectrum == "tiny" else 1e25)
    A = (Q @ torch.diag_embed(torch.stack(lam, -1) * scale) @ Q.transpose(1, 2)).float()
    A = 0.5 * (A + A.transpose(1, 2))
    o64, o32 = ref(O.eigenvalues, A)
    assert_2x(ops.eigenvalues(A.cuda()), o64, o32, f"eig {spectrum}")


@pytest.mark.parametrize("bc", ["periodic", "neumann", "dirichlet", "reflect"])
@pytest.mark.parametrize("window", [3, 5])
def test_moving_average(bc, window):
    f = field((18, 22, 40), 21)
    out = ops.moving_average(f, window, bc)
    f64 = f.double().cpu()
    assert_2x(out, O.moving_average(f64, window, bc), O.moving_average(f64.float(), window, bc),
              f"moving_average {bc} w={window}")
    with pytest.raises(ValueError):
        ops.moving_average(f, 4, bc)


def test_conv3d_spectral_and_expressions():
    f = field((18, 22, 40), 22)
    out = ops.conv3d_spectral(f)
    f64 = f.double().cpu()
    assert_2x(out, O.conv3d_spectral(f64), O.conv3d_spectral(f64.float()), "conv3d_spectral")
    # both are reachable from expressions (eager plain kernels inside a fused graph)
    from oracle import graph_eval
    for expr in ["moving_average(f) + laplacian(f)", "conv3d_spectral(tanh(f)) * 2"]:
        g, sink, *_ = sp.compile_expression(expr, boundary="periodic", field_map={"f": f})
        rt = sp.Runtime(device="cuda")
        res = rt.run(g)[sink]
        assert_2x(res, graph_eval.evaluate(g, torch.float64)[sink], graph_eval.evaluate(g, torch.float32)[sink], expr)


# ----------------------------------------------------------------------------- spectral
def _ri(t):
    """Complex results compared component-wise under the 2x rule."""
    return torch.view_as_real(t.to(torch.complex128)) if torch.is_complex(t) else t


@pytest.mark.parametrize("shape", [(64,), (12, 20), (10, 12, 15)])
@pytest.mark.parametrize("norm", ["backward", "ortho", "forward"])
def test_fft_family(shape, norm):
    f = field(shape if len(shape) == 3 else (1, 1) + shape if len(shape) == 1 else (1,) + shape, 41)
    f = f.reshape(shape).contiguous()
    F_ = ops.fft(f, norm=norm)
    assert F_.dtype == torch.complex64 and F_.shape == f.shape
    o64, o32 = ref(O.fft, f, norm=norm)
    assert_2x(_ri(F_), _ri(o64), _ri(o32), f"fft {shape} {norm}")
    for name in ("ifft", "ifft_complex"):
        out = getattr(ops, name)(F_, norm=norm)
        o64, o32 = ref(getattr(O, name), F_, norm=norm)
        assert_2x(_ri(out), _ri(o64), _ri(o32), f"{name} {shape} {norm}")
    # a real input to the inverse is a complex one with zero imaginary part
    o64, o32 = ref(O.ifft, f, norm=norm)
    assert_2x(ops.ifft(f, norm=norm), o64, o32, f"ifft(real) {shape} {norm}")
    c = ops.fft_complex(f.to(torch.complex64), norm="numpy_compatible")
    o64, o32 = ref(O.fft_complex, f.to(torch.complex64), norm="numpy_compatible")
    assert_2x(_ri(c), _ri(o64), _ri(o32), f"fft_complex {shape}")


@pytest.mark.parametrize("shape,dx", [((16, 18, 20), 1.0), ((9, 14, 11), (0.5, 1.0, 0.25)), ((24, 30), (0.1, 0.2)),
                                      ((50,), 0.3)])
def test_spectral_derivatives(shape, dx):
    base = field((1, 1) + shape if len(shape) == 1 else ((1,) + shape if len(shape) == 2 else shape), 42)
    f = base.reshape(shape).contiguous()
    for dim in range(len(shape)):
        o64, o32 = ref(O.spectral_gradient, f, dx=dx, dim=dim)
        assert_2x(ops.spectral_gradient(f, dx=dx, dim=dim), o64, o32, f"spectral_gradient {shape} dim={dim}")
    o64, o32 = ref(O.spectral_laplacian, f, dx=dx)
    assert_2x(ops.spectral_laplacian(f, dx=dx), o64, o32, f"spectral_laplacian {shape}")


def test_spectral_ops_in_graphs_and_errors():
    from oracle import graph_eval
    shape = (12, 16, 14)
    f = field(shape, 43)
    for expr in ["spectral_laplacian(f) + laplacian(f)", "spectral_gradient(f, dim=2) * f", "ifft(fft(f)) - f"]:
        g, sink, *_ = sp.compile_expression(expr, boundary="periodic", field_map={"f": f})
        rt = sp.Runtime(device="cuda")
        out = rt.run(g)[sink]
        o64 = graph_eval.evaluate(g, torch.float64)[sink]
        o32 = graph_eval.evaluate(g, torch.float32)[sink]
        assert_2x(out, o64, o32, expr, slack=1e-12)
    with pytest.raises(ValueError, match="1-D, 2-D or 3-D"):
        ops.fft(torch.zeros((2, 2, 2, 2), device="cuda"))
    with pytest.raises(IndexError):
        ops.spectral_gradient(f, dim=3)


# ----------------------------------------------------------------------------- geometry
@pytest.mark.parametrize("shape,dx,density", [((20, 24, 28), 1.0, 0.3), ((17, 9, 33), (0.5, 1.0, 2.0), 0.05),
                                              ((40, 36), (0.3, 0.7), 0.5), ((64,), 0.25, 0.1),
                                              ((12, 14, 16), 1.0, 0.002)])
def test_distance_transform(shape, dx, density):
    g = torch.Generator().manual_seed(7)
    keep = torch.rand(shape, generator=g) < density
    t = torch.where(keep, torch.randn(shape, generator=g), torch.zeros(shape))
    t.view(-1)[0] = 1.0            # at least one voxel with t != 0
    tc = t.float().cuda()
    out = ops.distance_transform(tc, dx=dx)
    o64, o32 = ref(O.distance_transform, tc, dx=dx)
    assert_2x(out, o64, o32, f"distance_transform {shape} {dx}")
    assert torch.equal(out.cpu() == 0, t != 0)


def test_distance_transform_in_graphs_and_empty_background():
    from oracle import graph_eval
    f = field((18, 20, 22), 44)
    g, sink, *_ = sp.compile_expression("distance_transform(f - clamp(f, lo=-0.05, hi=0.05))", boundary="periodic",
                                        field_map={"f": f})
    out = sp.Runtime(device="cuda").run(g)[sink]
    o64 = graph_eval.evaluate(g, torch.float64)[sink]
    o32 = graph_eval.evaluate(g, torch.float32)[sink]
    assert_2x(out, o64, o32, "distance_transform graph")
    assert torch.isinf(ops.distance_transform(torch.zeros((4, 5, 6), device="cuda"))).all()


# ----------------------------------------------------------------------------- statistics and integrals
@pytest.mark.parametrize("shape,dim", [((12, 14, 16), 0), ((12, 14, 16), 1), ((12, 14, 16), 2), ((30, 1), 0), ((1, 9), 0)])
def test_cumulative_integral(shape, dim):
    f = field((1,) + shape if len(shape) == 2 else shape, 45).reshape(shape).contiguous()
    o64, o32 = ref(O.cumulative_integral, f, dx=0.5, dim=dim)
    assert_2x(ops.cumulative_integral(f, dx=0.5, dim=dim), o64, o32, f"cumulative_integral {shape} dim={dim}")


def test_pair_statistics_entropy_and_surface_integral():
    a, b = field((20, 24, 28), 46), field((20, 24, 28), 47)
    b = 0.4 * a + b
    for name in ("covariance", "correlation"):
        out = getattr(ops, name)(a, b)
        assert out.dtype == torch.float64 and out.ndim == 0
        o64, o32 = ref(getattr(O, name), a, b)
        assert_2x(out, o64, o32, name)
    o64, o32 = ref(O.entropy, a)
    assert_2x(ops.entropy(a), o64, o32, "entropy")
    for d in range(3):
        o64, o32 = ref(O.entropy, a, dim=d)
        assert_2x(ops.entropy(a, dim=d), o64, o32, f"entropy dim={d}")
    v, n = vfield((10, 12, 14)), vfield((10, 12, 14), seeds=(6, 7, 8))
    w = field((10, 12, 14), 48).abs()
    for t in (v, field((10, 12, 14), 49).unsqueeze(-1).contiguous()):
        o64, o32 = ref(O.surface_integral, t, n, w)
        assert_2x(ops.surface_integral(t, n, w), o64, o32, f"surface_integral t{tuple(t.shape)}")
    with pytest.raises(ValueError):
        ops.covariance(a, b[:-1])
    with pytest.raises(ValueError):
        ops.surface_integral(v, n, w[:-1])


def test_statistics_in_graphs():
    from oracle import graph_eval
    f = field((14, 16, 18), 50)
    for expr in ["covariance(f, tanh(f))", "correlation(f, f * f)", "entropy(f)", "cumulative_integral(f, dim=1) * f"]:
        g, sink, *_ = sp.compile_expression(expr, boundary="periodic", field_map={"f": f})
        out = sp.Runtime(device="cuda").run(g)[sink]
        o64 = graph_eval.evaluate(g, torch.float64)[sink]
        o32 = graph_eval.evaluate(g, torch.float32)[sink]
        assert_2x(out, o64, o32, expr)


@pytest.mark.parametrize("persistent", [False, True])
def test_hbm_probes_compute_their_mix(persistent):
    """The bandwidth probes read and write what they claim: the 4:1 mixes sum each voxel's f and
    its three u components (u planar as three thirds, or channel-last), the copy copies, the
    write fills i-dependent values, the read sums per block.  Small integers, so every sum is
    exact in FP32."""
    from stenpy_amd import _lib as L
    n = 4 * 3000
    g = torch.Generator().manual_seed(5)
    a = torch.randint(-50, 50, (n,), generator=g).float().cuda()
    u = torch.randint(-50, 50, (n, 3), generator=g).float().cuda()
    out = torch.empty_like(a)
    L.probe_stream("4r1w_aos", persistent, a=a, u=u, out=out)
    assert torch.equal(out, a + u.sum(1))
    planar = u.t().contiguous().reshape(-1)          # three thirds: component c at c*n + i
    L.probe_stream("4r1w", persistent, a=a, u=planar, out=out)
    # the planar probe reads its thirds as if they were interleaved: voxel v's "components" are
    # the floats 3v..3v+2 of the 4-voxel group's 12, drawn from the three thirds in turn
    grp = torch.stack([planar[:n].view(-1, 4), planar[n:2 * n].view(-1, 4),
                       planar[2 * n:].view(-1, 4)], 1).reshape(-1, 12).view(-1, 4, 3).sum(2).reshape(-1)
    assert torch.equal(out, a + grp)
    L.probe_stream("copy", persistent, a=a, out=out)
    assert torch.equal(out, a)
    L.probe_stream("write", persistent, a=out)
    i = torch.arange(n // 4, device="cuda")
    want = ((i & 1023).float().unsqueeze(1) + torch.arange(4, device="cuda").float()).reshape(-1)
    assert torch.equal(out, want)
    part = torch.empty(L.probe_blocks(persistent, n), dtype=torch.float64, device="cuda")
    L.probe_stream("read", persistent, a=a, partials=part)
    assert float(part.sum()) == float(a.double().sum())

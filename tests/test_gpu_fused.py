"""Fused hipRTC path vs the CPU oracle: whole graphs, 2x error rule."""
import os

import pytest
import torch

import stenpy_amd as sp
from stenpy_amd.fused import Schedule
from oracle import graph_eval
from tests._util import assert_2x, errs, field, vfield

pytestmark = pytest.mark.gpu
os.environ["STENPY_FUSED_STRICT"] = "1"

STENCIL_GRAPH = "divergence(gradient(f)) + dot(u, gradient(f)) + sqnorm(curl(u))"


def run_and_check(expr, fields, boundary="periodic", dx=1.0, reductions=(), schedule=None,
                  out_dtype=None, slack=None, edge_mode=None):
    g, sink, _, _, _ = sp.compile_expression(expr, dx=dx, boundary=boundary, field_map=fields)
    reds = [g.add(r, (sink,), {}) for r in reductions]
    rt = sp.Runtime(device="cuda", fusion_enabled=True)
    rt.fused.edge_mode = edge_mode
    res = rt.fused.run(g, rt, wanted=[sink] + reds, out_dtype=out_dtype, schedule=schedule)
    o64 = graph_eval.evaluate(g, torch.float64)
    o32 = graph_eval.evaluate(g, torch.float32)
    out = {}
    for nid in [sink] + reds:
        extra = 0.0
        if out_dtype == torch.bfloat16 and nid == sink:
            # BF16 storage: half a BF16 ulp of the largest magnitude on top of the 2x rule
            extra = 0.5 * 2.0 ** -7 * 2.0 ** torch.floor(torch.log2(o64[nid].abs().max())).item()
        if slack and nid in slack:
            extra += slack[nid]
        out[g.node(nid).op_name] = assert_2x(res[nid], o64[nid], o32[nid], f"{expr} {g.node(nid).op_name}", extra)
    return res, out


def test_single_op_expressions():
    f = field((32, 32, 32), 1)
    for expr in ["f**2 + tanh(f)*exp(f)", "laplacian(f)", "gradient(f)", "mean(f)", "norm_l2(f)"]:
        run_and_check(expr, {"f": f})


@pytest.mark.parametrize("expr", [
    "exp(f*110 - 15)",        # down into FP32 subnormals and zero, up to 2^116
    "exp(f*0.0001)",
    "tanh(f*12)",             # saturates to +-1 beyond 9.02
    "tanh(f*0.00001)",
    "log(f*f)", "log(f)",     # eps clamp of the negative half
    "log(f*0.001 + 0.002)",
    "sin(f*1000)", "cos(f*3)", "cos(f*0.001)",
    "sin(f*700000)",          # beyond the short argument reduction
])
def test_transcendentals_over_wide_ranges(expr):
    """The kernels' exp / tanh / log / sin / cos (field |f| <= 0.87) against the oracle."""
    run_and_check(expr, {"f": field((20, 24, 66), 7)})


@pytest.mark.parametrize("edge_mode", ["band", "variants"])
@pytest.mark.parametrize("shape", [(32, 32, 32), (21, 37, 70), (9, 8, 130)])
def test_stencil_graph(shape, edge_mode):
    f = field(shape, 2)
    u = vfield(shape)
    run_and_check(STENCIL_GRAPH, {"f": f, "u": u}, reductions=("sum", "mean", "variance", "norm_l2"),
                  edge_mode=edge_mode)


@pytest.mark.parametrize("sched", [Schedule(32, 8, 16, 256), Schedule(64, 8, 64, 256),
                                   Schedule(128, 4, 32, 256), Schedule(64, 8, 8, 512)])
def test_stencil_graph_schedules(sched):
    shape = (40, 45, 150)
    run_and_check(STENCIL_GRAPH, {"f": field(shape, 2), "u": vfield(shape)}, schedule=sched,
                  reductions=("sum", "variance"))


def test_stencil_graph_bf16_output():
    shape = (24, 30, 66)
    run_and_check(STENCIL_GRAPH, {"f": field(shape, 2), "u": vfield(shape)},
                  reductions=("sum", "mean", "variance", "norm_l2"), out_dtype=torch.bfloat16)


@pytest.mark.parametrize("edge_mode", ["band", "variants"])
@pytest.mark.parametrize("bc", ["periodic", "neumann", "dirichlet", "reflect"])
def test_stencil_compositions(bc, edge_mode):
    shape = (18, 22, 70)
    f = field(shape, 30)
    u = vfield(shape)
    for expr in ["laplacian(laplacian(f))", "divergence(curl(u)) + laplacian(exp(f))",
                 "gradient(f) * 2 + curl(u)", "laplacian(f) * tanh(f) - clamp(f, lo=-0.1, hi=0.2)",
                 "divergence(u) + sqrt(f**2 + 1) + log(f**2)"]:
        run_and_check(expr, {"f": f, "u": u}, boundary=bc, edge_mode=edge_mode)


@pytest.mark.parametrize("edge_mode", ["band", "variants"])
def test_hessian_tensor_ops(edge_mode):
    shape = (20, 24, 30)
    f = field(shape, 8)
    for expr in ["eigenvalues(hessian(f))", "determinant(hessian(f))", "trace(hessian(f))",
                 "deviatoric(hessian(f))", "mean_curvature(f)", "jacobian(curl(u))"]:
        run_and_check(expr, {"f": f, "u": vfield(shape)}, boundary="periodic", edge_mode=edge_mode)


@pytest.mark.parametrize("edge_mode", ["band", "variants"])
@pytest.mark.parametrize("sched", [Schedule(128, 8, 16, 256), Schedule(64, 4, 16, 256), Schedule(32, 8, 16, 128),
                                   Schedule(16, 8, 16, 128), Schedule(32, 8, 8, 128)])
def test_wide_outputs_staged_stores(sched, edge_mode):
    # 9-channel outputs go through the per-wave LDS transpose (a wave spans 64 / TX rows of
    # narrow tiles); x = 200 and y = 20 leave partial tiles whose lanes past the edge store nothing
    shape = (12, 20, 200)
    f = field(shape, 8)
    for bc in ("periodic", "neumann"):
        for expr in ["hessian(f)", "deviatoric(hessian(f)) * 2", "jacobian(u)"]:
            run_and_check(expr, {"f": f, "u": vfield(shape)}, boundary=bc, schedule=sched,
                          edge_mode=edge_mode)


def test_scalar_from_reduction_feeds_next_group():
    shape = (16, 18, 66)
    f = field(shape, 31)
    run_and_check("f - mean(f)", {"f": f}, reductions=("variance",))


def test_fused_matches_eager_path():
    shape = (20, 22, 68)
    f, u = field(shape, 2), vfield(shape)
    g, sink, _, _, _ = sp.compile_expression(STENCIL_GRAPH, boundary="periodic", field_map={"f": f, "u": u})
    rt = sp.Runtime(device="cuda")
    a = rt.fused.run(g, rt, wanted=[sink])[sink]
    b = rt.run_eager(g)[sink]
    o64 = graph_eval.evaluate(g, torch.float64)[sink]
    ea, _ = errs(a, o64, o64.float())
    eb, _ = errs(b, o64, o64.float())
    assert ea <= 2 * eb + 1e-12, (ea, eb)


def test_determinant_regularises_only_when_a_diagonal_is_small(monkeypatch):
    """The fused determinant adds 1e-12 I only when the any-small-diagonal flag is set
    (stenpy_engine.py:2683-2712): the flag pass runs first and the main pass reads its device
    flag -- 0 for a generic field, 1 for f = x^2 (H_yy = H_zz = 0)."""
    from stenpy_amd import fused as F
    launches = []
    orig = F.FusedExecutor.prepare_launch

    def keep(self, cg, graph, results):
        ln = orig(self, cg, graph, results)
        launches.append(ln)
        return ln

    monkeypatch.setattr(F.FusedExecutor, "prepare_launch", keep)
    shape = (20, 24, 30)
    run_and_check("determinant(hessian(f))", {"f": field(shape, 8)}, boundary="periodic")
    assert float(launches[-1].detflag.item()) == 0.0
    x = torch.arange(shape[0], dtype=torch.float32).view(-1, 1, 1).expand(shape).contiguous()
    f = (x * x).cuda()
    g, sink, _, _, _ = sp.compile_expression("determinant(hessian(f))", boundary="periodic", field_map={"f": f})
    rt = sp.Runtime(device="cuda", fusion_enabled=True)
    out = rt.fused.run(g, rt, wanted=[sink])[sink]
    assert float(launches[-1].detflag.item()) == 1.0
    # H = diag(2, 0, 0) exactly, so det(H + 1e-12 I) = (2 + 1e-12) * 1e-24
    ref = graph_eval.evaluate(g, torch.float64)[sink]
    assert (out.double().cpu() - ref).abs().max().item() <= 1e-30


def test_captured_fused_run_replays_match_eager():
    """hipGraph capture of a fused evaluation (main + side-stream tail kernels, reductions,
    determinant flag pass): every replay reproduces the eager result exactly."""
    shape = (40, 45, 150)
    f, u = field(shape, 2), vfield(shape)
    g, sink, _, _, _ = sp.compile_expression(STENCIL_GRAPH, boundary="periodic", field_map={"f": f, "u": u})
    reds = [g.add(r, (sink,), {}) for r in ("sum", "variance")]
    g2, sink2, _, _, _ = sp.compile_expression("determinant(hessian(f)) + trace(hessian(f))",
                                               boundary="periodic", field_map={"f": f})
    for graph, wanted in ((g, [sink] + reds), (g2, [sink2])):
        rt = sp.Runtime(device="cuda", fusion_enabled=True)
        eager = {k: v.clone() for k, v in rt.fused.run(graph, rt, wanted=wanted).items() if k in wanted}
        cap, res = rt.fused.capture(graph, rt, wanted)
        for rep in range(3):
            for k in wanted:
                res[k].fill_(float("nan"))
            torch.cuda.synchronize()
            cap.replay()
            torch.cuda.synchronize()
            for k in wanted:
                assert torch.equal(res[k], eager[k]), (rep, k)
            torch.randn(1 << 20, device="cuda").sum().item()     # unrelated eager work


@pytest.mark.parametrize("edge_mode", ["band", "variants"])
@pytest.mark.parametrize("bc", ["periodic", "neumann", "dirichlet"])
def test_load_transformed_sources(bc, edge_mode):
    """Point-wise chains with transcendentals under stencils are applied once per loaded cell
    (queue and ring sources, partial tiles, every boundary): results stay within the rule."""
    shape = (22, 27, 70)
    f = field(shape, 12)
    for expr in ["laplacian(tanh(f)) * tanh(f)", "divergence(gradient(tanh(f)))", "gradient(exp(f)) * 2",
                 "hessian(log(f*f+1))", "laplacian(laplacian(exp(f*0.5)))"]:
        run_and_check(expr, {"f": f}, boundary=bc, edge_mode=edge_mode)


def test_reference_fused_entry_compile_and_execute():
    """The call the reference's Runtime._run_fused makes (stenpy_engine.py:966-981): a graph
    with FP64 field constants, a runtime exposing ``device``; the result holds the sink, widened
    to FP64 for the reference's downstream asserts, within the 2x rule."""
    from stenpy_amd import fused_compiler as fc
    shape = (20, 22, 68)
    f, u = field(shape, 2).double(), vfield(shape).double()
    g, sink, _, fm, _ = sp.compile_expression(STENCIL_GRAPH, boundary="periodic", field_map={"f": f, "u": u})

    class RefRuntime:
        device = torch.device("cuda")
    assert any(n.op_name in fc.FUSIBLE_OPS for n in g.nodes.values())
    res = fc.compile_and_execute(g, RefRuntime(), None)
    assert res[sink].dtype == torch.float64 and fm["f"] in res
    o64 = graph_eval.evaluate(g, torch.float64)[sink]
    o32 = graph_eval.evaluate(g, torch.float32)[sink]
    assert_2x(res[sink], o64, o32, "compile_and_execute")


def test_scalar_program_ops_and_chaining():
    """sp_scalar_program evaluates its ops in order in FP64: every unary and binary op against
    Python's math, constants on either side, and an op reading an earlier op's result."""
    import math
    from stenpy_amd import _lib as L
    dev = torch.device("cuda", 0)
    x = torch.tensor([0.37], dtype=torch.float64, device=dev)
    y = torch.tensor([-2.5], dtype=torch.float64, device=dev)
    outs = [torch.empty((), dtype=torch.float64, device=dev) for _ in range(16)]
    U, B = L.UNARY, L.BINARY

    def op(i, binary, code, a=None, b=None, av=0.0, bv=0.0, p0=0.0, p1=0.0):
        d = L.SpScalarOp()
        d.a = a.data_ptr() if a is not None else None
        d.b = b.data_ptr() if b is not None else None
        d.out, d.av, d.bv, d.p0, d.p1, d.binary, d.op = outs[i].data_ptr(), av, bv, p0, p1, binary, code
        return d
    prog = [op(0, 0, U["neg"], x), op(1, 0, U["exp"], x), op(2, 0, U["log"], x, p0=1e-15),
            op(3, 0, U["sqrt"], y), op(4, 0, U["sin"], y), op(5, 0, U["cos"], y), op(6, 0, U["tanh"], y),
            op(7, 0, U["clamp"], y, p0=-1.0, p1=1.0), op(8, 0, U["scale"], x, p0=3.0),
            op(9, 1, B["add"], x, y), op(10, 1, B["sub"], None, x, av=1.0), op(11, 1, B["mul"], x, None, bv=-4.0),
            op(12, 1, B["div"], x, None, bv=0.0, p0=1e-15), op(13, 1, B["div"], y, x, p0=1e-15),
            op(14, 1, B["mul"], outs[1], outs[9]),          # reads results made earlier in the program
            op(15, 0, U["neg"], outs[14])]
    L.scalar_program(prog, device=dev)
    got = [float(t) for t in outs]
    a, b = 0.37, -2.5
    want = [-a, math.exp(a), math.log(a), 0.0, math.sin(b), math.cos(b), math.tanh(b), -1.0, 3 * a,
            a + b, 1.0 - a, -4 * a, a / 1e-15, b / a, math.exp(a) * (a + b), -math.exp(a) * (a + b)]
    for i, (g_, w) in enumerate(zip(got, want)):
        assert g_ == pytest.approx(w, rel=1e-15, abs=0.0), (i, g_, w)


def test_scalar_steps_between_levels_run_as_one_program():
    """Scalar arithmetic on reduction results (neg, exp, log, mul by constants, clamp) between
    fused levels: one device program, results against the FP64 oracle under the 2x rule."""
    f = field((30, 26, 34), 7)
    expr = "f * exp(-mean(f)) - log(norm_l2(f) + 1) * 0.5 + clamp(mean(f*f))"
    run_and_check(expr, {"f": f}, reductions=("sum",))


@pytest.mark.parametrize("shape", [(3, 3, 3), (3, 4, 5), (2, 5, 6), (1, 6, 7), (4, 1, 5)])
@pytest.mark.parametrize("expr", ["f**2 + tanh(f)*exp(f)", "laplacian(f)", "gradient(f)",
                                  "divergence(gradient(f)) + mean(f)", "norm_l2(f) * f"])
def test_small_and_degenerate_grids(shape, expr):
    """Grids at the fused path's lower limit (every axis 3) and below it (an axis of 1 or 2,
    which the plain per-op kernels take), periodic: the runtime's result against the oracle."""
    f = field(shape, 23)
    g, sink, *_ = sp.compile_expression(expr, boundary="periodic", field_map={"f": f})
    rt = sp.Runtime(device="cuda", fusion_enabled=True)
    out = rt.run(g)[sink]
    o64 = graph_eval.evaluate(g, torch.float64)[sink]
    o32 = graph_eval.evaluate(g, torch.float32)[sink]
    assert tuple(out.shape) == tuple(o64.shape)
    assert_2x(out, o64, o32, f"{expr} on {shape}")


def test_empty_field_point_wise():
    """A field with no voxels evaluates to an empty result, as torch's element-wise ops give."""
    f = torch.empty((0, 4, 4), device="cuda")
    g, sink, *_ = sp.compile_expression("f**2 + tanh(f)*exp(f)", boundary="periodic", field_map={"f": f})
    out = sp.Runtime(device="cuda", fusion_enabled=True).run(g)[sink]
    assert tuple(out.shape) == (0, 4, 4)


@pytest.mark.parametrize("shape", [(21, 24, 30), (9, 8, 131), (3, 5, 6)])
def test_fused_quadrature_and_extrema(shape):
    """integrate (Simpson and trapezoid, anisotropic spacing; odd and even axis lengths, where
    Simpson's last interval differs), min_max and variance fused into the producing group.

    The summed field is a periodic Laplacian with power-of-two spacings, whose FP32 values the
    kernel reproduces bit for bit, so the reductions' errors are those of their own arithmetic.
    (A sum over a field whose FP32 values are rounded differently, e.g. f*f + ... contracted
    into an FMA, adds a random walk of per-voxel rounding differences that the rule, a
    comparison with one other random walk, does not bound.)"""
    f = field(shape, 31)
    dx = (0.5, 1.0, 0.25)
    g, sink, *_ = sp.compile_expression("laplacian(f)", dx=dx, boundary="periodic", field_map={"f": f})
    reds = [g.add("integrate", (sink,), {"dx": dx, "method": "simpson"}),
            g.add("integrate", (sink,), {"dx": dx, "method": "trapezoid"}),
            g.add("integrate", (sink,), {}),
            g.add("min_max", (sink,), {}), g.add("variance", (sink,), {})]
    rt = sp.Runtime(device="cuda", fusion_enabled=True)
    res = rt.fused.run(g, rt, wanted=[sink] + reds)
    fused_reds = {r for _, _, comp in rt.fused._plans.values() for cg in comp.values()
                  for r in cg.group.reductions}
    assert fused_reds == set(reds)
    o64 = graph_eval.evaluate(g, torch.float64)
    o32 = graph_eval.evaluate(g, torch.float32)
    for nid in [sink] + reds:
        assert_2x(res[nid], o64[nid], o32[nid], f"{shape} {nid}")

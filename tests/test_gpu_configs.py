"""Parity of the benchmarked workloads at the shapes, schedules and code paths that produce the
published numbers, against the FP64 CPU oracle under the 2x error rule.

* 128^3 plain eager: the five single expressions of the 128^3 workload, eager and fused.
* 512^3 stencil graph with the tuned schedule (the tuning database's entry, the one the
  benchmark runs): the whole field and the four reductions, FP32 and BF16 outputs.
* the five-sink tensor graph (eigenvalues, determinant, trace, deviatoric of the hessian, mean
  curvature) on a mixed field, lower half Fourier and upper half i^2 + j^2 + k^2, with the
  768^3 workload's schedule (32, 8, 32, 128), checked directly and on hipGraph replays.
* a forced split tail on a determinant + 9-channel output + reduction graph.
* seeded host scalars: one graph run with two different seeded values.
"""
import numpy as np
import pytest
import torch

import stenpy_amd as sp
from stenpy_amd import fused as F, gen, tune
from stenpy_amd.compiler import compile_expressions
from stenpy_amd.fused import Schedule
from oracle import graph_eval
from tests._util import assert_2x, field, vfield

pytestmark = pytest.mark.gpu

STENCIL_GRAPH = "divergence(gradient(f)) + dot(u, gradient(f)) + sqnorm(curl(u))"
TENSOR_EXPRS = ["eigenvalues(hessian(f))", "determinant(hessian(f))", "trace(hessian(f))",
                "deviatoric(hessian(f))", "mean_curvature(f)"]


def check_all(g, res, ids, what, bf16_sink=None):
    torch.set_num_threads(max(1, torch.get_num_threads()))
    o64 = graph_eval.evaluate(g, torch.float64)
    o32 = graph_eval.evaluate(g, torch.float32)
    out = {}
    for nid in ids:
        slack = 0.0
        if nid == bf16_sink:
            # BF16 storage: half a BF16 ulp of the largest magnitude on top of the 2x rule
            slack = 0.5 * 2.0 ** -7 * 2.0 ** torch.floor(torch.log2(o64[nid].abs().max())).item()
        out[nid] = assert_2x(res[nid], o64[nid], o32[nid], f"{what} {g.node(nid).op_name}", slack)
    return out


def group_of(rt, sink):
    for steps, info, comp in rt.fused._plans.values():
        for cg in comp.values():
            if sink in cg.group.outputs:
                return cg
    raise AssertionError("no fused group produces the sink")


@pytest.mark.parametrize("path", ["eager", "fused"])
def test_plain_eager_workload_128(path):
    f = field((128, 128, 128), 1)
    rt = sp.Runtime(device="cuda", fusion_enabled=(path == "fused"))
    for expr in ["f**2 + tanh(f)*exp(f)", "laplacian(f)", "gradient(f)", "mean(f)", "norm_l2(f)"]:
        g, sink, *_ = sp.compile_expression(expr, boundary="periodic", field_map={"f": f})
        res = rt.run(g)
        check_all(g, res, [sink], f"128^3 {path} {expr}")


@pytest.mark.parametrize("out_dtype", [None, torch.bfloat16])
def test_stencil_graph_512_tuned_schedule(out_dtype):
    n = 512
    dev = torch.device("cuda:0")
    f = gen.generate((n, n, n), 2, dev)
    u = gen.generate_vector((n, n, n), (3, 4, 5), dev)
    g, sink, *_ = sp.compile_expression(STENCIL_GRAPH, dx=1.0, boundary="periodic",
                                        field_map={"f": f, "u": u})
    reds = [g.add(r, (sink,), {}) for r in ("sum", "mean", "variance", "norm_l2")]
    rt = sp.Runtime(device=dev, fusion_enabled=True, schedule_fn=tune.schedule_fn(dev))
    res = rt.fused.run(g, rt, wanted=[sink] + reds, out_dtype=out_dtype)
    torch.cuda.synchronize()
    cg = group_of(rt, sink)
    # the schedule the benchmark runs: the tuning database's entry for this group
    tuned = tune.load_db().get(tune.group_signature(cg.group, g))
    assert tuned and cg.gen.sched.as_tuple() == tuple(tuned), (cg.gen.sched.as_tuple(), tuned)
    res = {k: (v.cpu() if isinstance(v, torch.Tensor) else v) for k, v in res.items()}
    del u
    torch.cuda.empty_cache()
    check_all(g, res, [sink] + reds, f"512^3 {out_dtype}",
              bf16_sink=sink if out_dtype is not None else None)


def mixed_field(shape, dev="cuda"):
    """Lower half (axis 0) the seeded Fourier field (seed 8), upper half i^2 + j^2 + k^2."""
    f = gen.generate(shape, 8, dev)
    h = shape[0] // 2
    f[h:] = torch.from_numpy(gen.quadratic_host(shape, h, shape[0] - h)).to(dev)
    return f


def test_tensor_graph_mixed_field_768_schedule():
    # large enough for the band edge mode and the face wave, the 768^3 workload's kernel path
    shape = (96, 96, 160)
    f = mixed_field(shape)
    g, sinks, *_ = compile_expressions(TENSOR_EXPRS, boundary="periodic", field_map={"f": f})
    rt = sp.Runtime(device="cuda", fusion_enabled=True)
    sched = Schedule(32, 8, 32, 128)
    res = rt.fused.run(g, rt, wanted=sinks, schedule=sched)
    cg = group_of(rt, sinks[0])
    assert all(s in cg.group.outputs for s in sinks), "the five sinks share one kernel"
    assert cg.gen.edge_mode == "band" and cg.gen.fw, "band mode with a face wave"
    # the register rule keeps the main kernel at 4 waves per SIMD with the 9-channel deviatoric
    # written through the per-wave LDS transpose (staged stores)
    regs = F.kernel_registers(F.compile_source(cg.source)[0])
    assert regs["fused_kernel"] <= F.MERGED_TAIL_MAX_VGPRS and cg.gen.stage, regs
    errs = check_all(g, res, sinks, "tensor graph")
    # the exactly degenerate half: H = 2I, so eigenvalues (2,2,2), trace 6, deviatoric 0
    h = shape[0] // 2 + 2
    ev = res[sinks[0]][h:-2, 2:-2, 2:-2]
    assert torch.equal(ev, torch.full_like(ev, 2.0))
    assert torch.equal(res[sinks[2]][h:-2, 2:-2, 2:-2], torch.full_like(ev[..., 0], 6.0))
    assert torch.equal(res[sinks[3]][h:-2, 2:-2, 2:-2], torch.zeros_like(res[sinks[3]][h:-2, 2:-2, 2:-2]))
    eager = {k: res[k].clone() for k in sinks}
    cap, cres = rt.fused.capture(g, rt, sinks, schedule=sched)
    for rep in range(3):
        for k in sinks:
            cres[k].fill_(float("nan"))
        torch.cuda.synchronize()
        cap.replay()
        torch.cuda.synchronize()
        for k in sinks:
            assert torch.equal(cres[k], eager[k]), (rep, g.node(k).op_name)
        torch.randn(1 << 20, device="cuda").sum().item()     # unrelated eager work


def test_tensor_graph_mixed_field_neumann_tuned():
    """The same five sinks with one-sided (Neumann) edges and the shape heuristic's schedule."""
    shape = (36, 30, 66)
    f = mixed_field(shape)
    g, sinks, *_ = compile_expressions(TENSOR_EXPRS, boundary="neumann", field_map={"f": f})
    rt = sp.Runtime(device="cuda", fusion_enabled=True)
    res = rt.fused.run(g, rt, wanted=sinks)
    check_all(g, res, sinks, "tensor graph neumann")


def test_forced_split_tail_with_flag_wide_output_and_reduction(monkeypatch):
    """Edge band in its own ``fused_tail`` (plus ``fused_flag_tail`` for the determinant flag)
    on the side stream: outputs, reductions and the flag agree with the oracle, directly and on
    hipGraph replays."""
    monkeypatch.setattr(F, "SPLIT_TAIL", "1")
    shape = (30, 34, 70)
    f = field(shape, 8)
    for bc in ("periodic", "neumann"):
        g, sinks, *_ = compile_expressions(["determinant(hessian(f))", "deviatoric(hessian(f))",
                                            "laplacian(f) * f"], boundary=bc, field_map={"f": f})
        reds = [g.add("sum", (sinks[2],), {}), g.add("norm_l2", (sinks[0],), {})]
        wanted = sinks + reds
        rt = sp.Runtime(device="cuda", fusion_enabled=True)
        res = rt.fused.run(g, rt, wanted=wanted)
        cg = group_of(rt, sinks[0])
        assert cg.gen.split_tail and cg.gen.boundary_blocks() > 0
        check_all(g, res, wanted, f"split tail {bc}")
        eager = {k: res[k].clone() for k in wanted}
        cap, cres = rt.fused.capture(g, rt, wanted)
        for rep in range(3):
            for k in wanted:
                cres[k].fill_(float("nan"))
            torch.cuda.synchronize()
            cap.replay()
            torch.cuda.synchronize()
            for k in wanted:
                assert torch.equal(cres[k], eager[k]), (bc, rep, g.node(k).op_name)


def test_seeded_scalars_do_not_reuse_a_plan_compiled_for_another_value():
    """``Runtime.run(graph, seed)``: a host scalar seeded for a node is compiled into the kernel,
    so a second run with another value must not reuse the first plan."""
    shape = (20, 22, 40)
    f = field(shape, 3)
    g, sink, _, fm, _ = sp.compile_expression("laplacian(f) * c + f", boundary="periodic",
                                              field_map={"f": f, "c": torch.tensor(1.0, dtype=torch.float64)})
    rt = sp.Runtime(device="cuda", fusion_enabled=True)
    for c in (2.0, 3.0, torch.tensor(-0.5)):
        res = rt.run(g, seed={fm["c"]: c})
        o64 = graph_eval.evaluate(g, torch.float64, seed={fm["c"]: c})
        o32 = graph_eval.evaluate(g, torch.float32, seed={fm["c"]: c})
        assert_2x(res[sink], o64[sink], o32[sink], f"seeded c={c}")


def test_captured_graph_keeps_its_kernels_alive(monkeypatch):
    """A captured fused evaluation replays correctly after the runtime and executor that made
    it are gone: the returned graph holds the modules of every kernel it launches (main, tail
    and flag kernels)."""
    import gc
    monkeypatch.setattr(F, "SPLIT_TAIL", "1")
    shape = (24, 26, 70)
    f = field(shape, 9)
    g, sinks, *_ = compile_expressions(["determinant(hessian(f))", "laplacian(f)"],
                                       boundary="periodic", field_map={"f": f})
    rt = sp.Runtime(device="cuda", fusion_enabled=True)
    eager = {k: v.clone() for k, v in rt.fused.run(g, rt, wanted=sinks).items() if k in sinks}
    cap, res = rt.fused.capture(g, rt, sinks)
    del rt
    gc.collect()
    for k in sinks:
        res[k].fill_(float("nan"))
    torch.cuda.synchronize()
    cap.replay()
    torch.cuda.synchronize()
    for k in sinks:
        assert torch.equal(res[k], eager[k]), g.node(k).op_name

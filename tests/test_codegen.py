"""Fused planner and hipRTC codegen on the CPU: plans have the expected shape and every
generated kernel compiles for gfx950 (hipRTC needs no GPU)."""
import pytest
import torch

import stenpy_amd as sp
from stenpy_amd import fused as F

STENCIL_GRAPH = "divergence(gradient(f)) + dot(u, gradient(f)) + sqnorm(curl(u))"
SHAPE = (20, 24, 30)


def plan(expr, bc="periodic", extra=(), shape=SHAPE):
    f, u = torch.zeros(shape), torch.zeros(shape + (3,))
    g, sink, *_ = sp.compile_expression(expr, boundary=bc, field_map={"f": f, "u": u})
    reds = [g.add(r, (sink,), {}) for r in extra]
    values = {n: g.node(n).params["value"] for n in g._order if g.node(n).op_name == "_constant"}
    ex = F.FusedExecutor(torch.device("cuda", 0))
    steps, info, comp = ex.prepare(g, values, [sink] + reds)
    return g, sink, reds, steps, comp


def test_stencil_graph_is_one_kernel_with_fused_reductions():
    g, sink, reds, steps, comp = plan(STENCIL_GRAPH, extra=("sum", "mean", "variance", "norm_l2"))
    assert [s.kind for s in steps] == ["group"]
    grp = steps[0].group
    assert grp.outputs == [sink] and sorted(grp.reductions) == sorted(reds)
    assert len(grp.sources) == 2 and grp.dscalars == []
    cg = comp[0]
    assert cg.gen.hz == 1 and [l["H"] for l in cg.gen.src_layout] == [1, 1]
    assert cg.gen.lds_bytes() <= F.LDS_LIMIT


def test_reduction_result_starts_a_new_level():
    g, sink, reds, steps, comp = plan("f - mean(f)")
    kinds = [s.kind for s in steps]
    assert kinds.count("group") == 2
    second = [s.group for s in steps if s.kind == "group"][1]
    assert len(second.dscalars) == 1


def test_periodic_over_one_sided_widens_the_halo():
    *_, comp = plan("mean_curvature(f)", bc="periodic")
    assert comp[0].gen.src_layout[0]["H"] == 3
    # band mode (a grid where the band is a small share): rows above and below the tile need
    # only the interior reach (the face wave's wrapped one-sided formulas run along x)
    *_, comp = plan("mean_curvature(f)", bc="periodic", shape=(64, 96, 128))
    assert comp[0].gen.edge_mode == "band"
    assert comp[0].gen.src_layout[0]["H"] == 3 and comp[0].gen.src_layout[0]["HY"] == 2
    *_, comp = plan("mean_curvature(f)", bc="neumann")
    assert comp[0].gen.src_layout[0]["H"] == 2


def test_deep_composition_is_cut():
    g, sink, reds, steps, comp = plan("laplacian(laplacian(laplacian(f)))", bc="neumann")
    assert sum(1 for s in steps if s.kind == "group") == 2


@pytest.mark.parametrize("expr,bc", [
    ("f**2 + tanh(f)*exp(f)", "periodic"), ("laplacian(f)", "periodic"), (STENCIL_GRAPH, "periodic"),
    ("laplacian(laplacian(f))", "neumann"), ("divergence(curl(u)) + laplacian(exp(f))", "dirichlet"),
    ("gradient(f) * 2 + curl(u)", "reflect"), ("eigenvalues(hessian(f))", "periodic"),
    ("determinant(hessian(f)) + trace(hessian(f))", "neumann"), ("mean_curvature(f)", "periodic"),
    ("jacobian(curl(u))", "periodic"), ("integrate(f*f)", "periodic"), ("min_max(f)", "periodic"),
])
def test_generated_kernels_compile(expr, bc):
    *_, comp = plan(expr, bc)
    for cg in comp.values():
        code, _ = F.compile_source(cg.source)
        assert len(code) > 1000


@pytest.mark.parametrize("edge_mode", ["band", "variants"])
@pytest.mark.parametrize("shape,sched", [((20, 24, 30), F.Schedule(32, 8, 16, 256)),
                                         ((20, 24, 64), F.Schedule(32, 8, 16, 256)),
                                         ((12, 40, 128), F.Schedule(64, 8, 64, 256)),
                                         ((9, 8, 130), F.Schedule(64, 8, 64, 256)),
                                         ((21, 37, 70), F.Schedule(16, 4, 64, 64)),
                                         ((40, 45, 150), F.Schedule(128, 4, 32, 256))])
def test_main_and_boundary_kernels_cover_every_voxel_once(shape, sched, edge_mode):
    """Voxels the main kernel evaluates plus the boundary boxes tile the grid exactly."""
    f, u = torch.zeros(shape), torch.zeros(shape + (3,))
    g, sink, *_ = sp.compile_expression(STENCIL_GRAPH, boundary="neumann", field_map={"f": f, "u": u})
    values = {n: g.node(n).params["value"] for n in g._order if g.node(n).op_name == "_constant"}
    steps, info = F.plan_graph(g, values, [sink])
    gen = F.KernelGen(g, steps[0].group, info, sched, {}, {}, edge_mode)
    assert gen.edge_mode == edge_mode
    n0, n1, n2 = shape
    hits = torch.zeros(shape, dtype=torch.int32)
    z0, z1, y0, y1, x0, x1 = gen.interior_bounds()
    # the main kernel: full tiles only, its boxes within them
    for (c0, c1, b0, b1, a0, a1) in gen.main_boxes():
        for ty in range(0, n1, sched.ty):
            for tx in range(0, n2, sched.tx):
                if (ty + sched.ty > n1 or tx + sched.tx > n2) and edge_mode != "band":
                    continue        # partial tiles: tail blocks (band mode: the main kernel)
                ya, yb = max(ty, b0), min(ty + sched.ty, b1)
                xa, xb = max(tx, a0), min(tx + sched.tx, a1)
                if ya < yb and xa < xb and c0 < c1:
                    hits[c0:c1, ya:yb, xa:xb] += 1
    for (a0, a1, b0, b1, c0, c1) in gen.boundary_boxes():
        hits[a0:a1, b0:b1, c0:c1] += 1
    assert int(hits.min()) == 1 and int(hits.max()) == 1
    # the interior box keeps every voxel at least the edge reach away from each face
    r = gen.redge
    assert z0 >= r and z1 <= n0 - r or z0 >= z1
    assert y0 >= min(r, n1) * (edge_mode == "band") and x0 >= min(r, n2) * (edge_mode == "band")
    assert gen.boundary_blocks() * gen.block_threads() * gen.TAIL_VPT >= sum((b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4])
                                              for b in gen.boundary_boxes())


def test_load_transforms_detected_only_where_they_pay():
    cases = {"laplacian(tanh(f))": ["tanh"], "laplacian(tanh(f)) * tanh(f)": ["tanh"],
             "gradient(exp(f*2))": ["mul", "exp"], "hessian(log(f*f+1))": ["mul", "add", "log"],
             "laplacian(f)": None, "tanh(f) + laplacian(f)": None, "laplacian(f*2)": None}
    for expr, chain in cases.items():
        g, sink, reds, steps, comp = plan(expr, bc="neumann")
        gen = comp[0].gen
        got = [g.node(n).op_name for n in gen.xform[0][0]] if gen.xform else None
        assert got == chain, (expr, got)


@pytest.mark.parametrize("sched,staged", [((128, 8, 64, 256), True), ((64, 4, 64, 256), True),
                                          ((32, 8, 64, 128), True), ((16, 8, 64, 128), True),
                                          ((32, 8, 64, 128), False)])
def test_wide_outputs_staged_through_lds(sched, staged, monkeypatch):
    # 9-channel hessian and 3-channel gradient: only outputs of >= 4 channels are staged; a
    # wave of a tile narrower than 64 covers 64 / TX consecutive rows
    monkeypatch.setattr(F, "STAGE_STORES", staged)
    f = torch.zeros((16, 24, 200))
    g, hes, *_ = sp.compile_expression("hessian(f)", boundary="periodic", field_map={"f": f})
    grad = g.add("gradient", (g.node(hes).input_ids[0],), dict(g.node(hes).params))
    values = {n: g.node(n).params["value"] for n in g._order if g.node(n).op_name == "_constant"}
    steps, info = F.plan_graph(g, values, [hes, grad])
    gen = F.KernelGen(g, steps[0].group, info, F.Schedule(*sched), {}, {})
    outs = steps[0].group.outputs
    assert gen.stage == ([outs.index(hes)] if staged else [])
    src = gen.source()
    assert ("so_stage" in src) == staged
    F.compile_source(src)


@pytest.mark.parametrize("expr", ["laplacian(f) + tanh(f)",
                                  "eigenvalues(hessian(f)) * 1.5 + mean_curvature(f) + determinant(hessian(f))"])
def test_tail_blocks_join_the_main_grid_unless_registers_grow(expr):
    # the plan keeps the edge band's blocks in the main launch unless that code pushes the main
    # kernel past the register limit (read from the code object's metadata, no GPU needed)
    g, sink, reds, steps, comp = plan(expr, bc="periodic", shape=(40, 48, 96))
    gen = comp[0].gen
    merged = F.KernelGen(g, steps[0].group, gen.info, gen.sched, gen.out_dtypes, {}, gen.edge_mode)
    merged.want_split = False
    regs = F.kernel_registers(F.compile_source(merged.source())[0])
    assert regs["fused_kernel"] > 0 and "fused_tail" not in regs
    assert gen.split_tail == (regs["fused_kernel"] > F.MERGED_TAIL_MAX_VGPRS)
    if gen.split_tail:
        assert F.kernel_registers(F.compile_source(comp[0].source)[0])["fused_tail"] > 0


def test_register_rule_splits_the_tail_before_dropping_staging(monkeypatch):
    # the five-sink tensor group (band edge mode) is staged with its tail in the main grid; the
    # rule keeps that when the kernel fits the register limit, and otherwise first moves the edge
    # band to its own kernel, then stores directly.  The limit is lowered to exercise each step.
    f = torch.zeros((96, 96, 160))
    exprs = ["eigenvalues(hessian(f))", "determinant(hessian(f))", "trace(hessian(f))",
             "deviatoric(hessian(f))", "mean_curvature(f)"]
    from stenpy_amd.compiler import compile_expressions
    g, sinks, *_ = compile_expressions(exprs, boundary="periodic", field_map={"f": f})
    values = {n: g.node(n).params["value"] for n in g._order if g.node(n).op_name == "_constant"}
    steps, info = F.plan_graph(g, values, sinks)
    regs = lambda s: F.kernel_registers(F.compile_source(s)[0])["fused_kernel"]

    def fitted(limit):
        monkeypatch.setattr(F, "MERGED_TAIL_MAX_VGPRS", limit)
        gen = F.KernelGen(g, steps[0].group, info, F.Schedule(32, 8, 16, 128), {}, {})
        assert gen.stage, "the 9-channel deviatoric is eligible for staging"
        return gen, gen.source(), F.FusedExecutor._fit_registers(gen, gen.source())

    gen, before_src, src = fitted(128)
    before = regs(before_src)
    if before <= 128:
        assert src == before_src and gen.stage and not gen.split_tail
    # a limit just below the merged, staged kernel: the tail moves out first, staging stays
    gen, _, src = fitted(before - 1)
    split_regs = regs(src)
    assert gen.split_tail and "fused_tail" in F.kernel_registers(F.compile_source(src)[0])
    if split_regs <= before - 1:
        assert gen.stage
    # a limit below the split, staged kernel: staging goes too
    gen, _, src = fitted(min(split_regs, before) - 1)
    assert gen.split_tail and not gen.stage
    # a light staged group keeps both its staging and its in-grid tail
    monkeypatch.setattr(F, "MERGED_TAIL_MAX_VGPRS", 128)
    g2, s2, *_ = compile_expressions(["hessian(f)"], boundary="periodic", field_map={"f": f})
    values = {n: g2.node(n).params["value"] for n in g2._order if g2.node(n).op_name == "_constant"}
    steps2, info2 = F.plan_graph(g2, values, s2)
    gen2 = F.KernelGen(g2, steps2[0].group, info2, F.Schedule(32, 8, 16, 128), {}, {})
    F.FusedExecutor._fit_registers(gen2, gen2.source())
    assert gen2.stage and not gen2.split_tail


def test_schedule_features_count_each_transcendental_call():
    """The schedule model's ``transc`` feature counts the math calls the emitter writes, one per
    op (a renamed math function would otherwise silently zero it)."""
    from stenpy_amd import tune
    for expr, want in [("tanh(f)*exp(f) + laplacian(f)", 2), ("sin(f) + cos(f) + log(f*f)", 3),
                       ("laplacian(f) * 2", 0)]:
        g, sink, reds, steps, comp = plan(expr)
        assert sum(tune.group_stats(c.gen)["transc"] for c in comp.values()) == want, expr

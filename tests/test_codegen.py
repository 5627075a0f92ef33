"""Fused planner and hipRTC codegen on the CPU: plans have the expected shape and every
generated kernel compiles for gfx950 (hipRTC needs no GPU)."""
import pytest
import torch

import stenpy_amd as sp
from stenpy_amd import fused as F

C2 = "divergence(gradient(f)) + dot(u, gradient(f)) + sqnorm(curl(u))"
SHAPE = (20, 24, 30)


def plan(expr, bc="periodic", extra=(), shape=SHAPE):
    f, u = torch.zeros(shape), torch.zeros(shape + (3,))
    g, sink, *_ = sp.compile_expression(expr, boundary=bc, field_map={"f": f, "u": u})
    reds = [g.add(r, (sink,), {}) for r in extra]
    values = {n: g.node(n).params["value"] for n in g._order if g.node(n).op_name == "_constant"}
    ex = F.FusedExecutor(torch.device("cuda", 0))
    steps, info, comp = ex.prepare(g, values, [sink] + reds)
    return g, sink, reds, steps, comp


def test_c2_is_one_kernel_with_fused_reductions():
    g, sink, reds, steps, comp = plan(C2, extra=("sum", "mean", "variance", "norm_l2"))
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
    *_, comp = plan("mean_curvature(f)", bc="neumann")
    assert comp[0].gen.src_layout[0]["H"] == 2


def test_deep_composition_is_cut():
    g, sink, reds, steps, comp = plan("laplacian(laplacian(laplacian(f)))", bc="neumann")
    assert sum(1 for s in steps if s.kind == "group") == 2


@pytest.mark.parametrize("expr,bc", [
    ("f**2 + tanh(f)*exp(f)", "periodic"), ("laplacian(f)", "periodic"), (C2, "periodic"),
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

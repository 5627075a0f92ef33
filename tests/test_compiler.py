"""SymPy -> Graph compiler and Graph IR (CPU).

Node lists follow the reference compiler's lowering rules (stenpy_engine.py:3618-3759) for
the benchmark expressions, with SymPy 1.14's canonical argument order.
"""
import pytest
import sympy
import torch

import stenpy_amd as sp
from stenpy_amd.compiler import _decode_kw_value, _encode_kwargs
from stenpy_amd.graph import Graph


def ops_of(g):
    return [(n.op_name, tuple(g.node(i).op_name for i in n.input_ids)) for n in g.topological_sort()]


def fmap(**kw):
    return {k: torch.zeros(4, 4, 4) for k in kw} if kw else {"f": torch.zeros(4, 4, 4)}


def test_sympy_version_pinned():
    assert sympy.__version__ == "1.14.0"


def test_c1_pointwise_expression():
    g, sink, w, fm, e = sp.compile_expression("f**2 + tanh(f)*exp(f)", field_map=fmap())
    assert ops_of(g) == [("_constant", ()), ("mul", ("_constant", "_constant")),
                         ("exp", ("_constant",)), ("tanh", ("_constant",)), ("mul", ("exp", "tanh")),
                         ("add", ("mul", "mul"))]
    assert g.node(sink).op_name == "add"
    assert all(n.params == {} for n in g.nodes.values() if n.op_name != "_constant")


def test_stencil_params_and_defaults():
    g, sink, _, _, _ = sp.compile_expression("laplacian(f)", field_map=fmap())
    assert g.node(sink).params == {"dx": 1.0, "boundary": "neumann"}
    g, sink, _, _, _ = sp.compile_expression("lap(f)", dx=0.5, boundary="periodic", field_map=fmap())
    assert g.node(sink).op_name == "laplacian" and g.node(sink).params == {"dx": 0.5, "boundary": "periodic"}


def test_c4_expression_order():
    g, sink, _, _, _ = sp.compile_expression("f**2 + tanh(f)*exp(f) + laplacian(f)", field_map=fmap())
    s = g.node(sink)
    assert s.op_name == "add"
    assert g.node(s.input_ids[0]).op_name == "add" and g.node(s.input_ids[1]).op_name == "laplacian"


def test_c2_expression_rewrite_cse_and_dce():
    g, sink, w, fm, e = sp.compile_expression(
        "divergence(gradient(f)) + dot(u, gradient(f)) + sqnorm(curl(u))",
        boundary="periodic", field_map=fmap(f=1, u=1))
    assert any("Simplified divergence(gradient(f))" in x for x in w)
    names = sorted(n.op_name for n in g.nodes.values())
    # one gradient (shared by CSE), laplacian from the rewrite, no divergence; DCE kept the
    # gradient because dot still uses it
    assert names.count("gradient") == 1 and "divergence" not in names and "laplacian" in names
    # SymPy order: Add(divergence(...), dot(...), sqnorm(...)) folded left
    s = g.node(sink)
    left = g.node(s.input_ids[0])
    assert g.node(s.input_ids[1]).op_name == "sqnorm"
    assert [g.node(i).op_name for i in left.input_ids] == ["laplacian", "dot"]


def test_dead_gradient_removed():
    g, sink, *_ = sp.compile_expression("divergence(gradient(f))", field_map=fmap())
    assert sorted(n.op_name for n in g.nodes.values()) == ["_constant", "laplacian"]
    g2, sink2, *_ = sp.compile_expression("divergence(gradient(f))", field_map=fmap(), eliminate_dead=False)
    assert "gradient" in [n.op_name for n in g2.nodes.values()]


def test_kwargs_encoding():
    s = _encode_kwargs("clamp(f, lo=-0.5, hi=2)")
    assert s == f"clamp(f, _kw_lo__{'-0.5'.encode().hex()}, _kw_hi__{'2'.encode().hex()})"
    assert _decode_kw_value("2".encode().hex()) == 2
    assert _decode_kw_value("0.25".encode().hex()) == 0.25
    assert _decode_kw_value("abc".encode().hex()) == "abc"
    g, sink, *_ = sp.compile_expression("clamp(f, lo=-0.5, hi=2)", field_map=fmap())
    assert g.node(sink).params == {"lo": -0.5, "hi": 2}
    g, sink, *_ = sp.compile_expression("gradient(f, dim=1)", field_map=fmap())
    assert g.node(sink).params == {"dx": 1.0, "boundary": "neumann", "dim": 1}


def test_power_rules():
    g, sink, *_ = sp.compile_expression("f**3", field_map=fmap())
    assert [n.op_name for n in g.topological_sort()] == ["_constant", "mul", "mul"]
    g, sink, *_ = sp.compile_expression("f**0.5", field_map=fmap())
    assert g.node(sink).op_name == "sqrt"
    g, sink, *_ = sp.compile_expression("f**-2", field_map=fmap())
    s = g.node(sink)
    assert s.op_name == "div" and g.node(s.input_ids[0]).op_name == "_constant"
    g, sink, *_ = sp.compile_expression("f**2.5", field_map=fmap())
    s = g.node(sink)
    assert s.op_name == "exp" and g.node(s.input_ids[0]).op_name == "mul"
    g, sink, *_ = sp.compile_expression("-f", field_map=fmap())
    assert g.node(sink).op_name == "neg"
    g, sink, *_ = sp.compile_expression("f - laplacian(f)", field_map=fmap())
    assert g.node(sink).op_name == "add"


def test_constants_shared_and_symbols():
    g, sink, w, *_ = sp.compile_expression("2*f + 2*laplacian(f) + q", field_map=fmap())
    consts = [n for n in g.nodes.values() if n.op_name == "_constant" and n.params["value"].ndim == 0]
    assert len(consts) == 1 and float(consts[0].params["value"]) == 2.0
    assert any("Unknown symbol 'q'" in x for x in w)
    g, sink, w, *_ = sp.compile_expression("g + h", field_map=fmap())
    assert not w  # f g h u v alias the primary field silently


def test_aliases_and_registered_names():
    for expr, op in [("grad(f)", "gradient"), ("div(gradient(f))", "laplacian"), ("var(f)", "variance"),
                     ("minmax(f)", "min_max"), ("hessian(f)", "hessian"), ("cos(f)", "cos"),
                     ("mean_curvature(f)", "mean_curvature")]:
        g, sink, *_ = sp.compile_expression(expr, field_map=fmap())
        assert g.node(sink).op_name == op, expr


def test_every_registered_op_accepts_the_params_the_compiler_gives_it():
    """Ops outside the no-stencil set receive dx and boundary; each must accept them, so that
    every registered name is callable from an expression (cos, fft_complex and ifft_complex
    included, which the reference's set omits)."""
    import inspect
    from stenpy_amd.compiler import _NO_STENCIL_OPS
    from stenpy_amd.registry import OP_REGISTRY
    bad = []
    for name, fn in OP_REGISTRY.items():
        if name in _NO_STENCIL_OPS or name == "_constant":
            continue
        ps = inspect.signature(fn).parameters
        if not (("dx" in ps and "boundary" in ps) or any(p.kind == p.VAR_KEYWORD for p in ps.values())):
            bad.append(name)
    assert not bad, bad
    g, sink, *_ = sp.compile_expression("cos(f) + fft_complex(f)", boundary="periodic", field_map=fmap())
    assert all(not n.params for n in g.nodes.values() if n.op_name in ("cos", "fft_complex"))


def test_errors():
    with pytest.raises(ValueError, match="Unknown operator"):
        sp.compile_expression("nosuchop(f)", field_map=fmap())
    with pytest.raises(ValueError, match="Unsupported SymPy node"):
        sp.compile_expression("f*pi", field_map=fmap())
    with pytest.raises(ValueError, match="Could not parse"):
        sp.compile_expression("f +* 2", field_map=fmap())


def test_graph_topology_and_clone():
    g = Graph()
    a = g.add("_constant", (), {"value": torch.zeros(2)})
    b = g.add("exp", (a,))
    c = g.add("add", (a, b))
    assert [n.id for n in g.topological_sort()] == [a, b, c]
    assert g.sinks() == [c]
    assert g.build_consumer_map()[a] == [b, c]
    h = g.clone_with_replacement(a, torch.ones(3))
    assert list(h._order) == list(g._order) and h.node(a).params["value"].shape == (3,)
    g.add("neg", ("missing",))
    with pytest.raises(ValueError, match="unknown input"):
        g.topological_sort()
    g2 = Graph()
    g2.add("neg", ("y",), node_id="x")
    g2.add("neg", ("x",), node_id="y")
    with pytest.raises(ValueError, match="cycle"):
        g2.topological_sort()


def test_canonical_key_distinguishes_structure():
    f = torch.zeros(4, 4, 4)
    k1 = sp.compile_expression("laplacian(f) + f", field_map={"f": f})[0].canonical_key()
    k2 = sp.compile_expression("laplacian(f) * f", field_map={"f": f})[0].canonical_key()
    k3 = sp.compile_expression("laplacian(f) + f", boundary="periodic", field_map={"f": f})[0].canonical_key()
    k4 = sp.compile_expression("laplacian(f) + f", field_map={"f": f})[0].canonical_key()
    assert len({k1, k2, k3}) == 3 and k1 == k4


def test_operator_dispatch_returns_pipeline():
    p = sp.gradient("f**2")
    assert isinstance(p, sp.Pipeline) and p._expr_str == "f**2" and p._op_name == "gradient"
    assert repr(p).startswith("Pipeline(gradient(f**2)")
    with pytest.raises(AttributeError):
        sp.no_such_operator


def test_keyword_arguments_need_a_call_without_nested_parentheses():
    """The front-end encodes ``k=v`` only inside innermost calls (engine 3525-3543): bounds on
    a bare symbol parse, bounds on a nested call are rejected as unknown sympify options."""
    f = torch.zeros(4, 4, 4, dtype=torch.float64)
    g, sink, *_ = sp.compile_expression("clamp(f, lo=-1, hi=1)", field_map={"f": f})
    node = g.node(sink)
    assert node.op_name == "clamp" and node.params["lo"] == -1 and node.params["hi"] == 1
    with pytest.raises(ValueError):
        sp.compile_expression("clamp(laplacian(f), lo=-1, hi=1)", field_map={"f": f})


def test_compile_expressions_shares_subexpressions_across_sinks():
    from stenpy_amd.compiler import compile_expressions
    f = torch.zeros((6, 7, 8))
    exprs = ["eigenvalues(hessian(f))", "trace(hessian(f))", "mean_curvature(f)", "laplacian(f) + 1"]
    g, sinks, _, fm = compile_expressions(exprs, boundary="periodic", field_map={"f": f})
    ops = [g.node(n).op_name for n in g._order]
    assert ops.count("hessian") == 1 and ops.count("_constant") == 2
    assert [g.node(s).op_name for s in sinks] == ["eigenvalues", "trace", "mean_curvature", "add"]
    # each sink is the same node structure the single-expression compiler builds
    for e, s in zip(exprs, sinks):
        g1, s1, *_ = sp.compile_expression(e, boundary="periodic", field_map={"f": f})
        assert g1.node(s1).op_name == g.node(s).op_name and g1.node(s1).params == g.node(s).params


def test_random_pointwise_expressions_lower_to_their_values():
    """Seeded random point-wise expressions: the compiled graph, evaluated by the FP64 oracle,
    equals the expression evaluated directly by SymPy's numpy printer.  Field values lie in
    [0.5, 1.5], so the eps clamps of div and log never act."""
    import random
    import numpy as np
    rng = random.Random(11)
    f = torch.rand(3, 4, 5, dtype=torch.float64) + 0.5
    from oracle import graph_eval

    def rand_expr(depth):
        if depth == 0 or rng.random() < 0.25:
            return rng.choice(["f", f"{rng.uniform(0.5, 2.0):.3f}"])
        kind = rng.random()
        if kind < 0.4:
            op = rng.choice(["exp", "tanh", "sin", "cos", "sqrt", "log"])
            inner = rand_expr(depth - 1)
            if op in ("sqrt", "log"):
                inner = f"({inner})**2 + 0.5"
            if op == "exp":
                inner = f"tanh({inner})"
            return f"{op}({inner})"
        if kind < 0.55:
            return f"({rand_expr(depth - 1)})**{rng.choice([2, 3, -1, -2, 0.5])}"
        a, b = rand_expr(depth - 1), rand_expr(depth - 1)
        op = rng.choice(["+", "-", "*", "/"])
        if op == "/":
            b = f"(({b})**2 + 0.5)"
        return f"({a}) {op} ({b})"

    checked = 0
    for _ in range(60):
        expr = rand_expr(4)
        if "f" not in expr:
            continue
        g, sink, *_ = sp.compile_expression(expr, field_map={"f": f})
        got = graph_eval.evaluate(g, torch.float64)[sink]
        fn = sympy.lambdify(sympy.Symbol("f"), sympy.sympify(expr), "numpy")
        want = np.broadcast_to(np.asarray(fn(f.numpy()), dtype=np.float64), f.shape)
        np.testing.assert_allclose(got.numpy(), want, rtol=1e-12, atol=1e-12, err_msg=expr)
        checked += 1
    assert checked >= 40


def test_fused_compiler_boundary_names():
    """The reference's fused-path names: FUSIBLE_OPS covers the reference's set
    (fused_compiler.py:194-195) and compile_and_execute rejects a non-HIP device with
    ValueError (fused_compiler.py:634-636), which the reference's Runtime.run turns into its
    eager fallback."""
    from stenpy_amd import fused_compiler as fc
    reference_set = {"add", "sub", "mul", "div", "exp", "log", "sqrt", "sin", "cos", "tanh", "neg",
                     "laplacian", "gradient", "divergence", "curl", "hessian"}
    assert reference_set <= fc.FUSIBLE_OPS
    assert {"dot", "sqnorm", "eigenvalues", "determinant", "sum", "variance"} <= fc.FUSIBLE_OPS

    class CpuRuntime:
        device = torch.device("cpu")
    g, sink, *_ = sp.compile_expression("laplacian(f)", field_map={"f": torch.zeros(4, 4, 4)})
    with pytest.raises(ValueError, match="CUDA"):
        fc.compile_and_execute(g, CpuRuntime(), None)


def test_fused_compiler_reads_a_foreign_graph():
    """compile_and_execute accepts any graph exposing the reference's node table."""
    import collections
    from stenpy_amd import fused_compiler as fc
    Node = collections.namedtuple("Node", "id op_name input_ids params")

    class ForeignGraph:
        def __init__(self, nodes):
            self._nodes = {n.id: n for n in nodes}
    f = torch.zeros(4, 4, 4)
    fg = ForeignGraph([Node("c", "_constant", (), {"value": f}),
                       Node("lap", "laplacian", ("c",), {"dx": 1.0, "boundary": "periodic"}),
                       Node("out", "exp", ("lap",), {})])
    g = fc._as_graph(fg)
    assert [n.id for n in g.topological_sort()] == ["c", "lap", "out"] and g.sinks() == ["out"]
    assert g.node("lap").params == {"dx": 1.0, "boundary": "periodic"}

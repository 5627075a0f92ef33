"""SymPy -> Graph compiler (expression front end).

Behaviour follows the reference compiler (stenpy_engine.py:3458-3778):

* ``name(k=v)`` keyword arguments of innermost calls are hex-encoded into ``_kw_<k>__<hex>``
  symbols before SymPy sees them (stenpy_engine.py:3525-3543) and decoded back into
  int/float/str params (3515-3522, 3596-3610);
* the SymPy namespace declares the reference alias table as undefined functions, the
  reserved coordinates ``x y z t`` (unless shadowed by a field), the fields, ``pi`` and ``E``
  (3546-3573); names outside the table become SymPy auto-functions, so any registered op
  name (``hessian``, ``mean_curvature``, ``dot`` ...) is callable;
* lowering (3618-3759): one ``_constant`` per distinct scalar; ``Add``/``Mul`` fold left into
  binary ``add``/``mul``; ``-1*x`` becomes ``neg``; ``x**0.5`` -> ``sqrt``; integer powers
  ``|n| <= 16`` -> repeated ``mul`` (negative -> ``div(1, .)``); other powers ->
  ``exp(c*log(x))``; ``divergence(gradient(X))`` -> ``laplacian(X)``; stencil ops receive
  ``{dx, boundary}``; unknown symbols ``f g h u v`` alias the primary field; every op node is
  hash-consed on ``(op, child ids, scalar params)``.

Deliberate differences: ``dot`` and ``sqnorm`` are extra point-wise op names (the reference
cannot express ``u . grad f`` and ``|curl u|^2``), and dead nodes (such as the
gradient the divergence(gradient) rewrite leaves behind) are removed after lowering.
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field as dc_field
from typing import Any, Dict, List, Optional, Sequence, Set, Tuple

import sympy as sp
import torch

from .graph import Graph
from .registry import OP_REGISTRY

# Alias table of the expression language (stenpy_engine.py:3462-3475); an interface table,
# so its entries are the reference's.
_ALIAS: Dict[str, str] = {
    "gradient": "gradient", "grad": "gradient", "divergence": "divergence", "div": "divergence",
    "laplacian": "laplacian", "laplace": "laplacian", "lap": "laplacian", "curl": "curl",
    "spectral_gradient": "spectral_gradient", "spectral_laplacian": "spectral_laplacian",
    "fft": "fft", "ifft": "ifft", "add": "add", "sub": "sub", "mul": "mul", "div_op": "div",
    "neg": "neg", "clamp": "clamp", "exp": "exp", "log": "log", "sqrt": "sqrt", "sin": "sin",
    "tanh": "tanh", "sum": "sum", "mean": "mean", "norm_l2": "norm_l2", "variance": "variance",
    "entropy": "entropy", "integrate": "integrate", "cumulative_integral": "cumulative_integral",
    "trace": "trace", "determinant": "determinant", "eigenvalues": "eigenvalues",
    "inverse": "inverse", "deviatoric": "deviatoric", "var": "variance",
    "moving_average": "moving_average", "min_max": "min_max", "minmax": "min_max",
    "covariance": "covariance", "cov": "covariance", "correlation": "correlation",
    "corr": "correlation", "surface_integral": "surface_integral",
    "distance_transform": "distance_transform",
}

# Ops that take no {dx, boundary} params (stenpy_engine.py:3476-3484), plus the point-wise
# extensions of this backend and three ops that set omits although their functions take no
# dx or boundary (cos, fft_complex, ifft_complex: called from an expression there, they raise
# TypeError on the unexpected keywords).
_NO_STENCIL_OPS: Set[str] = {
    "add", "sub", "mul", "div", "neg", "clamp", "exp", "log", "sqrt", "sin", "tanh",
    "sum", "mean", "norm_l2", "min_max", "variance", "entropy", "trace", "determinant",
    "eigenvalues", "inverse", "deviatoric", "fft", "ifft", "covariance", "correlation",
    "moving_average", "surface_integral", "distance_transform", "stack_components",
    "select_component", "norm_last", "div_last", "neg_stack3", "scale_eye",
    "velocity_dot_grad", "integrate", "cumulative_integral", "spectral_gradient",
    "spectral_laplacian",
    "dot", "sqnorm", "cast_bf16", "cos", "fft_complex", "ifft_complex",
}

_SIMPLIFY_RULES: Dict[Tuple[str, str], str] = {("divergence", "gradient"): "laplacian"}
_RESERVED_SYMBOLS = frozenset({"x", "y", "z", "t"})
_PRIMARY_ALIASES = ("f", "g", "h", "u", "v")
_CALL_RE = re.compile(r"(\w+)\s*\(([^()]*?)\)", re.DOTALL)


class UserVar:
    """Named sweep variable declared in the expression namespace (stenpy_engine.py:3499-3510)."""

    def __init__(self, name: str, start: float, end: float, steps: int, current: float = 0.0):
        self.name, self.start, self.end, self.steps, self.current = name, start, end, steps, current

    @property
    def values(self):
        import numpy as np
        return np.linspace(self.start, self.end, self.steps)

    @property
    def step_size(self) -> float:
        return 0.0 if self.steps < 2 else (self.end - self.start) / (self.steps - 1)


_USER_VARS: Dict[str, UserVar] = {}


def _decode_kw_value(enc: str) -> Any:
    try:
        text = bytes.fromhex(enc).decode()
    except Exception:
        text = enc
    for conv in (int, float):
        try:
            return conv(text)
        except ValueError:
            pass
    return text


def _encode_kwargs(s: str) -> str:
    """Rewrite ``k=v`` arguments of innermost calls into ``_kw_k__<hex(v)>`` symbols."""

    def one_call(m: re.Match) -> str:
        positional, encoded = [], []
        for raw in m.group(2).split(","):
            arg = raw.strip()
            if "=" in arg and not arg.startswith("_kw_"):
                key, val = arg.split("=", 1)
                val = val.strip().strip("'\"")
                encoded.append(f"_kw_{key.strip()}__{val.encode().hex()}")
            else:
                positional.append(arg)
        return f"{m.group(1)}({', '.join(positional + encoded)})"

    while True:
        nxt = _CALL_RE.sub(one_call, s)
        if nxt == s:
            return s
        s = nxt


def _build_sympy_namespace(field_names: Optional[Set[str]] = None) -> Dict[str, Any]:
    field_names = set(field_names or ())
    ns: Dict[str, Any] = {}
    for sym in _RESERVED_SYMBOLS - field_names:
        ns[sym] = sp.Symbol(sym)
    for name in field_names:
        ns[name] = sp.Symbol(name)
    for name in _USER_VARS:
        ns.setdefault(name, sp.Symbol(name))
    for alias, canonical in _ALIAS.items():
        ns.setdefault(alias, sp.Function(alias))
        ns.setdefault(canonical, sp.Function(canonical))
    # Registered names that SymPy would otherwise resolve to its own builtins (``hessian``,
    # ``cos``) -- the reference's ``_hessian`` special case (stenpy_engine.py:3562-3566) never
    # triggers, so there ``hessian(f)`` fails to parse; here every op name is callable.
    for name in OP_REGISTRY:
        if not name.startswith("_"):
            ns.setdefault(name, sp.Function(name))
    ns["pi"] = sp.pi
    ns["E"] = sp.E
    return ns


def parse_expression(expr_str: str, field_names: Optional[Set[str]] = None) -> sp.Expr:
    """String -> SymPy expression (stenpy_engine.py:3576-3582)."""
    cleaned = _encode_kwargs(expr_str.strip())
    ns = _build_sympy_namespace(field_names)
    for kw in re.findall(r"_kw_\w+", cleaned):
        ns[kw] = sp.Symbol(kw)
    try:
        return sp.sympify(cleaned, locals=ns)
    except Exception as exc:
        raise ValueError(f"Could not parse '{expr_str}': {exc}\nProcessed form: '{cleaned}'") from exc


@dataclass
class _CompileCtx:
    graph: Graph
    dx: Any
    boundary: str
    cse: Dict[Tuple, str] = dc_field(default_factory=dict)
    src_id: str = ""
    warnings: List[str] = dc_field(default_factory=list)
    field_map: Dict[str, str] = dc_field(default_factory=dict)

    def node(self, key: Tuple, op: str, inputs: Tuple[str, ...], params: Dict[str, Any]) -> str:
        """Hash-consed node insertion."""
        if key in self.cse:
            return self.cse[key]
        nid = self.graph.add(op, inputs, params)
        self.cse[key] = nid
        return nid


def _split_kwargs(func_name: str, args: Tuple) -> Tuple[str, Dict[str, Any], List]:
    real, kwargs = [], {}
    for a in args:
        s = str(a)
        if s.startswith("_kw_"):
            rest = s[4:]
            sep = rest.index("__")
            kwargs[rest[:sep]] = _decode_kw_value(rest[sep + 2:])
        else:
            real.append(a)
    name = func_name.lower()
    return _ALIAS.get(name, name), kwargs, real


def _op_params(canonical: str, dx: Any, boundary: str) -> Dict[str, Any]:
    return {} if canonical in _NO_STENCIL_OPS else {"dx": dx, "boundary": boundary}


def _frozen(params: Dict[str, Any]) -> Tuple:
    return tuple(sorted((k, v) for k, v in params.items()
                        if isinstance(v, (int, float, str, bool))))


def _scalar_const(value: float) -> torch.Tensor:
    return torch.tensor(float(value), dtype=torch.float64)


def _lower(expr: sp.Expr, ctx: _CompileCtx) -> str:
    if isinstance(expr, sp.Symbol):
        name = str(expr)
        if name in ctx.field_map:
            return ctx.field_map[name]
        if name not in _PRIMARY_ALIASES:
            ctx.warnings.append(f"Unknown symbol '{name}' — falling back to primary field.")
        return ctx.src_id

    if isinstance(expr, sp.Number):
        val = float(expr)
        return ctx.node(("_scalar", val), "_constant", (), {"value": _scalar_const(val)})

    if isinstance(expr, sp.core.function.AppliedUndef):
        canonical, kwargs, real_args = _split_kwargs(type(expr).__name__, expr.args)
        child_ids = tuple(_lower(a, ctx) for a in real_args)
        params = _op_params(canonical, ctx.dx, ctx.boundary)
        params.update(kwargs)
        if len(child_ids) == 1:
            child = ctx.graph.node(child_ids[0])
            rewrite = _SIMPLIFY_RULES.get((canonical, child.op_name))
            if rewrite is not None:
                ctx.warnings.append(f"Simplified {canonical}({child.op_name}(f)) → {rewrite}(f)")
                canonical, child_ids = rewrite, child.input_ids
        if (canonical, child_ids, _frozen(params)) not in ctx.cse and canonical not in OP_REGISTRY:
            raise ValueError(f"Unknown operator '{canonical}' (from '{type(expr).__name__}').\n"
                             f"Available: {sorted(OP_REGISTRY)}")
        return ctx.node((canonical, child_ids, _frozen(params)), canonical, child_ids, params)

    if isinstance(expr, (sp.Add, sp.Mul)):
        op = "add" if isinstance(expr, sp.Add) else "mul"
        operands = list(expr.args)
        if op == "mul" and len(operands) == 2 and sp.Integer(-1) in operands:
            other = operands[1] if operands[0] == sp.Integer(-1) else operands[0]
            cid = _lower(other, ctx)
            return ctx.node(("neg", (cid,), ()), "neg", (cid,), {})
        ids = [_lower(o, ctx) for o in operands]
        acc = ids[0]
        for nxt in ids[1:]:
            acc = ctx.node((op, (acc, nxt), ()), op, (acc, nxt), {})
        return acc

    if isinstance(expr, sp.Pow):
        base_id = _lower(expr.args[0], ctx)
        e = float(expr.args[1])
        if abs(e - 0.5) < 1e-9:
            return ctx.node(("sqrt", (base_id,), ()), "sqrt", (base_id,), {})
        n = int(e)
        if abs(e - n) < 1e-9 and abs(n) <= 16:
            if n == 0:
                return ctx.graph.add("_constant", (), {"value": _scalar_const(1.0)})
            if n < 0:
                pos = _lower(sp.Pow(expr.args[0], sp.Integer(-n)), ctx)
                one = ctx.graph.add("_constant", (), {"value": _scalar_const(1.0)})
                return ctx.graph.add("div", (one, pos), {})
            acc = base_id
            for _ in range(n - 1):
                acc = ctx.node(("mul", (acc, base_id), ()), "mul", (acc, base_id), {})
            return acc
        log_id = ctx.graph.add("log", (base_id,), {})
        c_id = _lower(sp.Float(e), ctx)
        mul_id = ctx.graph.add("mul", (log_id, c_id), {})
        return ctx.graph.add("exp", (mul_id,), {})

    raise ValueError(f"Unsupported SymPy node {type(expr).__name__}: {expr}\n"
                     "Use recognised operators and arithmetic (+, *, **).")


def compile_expression(expr_str: str, dx: Any = 1.0, boundary: str = "neumann",
                       field_map: Optional[Dict[str, Any]] = None,
                       eliminate_dead: bool = True
                       ) -> Tuple[Graph, str, List[str], Dict[str, str], sp.Expr]:
    """Compile ``expr_str`` into ``(graph, sink_id, warnings, field ids, sympy expr)``.

    Same contract as stenpy_engine.py:3762-3778 (defaults dx=1.0, boundary="neumann";
    a missing ``field_map`` means one placeholder field ``f``).
    """
    if field_map is None:
        field_map = {"f": torch.zeros(1, dtype=torch.float64)}
    sym = parse_expression(expr_str, field_names=set(field_map))
    g = Graph()
    ctx = _CompileCtx(graph=g, dx=dx, boundary=boundary)
    for name, value in field_map.items():
        ctx.field_map[name] = g.add("_constant", (), {"value": value})
    ctx.src_id = ctx.field_map.get("f", next(iter(ctx.field_map.values())))
    sink = _lower(sym, ctx)
    if eliminate_dead:
        keep = [sink] + list(ctx.field_map.values())
        g = g.dead_code_eliminate(keep)
    return g, sink, ctx.warnings, dict(ctx.field_map), sym


def compile_expressions(exprs: Sequence[str], dx: Any = 1.0, boundary: str = "neumann",
                        field_map: Optional[Dict[str, Any]] = None
                        ) -> Tuple[Graph, List[str], List[str], Dict[str, str]]:
    """Compile several expressions over the same fields into ONE graph with one sink each,
    sharing common subexpressions across them (``hessian(f)`` under both
    ``eigenvalues(hessian(f))`` and ``trace(hessian(f))`` is one node), so a fused run reads
    each field once for all of them.  Returns ``(graph, sink_ids, warnings, field ids)``.
    The reference compiles one expression per graph (stenpy_engine.py:3762-3778)."""
    if field_map is None:
        field_map = {"f": torch.zeros(1, dtype=torch.float64)}
    g = Graph()
    ctx = _CompileCtx(graph=g, dx=dx, boundary=boundary)
    for name, value in field_map.items():
        ctx.field_map[name] = g.add("_constant", (), {"value": value})
    ctx.src_id = ctx.field_map.get("f", next(iter(ctx.field_map.values())))
    sinks = [_lower(parse_expression(e, field_names=set(field_map)), ctx) for e in exprs]
    g = g.dead_code_eliminate(sinks + list(ctx.field_map.values()))
    return g, sinks, ctx.warnings, dict(ctx.field_map)

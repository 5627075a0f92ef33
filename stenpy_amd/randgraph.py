"""Random type-correct expression graphs (batched workloads and schedule-model training data).

Ops are drawn from {gradient, laplacian, curl, divergence, exp, tanh, clamp, mul, add, mean}
with scalar/vector typing: gradient S->V, curl V->V, divergence V->S, laplacian S->S,
exp/tanh/clamp keep the type, mul/add combine equal types, mean ends a chain (S or V ->
reduction).  Such graphs never hit the reference's vector-Laplacian quirk, and the result
is always a compilable expression string for ``compile_expression``.
"""
from __future__ import annotations

import random
from typing import List, Tuple

UNARY_SAME = ["exp", "tanh", "clamp"]


def random_expression(rng: random.Random, n_ops: int, want_mean: bool = None,
                      vector_leaf: bool = True, mid_mean: float = 0.0) -> Tuple[str, int]:
    """Expression over field ``f`` (scalar) and, with ``vector_leaf``, ``u`` (vector).

    ``mid_mean`` is the probability that a scalar step adds the mean of another scalar
    expression (a reduction feeding later pointwise work).  Returns (expr, op count)."""
    pool: List[Tuple[str, str]] = [("f", "S"), ("u", "V")] if vector_leaf else [("f", "S")]
    ops = 0
    while ops < n_ops:
        e, t = rng.choice(pool)
        r = rng.random()
        if t == "S" and mid_mean and rng.random() < mid_mean:
            e2 = rng.choice([p for p in pool if p[1] == "S"])[0]
            cand = (f"({e} - mean({e2}))", "S")
            ops += 1
        elif t == "S":
            if r < 0.25:
                cand = (f"gradient({e})", "V")
            elif r < 0.45:
                cand = (f"laplacian({e})", "S")
            elif r < 0.75:
                fn = rng.choice(UNARY_SAME)
                if fn == "exp":
                    cand = (f"exp(tanh({e}))", "S")
                    ops += 1
                elif fn == "clamp":
                    # keyword arguments are only encoded on calls without nested parentheses
                    # (the reference front-end's rule), so bounds are given to bare symbols only
                    cand = (f"clamp({e}, lo=-1, hi=1)", "S") if e.isidentifier() else (f"clamp({e})", "S")
                else:
                    cand = (f"tanh({e})", "S")
            else:
                e2 = rng.choice([p for p in pool if p[1] == "S"])[0]
                cand = (f"({e} {rng.choice(['*', '+'])} {e2})", "S")
        else:
            if r < 0.3:
                cand = (f"curl({e})", "V")
            elif r < 0.6:
                cand = (f"divergence({e})", "S")
            elif r < 0.8:
                cand = (f"tanh({e})", "V")
            else:
                e2 = rng.choice([p for p in pool if p[1] == "V"])[0]
                cand = (f"({e} {rng.choice(['*', '+'])} {e2})", "V")
        pool.append(cand)
        ops += 1
    # result: the newest expression, reduced with mean sometimes
    e, t = pool[-1]
    if want_mean is None:
        want_mean = rng.random() < 0.2
    if want_mean:
        return f"mean({e})", ops + 1
    if t == "V" and rng.random() < 0.5:
        return f"divergence({e})", ops + 1
    return e, ops


def random_shape(rng: random.Random, lo: int = 32, hi: int = 384) -> Tuple[int, int, int]:
    return tuple(rng.randint(lo, hi) for _ in range(3))

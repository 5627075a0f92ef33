"""The five benchmark workloads as graphs on seeded synthetic fields.

Shared by the benchmark and by the schedule study (``python -m stenpy_amd.tune configs``), so
that both measure the same fused groups:

  1. plain eager, 128^3: five single expressions on the Fourier field of seed 1;
  2. the stencil graph, 512^3: lap(f) + u.grad(f) + |curl u|^2 with sum/mean/variance/L2;
  3. batched: 96 fields with axes U{32..384}, 6 random graphs of 3-12 ops;
  4. out of core: f**2 + tanh(f)*exp(f) + laplacian(f) on 256-plane slabs of a
     4608 x 4096 x 4096 field (plus the streamed L2 / sum / min-max);
  5. tensor ops, 768^3: hessian -> eigenvalues, determinant, trace, deviatoric + mean curvature
     on a field whose upper half is i^2 + j^2 + k^2.
"""
from __future__ import annotations

import random
from typing import List, Sequence, Tuple

import torch

PLAIN_EXPRS = ["f**2 + tanh(f)*exp(f)", "laplacian(f)", "gradient(f)", "mean(f)", "norm_l2(f)"]
STENCIL_GRAPH = "divergence(gradient(f)) + dot(u, gradient(f)) + sqnorm(curl(u))"
REDUCTIONS = ("sum", "mean", "variance", "norm_l2")
STREAM_EXPR = "f**2 + tanh(f)*exp(f) + laplacian(f)"
STREAM_SHAPE = (4608, 4096, 4096)
STREAM_SLAB = 256
TENSOR_EXPRS = ["eigenvalues(hessian(f))", "determinant(hessian(f))", "trace(hessian(f))",
                "deviatoric(hessian(f))", "mean_curvature(f)"]


def plain_graphs(dev, n: int = 128):
    """Config 1: one graph per expression on the n^3 Fourier field of seed 1 (periodic, dx 1)."""
    from . import gen
    from .compiler import compile_expression
    f = gen.generate((n, n, n), 1, dev)
    out = []
    for e in PLAIN_EXPRS:
        g, sink, *_ = compile_expression(e, dx=1.0, boundary="periodic", field_map={"f": f})
        out.append((e, g, [sink]))
    return out


def stencil_graph(f, u):
    """Config 2's graph on given f (n0, n1, n2) and u (n0, n1, n2, 3): (graph, sink, reductions)."""
    from .compiler import compile_expression
    g, sink, _, _, _ = compile_expression(STENCIL_GRAPH, dx=1.0, boundary="periodic",
                                          field_map={"f": f, "u": u})
    reds = [g.add(r, (sink,), {}) for r in REDUCTIONS]
    return g, sink, reds


def stencil_inputs(dev, n: int = 512):
    """Config 2's inputs: f of seed 2, u with components of seeds 3, 4, 5."""
    from . import gen
    return gen.generate((n, n, n), 2, dev), gen.generate_vector((n, n, n), (3, 4, 5), dev)


def batched_items(dev) -> Tuple[List[str], List[Tuple[object, List[str]]]]:
    """Config 3: 96 fields with axes U{32..384} (seed 6) and 6 random graphs of 3-12 ops
    (seed 6); field i (seed 1000 + i) uses graph i mod 6.  Returns (exprs, [(graph, wanted)])."""
    from . import gen, randgraph
    from .compiler import compile_expression
    rng = random.Random(6)
    exprs = [randgraph.random_expression(rng, rng.randint(3, 12), vector_leaf=False, mid_mean=0.15)[0]
             for _ in range(6)]
    shapes = [tuple(rng.randint(32, 384) for _ in range(3)) for _ in range(96)]
    items = []
    for i, sh in enumerate(shapes):
        f = gen.generate(sh, 1000 + i, dev)
        g, sink, *_ = compile_expression(exprs[i % 6], boundary="periodic", field_map={"f": f})
        items.append((g, [sink]))
    return exprs, items


def stream_slab_graph(dev, planes: int = STREAM_SLAB + 2):
    """Config 4's per-slab group: the expression on a padded slab of ``planes`` planes of the
    4096 x 4096 cross-section (256 planes plus the 1-plane halo on each side), with the
    streamed reductions of the sink.  The slab holds the first planes of the seed-7 field."""
    from . import gen
    from .compiler import compile_expression
    from .stream import STREAM_REDUCTIONS
    f = gen.generate((planes,) + STREAM_SHAPE[1:], 7, dev)
    g, sink, *_ = compile_expression(STREAM_EXPR, boundary="periodic", field_map={"f": f})
    reds = [g.add(k, (sink,), {}) for k in STREAM_REDUCTIONS]
    return g, [sink] + reds


def tensor_graph(dev, n: int = 768):
    """Config 5: n^3 FP32 field, lower half the seeded Fourier field (seed 8), upper half
    i^2 + j^2 + k^2; one graph whose five sinks share the hessian.  Returns (g, sinks, f)."""
    from . import gen
    from .compiler import compile_expressions
    f = gen.generate((n, n, n), 8, dev)
    i = torch.arange(n, device=dev, dtype=torch.float32)
    q = (i.view(-1, 1, 1) ** 2 + i.view(1, -1, 1) ** 2 + i.view(1, 1, -1) ** 2)
    f[n // 2:] = q[n // 2:]
    del q
    g, sinks, *_ = compile_expressions(TENSOR_EXPRS, boundary="periodic", field_map={"f": f})
    return g, sinks, f

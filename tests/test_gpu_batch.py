"""Batched persistent launches vs the FP64 oracle and the per-graph fused path (GPU)."""
import random

import pytest
import torch

import stenpy_amd as sp
from stenpy_amd import batch as B, gen, randgraph
from oracle import graph_eval
from tests._util import assert_2x

pytestmark = pytest.mark.gpu


def _items(n_items, n_graphs, seed, lo, hi, device="cuda"):
    rng = random.Random(seed)
    exprs = [randgraph.random_expression(rng, rng.randint(3, 12), vector_leaf=False, mid_mean=0.15)[0]
             for _ in range(n_graphs)]
    items = []
    for i in range(n_items):
        shape = tuple(rng.randint(lo, hi) for _ in range(3))
        f = gen.generate(shape, 1000 + i, device)
        g, sink, *_ = sp.compile_expression(exprs[i % n_graphs], boundary="periodic", field_map={"f": f})
        items.append((g, [sink]))
    return exprs, items


def _check_against_oracle(br, items):
    for ii, (g, wanted) in enumerate(items):
        out = br.outputs(ii)
        o64 = graph_eval.evaluate(g, torch.float64)
        o32 = graph_eval.evaluate(g, torch.float32)
        for w in wanted:
            assert_2x(out[w], o64[w], o32[w], f"item {ii} {g.node(w).op_name}")


def test_batch_matches_oracle_and_fused_path():
    exprs, items = _items(12, 6, seed=11, lo=32, hi=80)
    print(exprs, flush=True)
    br = B.BatchRunner(items, device="cuda")
    br.run()
    torch.cuda.synchronize()
    _check_against_oracle(br, items)
    rt = sp.Runtime(device="cuda", fusion_enabled=True)
    for ii, (g, wanted) in enumerate(items):
        ref = rt.fused.run(g, rt, wanted=wanted)
        out = br.outputs(ii)
        for w in wanted:
            a, b = out[w].double(), ref[w].double()
            assert (a - b).abs().max().item() <= 1e-5 * max(1.0, b.abs().max().item()), (ii, exprs[ii % 6])


def test_batch_graph_replays_match_eager():
    exprs, items = _items(10, 5, seed=12, lo=32, hi=64)
    br = B.BatchRunner(items, device="cuda")
    br.run()
    torch.cuda.synchronize()
    eager = [{w: v.clone() for w, v in br.outputs(i).items()} for i in range(len(items))]
    graph = br.capture()
    for rep in range(3):
        for i in range(len(items)):                      # poison the outputs between replays
            for v in br.outputs(i).values():
                v.fill_(float("nan"))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for i in range(len(items)):
            for w, v in br.outputs(i).items():
                assert torch.equal(v, eager[i][w]), (rep, i, w)
        # unrelated eager work between replays
        torch.randn(1 << 20, device="cuda").sum().item()


def test_batch_with_vector_outputs_and_reductions():
    shapes = [(33, 40, 47), (64, 35, 52), (41, 41, 70)]
    items = []
    for i, sh in enumerate(shapes):
        f = gen.generate(sh, 50 + i, "cuda")
        for expr in ["gradient(f) + curl(gradient(tanh(f)))", "f - mean(laplacian(f))", "divergence(gradient(f*f))"]:
            g, sink, *_ = sp.compile_expression(expr, boundary="periodic", field_map={"f": f})
            r = g.add("variance", (sink,), {}) if expr != "gradient(f) + curl(gradient(tanh(f)))" else None
            items.append((g, [sink] + ([r] if r else [])))
    br = B.BatchRunner(items, device="cuda")
    br.run()
    torch.cuda.synchronize()
    _check_against_oracle(br, items)

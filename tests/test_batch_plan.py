"""Batched plans on the CPU: levels, work lists, parameter packing, and the batch module
compiles for gfx950 (no GPU needed)."""
import random
import struct

import numpy as np
import pytest
import torch

import stenpy_amd as sp
from stenpy_amd import batch as B, fused as F, randgraph


def _graphs(n_items=8, n_graphs=4, seed=3, lo=32, hi=72):
    rng = random.Random(seed)
    exprs = [randgraph.random_expression(rng, rng.randint(3, 12), vector_leaf=False, mid_mean=0.15)[0]
             for _ in range(n_graphs)]
    out = []
    for i in range(n_items):
        shape = tuple(rng.randint(lo, hi) for _ in range(3))
        g, sink, *_ = sp.compile_expression(exprs[i % n_graphs], boundary="periodic",
                                            field_map={"f": torch.randn(shape)})
        out.append((g, [sink]))
    return exprs, out


def test_queue_split_balances_cost_in_contiguous_runs():
    rng = np.random.default_rng(0)
    for n in (0, 1, 7, 100, 2049):
        c = rng.uniform(1, 10, n)
        b = B.queue_split(c, 8)
        assert b[0] == 0 and b[-1] == n and len(b) == 9 and np.all(np.diff(b) >= 0)
        if n >= 100:
            sums = [c[b[q]:b[q + 1]].sum() for q in range(8)]
            assert max(sums) - min(sums) <= 2 * c.max()


def test_batch_plan_units_cover_every_tile_and_tail_block():
    exprs, graphs = _graphs()
    br = B.BatchRunner(graphs, device="cpu")
    st = br.stats()
    assert st["levels"] >= 1 and sum(st["units"]) > 0
    for lvl, entries in enumerate(br.levels):
        ln = br.launches[lvl]
        if not entries:
            continue
        u = ln["units"].numpy()
        assert u.shape == (ln["nunits"], 4)
        qs = ln["qstart"].numpy()
        assert qs[0] == 0 and qs[-1] == ln["nunits"] and np.all(np.diff(qs) >= 0)
        assert ln["counters"].numpy().tolist() == [0] * 9
        keys = br.mods[lvl]["keys"]
        for ki, k in enumerate(keys):
            es = [e for e in entries if e.key == k]
            for ii, e in enumerate(es):
                bs = np.sort(u[(u[:, 0] == ki) & (u[:, 1] == ii), 2])
                assert bs.tolist() == list(range(e.nb + e.nbt))
                # tail boxes plus interior tiles cover the grid exactly once
                gen = br.mods[lvl]["gens"][ki]
                n0, n1, n2 = e.group.grid
                hits = np.zeros((n0, n1, n2), np.int8)
                for b in gen.boundary_boxes(e.group.grid):
                    hits[b[0]:b[1], b[2]:b[3], b[4]:b[5]] += 1
                for (z0, z1, y0, y1, x0, x1) in gen.main_boxes(e.group.grid):
                    hits[z0:z1, y0:y1, x0:x1] += 1
                assert hits.min() == 1 and hits.max() == 1
        # one parameter record per grid of each structure
        for ki, prm in enumerate(ln["params"]):
            n_items = sum(1 for e in entries if e.key == keys[ki])
            assert prm.numel() % n_items == 0 and (prm.numel() // n_items) % 8 == 0


def test_batch_module_compiles():
    exprs, graphs = _graphs(n_items=6, n_graphs=3, seed=5)
    br = B.BatchRunner(graphs, device="cpu")
    for m in br.mods:
        if m["source"]:
            code, _ = F.compile_source(m["source"], "batch_kernel")
            assert len(code) > 1000

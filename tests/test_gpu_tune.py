"""Autotuner and learned schedule model on the GPU.

* exhaustive search picks the fastest candidate and records it in the database;
* the schedule MLP, trained by the project's HIP kernels on the committed timing rows of
  random graphs, is evaluated on held-out graphs: regret against exhaustive search and the gain
  of its pick over the shape heuristic used without it.
"""
import json
import math

import numpy as np
import pytest
import torch

import stenpy_amd as sp
from stenpy_amd import fused as F, gen, tune

pytestmark = pytest.mark.gpu


def test_autotune_records_the_fastest_candidate(tmp_path, monkeypatch):
    monkeypatch.setenv("STENPY_TUNE_DB", str(tmp_path / "db.json"))
    shape = (40, 48, 96)
    f = gen.generate(shape, 1, "cuda")
    g, sink, *_ = sp.compile_expression("laplacian(f) + tanh(f)", boundary="periodic", field_map={"f": f})
    rt = sp.Runtime(device="cuda", fusion_enabled=True)
    cands = [F.Schedule(32, 8, 32, 128), F.Schedule(64, 8, 64, 256), F.Schedule(32, 4, 32, 128)]
    rows = tune.autotune(g, rt, [sink], cands=cands, reps=5, workers=1)
    assert {tuple(r["schedule"]) for r in rows} >= {c.as_tuple() for c in cands}
    assert all(r["ms"] > 0 and len(r["features"]) == tune.N_FEATURES for r in rows)
    db = json.loads((tmp_path / "db.json").read_text())
    best = min(rows, key=lambda r: r["ms"])
    assert db[best["sig"]] == best["schedule"]
    # the executor now takes the recorded schedule, and still matches the eager path
    rt2 = sp.Runtime(device="cuda", fusion_enabled=True, schedule_fn=tune.schedule_fn())
    out = rt2.fused.run(g, rt2, wanted=[sink])[sink]
    steps, info, comp = next(iter(rt2.fused._plans.values()))
    assert list(comp[0].gen.sched.as_tuple()) == best["schedule"]
    ref = rt2.run_eager(g)[sink]
    assert (out - ref).abs().max().item() <= 1e-4 * ref.abs().max().item()


@pytest.mark.skipif(not tune.ROWS_PATH.exists(), reason="no committed timing rows")
def test_learned_schedules_beat_the_heuristic_on_heldout_graphs():
    rows = tune._read_rows(tune.ROWS_PATH)
    model, rep = tune.train_and_report(rows, device="cuda", epochs=40)
    print(json.dumps(rep))
    held, top3 = rep["heldout"], rep["heldout_top3"]
    assert held["groups"] >= 5
    assert rep["loss_last"] < 0.5 * rep["loss_first"]
    # the network's pick is on average faster than the heuristic's and close to exhaustive;
    # timing its three best proposals comes closer still.  The gain is measured against the
    # most any pick can gain over the heuristic (the exhaustive search's own gain), in log terms
    share = lambda r: math.log(r["geomean_gain_vs_default"]) / math.log(r["geomean_ceiling_gain_vs_default"])
    assert held["geomean_gain_vs_default"] > 1.0 and share(held) > 0.4, held
    assert share(top3) > 0.7, top3
    assert held["mean_regret"] < 0.1, held
    assert top3["mean_regret"] < 0.05 and top3["mean_regret"] <= held["mean_regret"], top3


@pytest.mark.skipif(not tune.MODEL_PATH.exists(), reason="no committed model")
def test_committed_model_loads_and_proposes_valid_schedules():
    model = tune.ScheduleModel.load(tune.MODEL_PATH, "cuda")
    shape = (64, 80, 200)
    f = gen.generate(shape, 3, "cuda")
    u = gen.generate_vector(shape, (4, 5, 6), "cuda")
    g, sink, *_ = sp.compile_expression("divergence(gradient(f)) + dot(u, gradient(f)) + sqnorm(curl(u))",
                                        boundary="periodic", field_map={"f": f, "u": u})
    values = {n: g.node(n).params["value"] for n in g._order if g.node(n).op_name == "_constant"}
    steps, info = F.plan_graph(g, values, [sink])
    s = model.best(steps[0].group, g)
    assert s is not None and s in model.cands
    gen_ = F.KernelGen(g, steps[0].group, info, s, {}, {})
    assert gen_.lds_bytes() <= F.LDS_LIMIT


@pytest.mark.skipif(not tune.MODEL_PATH.exists(), reason="no committed model")
def test_shortlist_search_times_only_the_proposals(tmp_path, monkeypatch):
    monkeypatch.setenv("STENPY_TUNE_DB", str(tmp_path / "db.json"))
    model = tune.ScheduleModel.load(tune.MODEL_PATH, "cuda")
    shape = (48, 64, 160)
    f = gen.generate(shape, 9, "cuda")
    g, sink, *_ = sp.compile_expression("laplacian(tanh(f)) * f", boundary="neumann", field_map={"f": f})
    rt = sp.Runtime(device="cuda", fusion_enabled=True)
    rows = tune.autotune(g, rt, [sink], cands=lambda grp, graph: model.shortlist(grp, graph, 3),
                         reps=5, workers=1)
    scheds = {tuple(r["schedule"]) for r in rows}
    assert 1 <= len(scheds) <= 4            # three proposals (+ the heuristic when not among them)
    db = json.loads((tmp_path / "db.json").read_text())
    assert db[rows[0]["sig"]] == min(rows, key=lambda r: r["ms"])["schedule"]

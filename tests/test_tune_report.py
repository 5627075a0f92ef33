"""Schedule-study scoring on the host: the model's picks are ranked inside its own candidate
set, and regrets are taken against the best of every timed candidate."""
import numpy as np

from stenpy_amd import tune


class _RankModel:
    """Predicts the score stored in each row (lower = faster)."""

    def predict_rows(self, rows):
        return np.asarray([r["score"] for r in rows], dtype=np.float32)


def _restrict(monkeypatch):
    """A model candidate set without the 512-plane z-chunks, so that some timed rows lie
    outside it."""
    monkeypatch.setattr(tune, "MODEL_CANDIDATES", [c for c in tune.CANDIDATES if c.zc <= 256])


def _rows():
    own = tune.MODEL_CANDIDATES[:5]
    wide = [c for c in tune.CANDIDATES if c.zc == 512][:2]     # outside the model's set
    ms = [3.0, 2.0, 2.5, 4.0, 1.9, 1.5, 5.0]
    score = [0.1, 0.3, 0.0, 0.5, 0.2, -9.0, -9.0]            # the model never sees the last two
    return [{"schedule": list(c.as_tuple()), "ms": m, "grid": [64, 64, 64], "score": s}
            for c, m, s in zip(own + wide, ms, score)]


def test_config_choice_scores_top1_and_top3_against_the_exhaustive_best(monkeypatch):
    _restrict(monkeypatch)
    rows = _rows()
    r = tune.config_choice(rows, _RankModel(), db_sched=rows[5]["schedule"])
    assert r["timed"] == 7 and r["model_candidates"] == 5
    assert r["best"]["ms"] == 1.5 and r["best_in_model_set"]["ms"] == 1.9
    # top-1 = lowest score inside the model's set (score 0.0 -> 2.5 ms)
    assert r["model_top1"]["ms"] == 2.5
    assert np.isclose(r["model_top1"]["regret"], 2.5 / 1.5 - 1)
    assert np.isclose(r["model_top1"]["regret_in_model_set"], 2.5 / 1.9 - 1)
    # top-3 = scores 0.0, 0.1, 0.2 -> the fastest of 2.5, 3.0, 1.9
    assert r["model_top3"]["ms"] == 1.9 and np.isclose(r["model_top3"]["regret_in_model_set"], 0.0)
    assert r["database"]["regret"] == 0.0


def test_config_choice_never_picks_outside_the_model_candidates(monkeypatch):
    _restrict(monkeypatch)
    rows = _rows()
    r = tune.config_choice(rows, _RankModel())
    own = {tuple(c.as_tuple()) for c in tune.MODEL_CANDIDATES}
    assert tuple(r["model_top1"]["schedule"]) in own and tuple(r["model_top3"]["schedule"]) in own
    assert "database" not in r


def test_model_shortlist_respects_a_block_size_filter(monkeypatch):
    """The batched launches share one block size; the model's pick is restricted to it."""
    import torch
    import stenpy_amd as sp
    from stenpy_amd import fused as F

    f = torch.zeros((40, 36, 48))
    g, sink, *_ = sp.compile_expression("laplacian(f) + exp(f)", boundary="periodic", field_map={"f": f})
    values = {n: g.node(n).params["value"] for n in g._order if g.node(n).op_name == "_constant"}
    steps, info = F.plan_graph(g, values, [sink])
    grp = [s.group for s in steps if s.kind == "group"][0]

    class _Mlp:
        def predict(self, X):
            return -np.asarray(X)[:, 18]          # feature 18 = threads / 256: prefers big blocks

    m = tune.ScheduleModel(_Mlp())
    assert m.best(grp, g).threads == 512
    assert m.best(grp, g, threads=256).threads == 256

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


def test_fill_small_grid_halves_the_chunk_until_the_gpu_is_filled():
    from stenpy_amd.fused import Schedule, fill_small_grid
    # 128^3 with a 64 x 8 tile: 32 tiles per plane, 4 chunks of 32 planes -> 128 blocks
    s = fill_small_grid(Schedule(64, 8, 32, 256), (128, 128, 128))
    assert s == Schedule(64, 8, 8, 256)          # 16 chunks of 8 planes: 512 blocks
    # a large grid keeps its schedule (512 blocks of 64 x 16 x 256 at 512^3)
    big = Schedule(64, 16, 256, 512)
    assert fill_small_grid(big, (512, 512, 512)) is big
    # halving stops as soon as two blocks per CU are reached
    assert fill_small_grid(Schedule(32, 8, 64, 128), (128, 128, 128)) == Schedule(32, 8, 16, 128)
    # and at the 8-plane floor, whatever the block count
    assert fill_small_grid(Schedule(64, 16, 64, 512), (32, 32, 64)) == Schedule(64, 16, 8, 512)

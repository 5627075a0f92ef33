"""Schedule autotuning for fused groups: exhaustive search, timing database, learned model.

A schedule is (TX, TY, ZC, threads).  ``autotune`` compiles every candidate of a group in
parallel (hipRTC releases the GIL), times each kernel alone with HIP events (min over rounds
of the mean of ``reps`` launches), stores the winner in a JSON database keyed by the group's
structure, and returns the whole timing table -- the same rows train the schedule MLP
(``stenpy_amd.mlp``).  ``schedule_fn`` answers the fused executor from the database, then
from a trained model if one is installed, else leaves the default schedule.
"""
from __future__ import annotations

import concurrent.futures as cf
import json
import math
import os
import threading
from pathlib import Path
from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import fused as F
from .fused import Schedule

_DB_LOCK = threading.Lock()

# threads must cover whole tile rows: NT % TX == 0 and TY a multiple of NT / TX
CANDIDATES: List[Schedule] = [
    Schedule(tx, ty, zc, nt)
    for tx in (32, 64, 128) for ty in (4, 8, 16) for zc in (16, 32, 64, 128) for nt in (256, 512)
    if nt % tx == 0 and ty % (nt // tx) == 0
]
# a smaller grid of candidates for the training-data sweep
DATA_CANDIDATES: List[Schedule] = [c for c in CANDIDATES if c.zc in (32, 64, 128)]


# ------------------------------------------------------------------------------ database
def db_path() -> Path:
    d = os.environ.get("STENPY_TUNE_DB")
    return Path(d) if d else Path(__file__).resolve().parent / "tuning_db.json"


def load_db() -> Dict[str, list]:
    p = db_path()
    if p.exists():
        try:
            return json.loads(p.read_text())
        except (OSError, ValueError):
            return {}
    return {}


def save_db(db: Dict[str, list]) -> None:
    p = db_path()
    with _DB_LOCK:
        tmp = p.with_suffix(".tmp")
        tmp.write_text(json.dumps(db, indent=1, sort_keys=True))
        os.replace(tmp, p)


def group_signature(group, graph) -> str:
    """Structure of a fused group (ops, edges, params, grid) -- the tuning DB key."""
    idx = {}
    parts = [f"grid={tuple(group.grid)}"]
    for i, s in enumerate(group.sources):
        idx[s] = f"s{i}"
    for i, d in enumerate(group.dscalars):
        idx[d] = f"d{i}"
    for i, nid in enumerate(group.nodes):
        node = graph.node(nid)
        idx[nid] = f"n{i}"
        ins = ",".join(idx.get(x, "c") for x in node.input_ids)
        ps = ",".join(f"{k}={v!r}" for k, v in sorted(node.params.items())
                      if isinstance(v, (int, float, str, bool, tuple, list)))
        parts.append(f"{node.op_name}({ins})[{ps}]")
    parts.append("out=" + ",".join(idx[o] for o in group.outputs))
    return "|".join(parts)


_MODEL = None


def install_model(model) -> None:
    """Use ``model`` (a trained ``ScheduleModel``) for groups missing from the database."""
    global _MODEL
    _MODEL = model


def schedule_fn(device: Optional[torch.device] = None, use_model: bool = True):
    db = load_db()

    def pick(group, graph) -> Optional[Schedule]:
        hit = db.get(group_signature(group, graph))
        if hit:
            return Schedule(*hit)
        if use_model and _MODEL is not None:
            return _MODEL.best(group, graph)
        return None

    return pick


# ------------------------------------------------------------------------------ features
def group_stats(gen: F.KernelGen) -> Dict[str, float]:
    """Static statistics of a group's generated interior code (schedule-independent)."""
    e = F._Emit(gen, ("lds", False, False)) if gen.use_lds else F._Emit(gen, ("gather",))
    pos = e.origin()
    for nid in gen.group.outputs:
        e.val(nid, pos)
    for nid in gen.group.reductions:
        e.val(gen.graph.node(nid).input_ids[0], pos)
    text = "\n".join(e.lines)
    g = gen.group
    src_ch = sum(l["C"] for l in gen.src_layout)
    out_ch = sum(max(gen.info[o].ch, 1) for o in g.outputs)
    return {
        "taps": text.count("lds_s") + text.count("p.src"),
        "lines": len(e.lines),
        "transc": sum(text.count(f) for f in ("exp(", "log(", "tanh(", "sin(", "cos(", "acos(")),
        "sqrt": text.count("sqrt"),
        "src_ch": src_ch, "out_ch": out_ch, "nred": len(g.reductions),
        "hz": gen.hz, "redge": gen.redge,
        "halo": max([l["H"] for l in gen.src_layout] + [0]),
        "bf16_out": sum(1 for o in g.outputs if gen.out_dtypes.get(o) == torch.bfloat16),
    }


def features(stats: Dict[str, float], grid: Sequence[int], s: Schedule, lds_bytes: int) -> List[float]:
    n0, n1, n2 = grid
    tiles = math.ceil(n2 / s.tx) * math.ceil(n1 / s.ty) * math.ceil(n0 / s.zc)
    h = stats["halo"]
    halo_ratio = (s.tx + 2 * h) * (s.ty + 2 * h) / (s.tx * s.ty)
    blocks_per_cu = min(160 * 1024 // max(lds_bytes, 1), 2048 // s.threads, 16)
    return [
        math.log2(n0), math.log2(n1), math.log2(n2), math.log2(n0 * n1 * n2),
        stats["taps"], stats["lines"], stats["transc"], stats["sqrt"], stats["src_ch"],
        stats["out_ch"], stats["nred"], stats["hz"], stats["redge"], h, stats["bf16_out"],
        math.log2(s.tx), math.log2(s.ty), math.log2(s.zc), s.threads / 256.0,
        lds_bytes / 1024.0, blocks_per_cu, math.log2(tiles), s.tx * s.ty / s.threads,
        halo_ratio, (2 * stats["hz"] + 1) / s.zc, tiles / 256.0,
        (n2 % s.tx != 0) + (n1 % s.ty != 0), min(n0, s.zc) / s.zc,
    ]


N_FEATURES = 28


# ------------------------------------------------------------------------------ search
def _compile_all(sources: List[str], workers: int) -> List[bytes]:
    with cf.ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        return [c for c, _ in ex.map(F.compile_source, sources)]


def candidates_for(graph, group, info, out_dtypes, cands: Sequence[Schedule] = None):
    gens = []
    for s in (cands or CANDIDATES):
        try:
            gen = F.KernelGen(graph, group, info, s, out_dtypes, {})
        except ValueError:
            continue
        if gen.lds_bytes() <= F.LDS_LIMIT:
            gens.append(gen)
    return gens


def autotune(graph, runtime, wanted: Sequence[str], seed: Optional[Dict[str, Any]] = None,
             cands: Sequence[Schedule] = None, reps: int = 10, workers: int = 16,
             update_db: bool = True, out_dtype: Optional[torch.dtype] = None) -> List[Dict[str, Any]]:
    """Exhaustive schedule search for every fused group of ``graph``; returns timing rows."""
    ex = runtime.fused
    results = ex.run(graph, runtime, seed, wanted=wanted, out_dtype=out_dtype)   # inputs per group
    values = {n: results[n] for n in graph._order if graph.node(n).op_name == "_constant"}
    steps, info, comp = ex.prepare(graph, values, wanted, out_dtype)
    rows: List[Dict[str, Any]] = []
    db = load_db() if update_db else {}
    for si, st in enumerate(steps):
        if st.kind != "group":
            continue
        base = comp[si]
        gens = candidates_for(graph, st.group, info, base.gen.out_dtypes, cands)
        srcs = [g.source() for g in gens]
        codes = _compile_all(srcs, workers)
        stats = group_stats(base.gen)
        sig = group_signature(st.group, graph)
        best = None
        for gen, src, code in zip(gens, srcs, codes):
            cg = F.CompiledGroup(gen, src)
            cg.module = F.load_module(src)
            ln = ex.prepare_launch(cg, graph, results)
            ms = ln.time_ms(reps=reps)
            nvox = int(np.prod(st.group.grid))
            rows.append({"sig": sig, "grid": list(st.group.grid), "schedule": list(gen.sched.as_tuple()),
                         "ms": ms, "ns_per_voxel": ms * 1e6 / nvox, "lds": gen.lds_bytes(),
                         "stats": stats, "features": features(stats, st.group.grid, gen.sched, gen.lds_bytes())})
            if best is None or ms < best[0]:
                best = (ms, gen.sched)
        if update_db and best is not None:
            db[sig] = list(best[1].as_tuple())
    if update_db:
        save_db(db)
    return rows


def default_time(rows: List[Dict[str, Any]], sched: Schedule = F.DEFAULT_SCHEDULE) -> Optional[float]:
    for r in rows:
        if tuple(r["schedule"]) == sched.as_tuple():
            return r["ms"]
    return None


# ------------------------------------------------------------------------------ learned model
def random_group_rows(n_graphs: int, seed: int, runtime, cands: Sequence[Schedule] = None,
                      shape_range: Tuple[int, int] = (64, 256), reps: int = 5, workers: int = 16,
                      log=None) -> List[Dict[str, Any]]:
    """Exhaustive timings of random config-3-style graphs on random grids (training data)."""
    import random
    from . import randgraph
    from .compiler import compile_expression
    from . import gen
    rng = random.Random(seed)
    rows: List[Dict[str, Any]] = []
    for gi in range(n_graphs):
        expr, nops = randgraph.random_expression(rng, rng.randint(3, 12))
        shape = randgraph.random_shape(rng, *shape_range)
        f = gen.generate(shape, 1000 + gi, runtime.device)
        u = gen.generate_vector(shape, (2000 + gi, 3000 + gi, 4000 + gi), runtime.device)
        g, sink, *_ = compile_expression(expr, boundary=rng.choice(["periodic", "neumann"]),
                                         field_map={"f": f, "u": u})
        try:
            r = autotune(g, runtime, [sink], cands=cands, reps=reps, workers=workers, update_db=False)
        except Exception as exc:  # a graph the planner cannot fuse is skipped
            if log:
                log(f"graph {gi} skipped: {exc}")
            continue
        for row in r:
            row["graph"] = gi
            row["expr"] = expr
        rows.extend(r)
        if log:
            log(f"graph {gi}: {expr[:60]} {shape} -> {len(r)} timings")
    return rows


class ScheduleModel:
    """MLP over (group, schedule) features predicting log(ns per voxel); picks the argmin."""

    def __init__(self, mlp=None, cands: Sequence[Schedule] = None):
        self.mlp = mlp
        self.cands = list(cands or CANDIDATES)

    def fit(self, rows: List[Dict[str, Any]], device="cuda", epochs: int = 150, seed: int = 0):
        from .mlp import ScheduleMLP
        X = np.asarray([r["features"] for r in rows], dtype=np.float32)
        y = np.log(np.asarray([r["ns_per_voxel"] for r in rows], dtype=np.float32))
        self.mlp = ScheduleMLP(X.shape[1], 256, 4, device, seed)
        return self.mlp.fit(X, y, epochs=epochs, batch=256, lr=2e-3, seed=seed)

    def predict_rows(self, rows: List[Dict[str, Any]]) -> np.ndarray:
        X = np.asarray([r["features"] for r in rows], dtype=np.float32)
        return self.mlp.predict(X)

    def best(self, group, graph) -> Optional[Schedule]:
        info = getattr(group, "_info", None)
        try:
            base = F.KernelGen(graph, group, info, F.DEFAULT_SCHEDULE, {}, {})
        except Exception:
            return None
        stats = group_stats(base)
        feats, scheds = [], []
        for s in self.cands:
            try:
                gen = F.KernelGen(graph, group, info, s, {}, {})
            except ValueError:
                continue
            if gen.lds_bytes() > F.LDS_LIMIT:
                continue
            feats.append(features(stats, group.grid, s, gen.lds_bytes()))
            scheds.append(s)
        if not scheds:
            return None
        pred = self.mlp.predict(np.asarray(feats, dtype=np.float32))
        return scheds[int(np.argmin(pred))]

    def save(self, path) -> None:
        torch.save({"mlp": self.mlp.state_dict(), "cands": [c.as_tuple() for c in self.cands]}, path)

    @classmethod
    def load(cls, path, device="cuda") -> "ScheduleModel":
        from .mlp import ScheduleMLP
        st = torch.load(path, map_location="cpu", weights_only=True)
        dims = st["mlp"]["dims"]
        mlp = ScheduleMLP(dims[0], dims[1], len(dims) - 1, device)
        mlp.load_state_dict(st["mlp"])
        return cls(mlp, [Schedule(*c) for c in st["cands"]])


def evaluate_choice(rows: List[Dict[str, Any]], model: ScheduleModel) -> Dict[str, float]:
    """Per group: time of the model's pick vs the exhaustive best and vs the default."""
    by: Dict[Tuple, List[Dict[str, Any]]] = {}
    for r in rows:
        by.setdefault((r.get("graph"), r["sig"]), []).append(r)
    regrets, gains, hits = [], [], 0
    for key, rs in by.items():
        pred = model.predict_rows(rs)
        pick = rs[int(np.argmin(pred))]["ms"]
        best = min(r["ms"] for r in rs)
        dflt = default_time(rs)
        regrets.append(pick / best - 1.0)
        if dflt is not None:
            gains.append(dflt / pick)
        hits += pick == best
    return {"groups": len(by), "mean_regret": float(np.mean(regrets)), "max_regret": float(np.max(regrets)),
            "top1": hits / max(len(by), 1), "geomean_gain_vs_default": float(np.exp(np.mean(np.log(gains)))) if gains else None}

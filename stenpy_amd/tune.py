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
    for tx in (32, 64, 128) for ty in (4, 8, 16) for zc in (32, 64, 128, 256, 512) for nt in (128, 256, 512)
    if nt % tx == 0 and ty % (nt // tx) == 0 and 1 <= tx * ty // nt <= 4
]
# the schedules the learned model was trained and is evaluated on (its timing rows predate the
# 512-plane z-chunks, which only the exhaustive search times)
MODEL_CANDIDATES: List[Schedule] = [c for c in CANDIDATES if c.zc <= 256]
# a smaller grid of candidates for the training-data sweep (the heuristic schedule of each
# grid is always added)
DATA_CANDIDATES: List[Schedule] = [c for c in CANDIDATES if c.zc in (64, 256) or (c.zc == 128 and c.tx == 64)]


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
    """Structure of a fused group (ops, edges, params, grid) -- the tuning DB key.

    Nodes are listed in a canonical order that depends only on the structure, not on the order
    in which a topological sort visits them: by depth inside the group, then by op name,
    params and the labels of their inputs (assigned in that same order)."""
    idx = {}
    parts = [f"grid={tuple(group.grid)}"]
    for i, s in enumerate(group.sources):
        idx[s] = f"s{i}"
    for i, d in enumerate(group.dscalars):
        idx[d] = f"d{i}"
    members = set(group.nodes)
    depth: Dict[str, int] = {}
    for nid in group.nodes:                     # group.nodes is topologically ordered
        depth[nid] = 1 + max([depth[x] for x in graph.node(nid).input_ids if x in members],
                             default=0)

    def params_of(node) -> str:
        return ",".join(f"{k}={v!r}" for k, v in sorted(node.params.items())
                        if isinstance(v, (int, float, str, bool, tuple, list)))

    n = 0
    for level in sorted(set(depth.values())):
        rows = []
        for nid in group.nodes:
            if depth[nid] != level:
                continue
            node = graph.node(nid)
            ins = ",".join(idx.get(x, "c") for x in node.input_ids)
            rows.append((node.op_name, params_of(node), ins, nid))
        for op, ps, ins, nid in sorted(rows):
            idx[nid] = f"n{n}"
            n += 1
            parts.append(f"{op}({ins})[{ps}]")
    parts.append("out=" + ",".join(idx[o] for o in group.outputs))
    return "|".join(parts)


_MODEL = None
DATA_DIR = Path(__file__).resolve().parent / "data"
MODEL_PATH = DATA_DIR / "schedule_mlp.npz"
ROWS_PATH = DATA_DIR / "tune_rows.jsonl.gz"


def install_model(model) -> None:
    """Use ``model`` (a trained ``ScheduleModel``) for groups missing from the database."""
    global _MODEL
    _MODEL = model


def load_default_model(device="cuda") -> Optional["ScheduleModel"]:
    """The committed trained schedule model (None when absent)."""
    global _MODEL
    if _MODEL is None and MODEL_PATH.exists():
        _MODEL = ScheduleModel.load(MODEL_PATH, device)
    return _MODEL


def schedule_fn(device: Optional[torch.device] = None, use_model: bool = True,
                threads: Optional[int] = None):
    """Schedule source for the fused executor: tuning database, then the learned model, then
    (by returning None) the executor's shape heuristic.  ``threads`` restricts both to one
    block size (the batched launches share one)."""
    db = load_db()

    def pick(group, graph) -> Optional[Schedule]:
        hit = db.get(group_signature(group, graph))
        if hit and (threads is None or hit[3] == threads):
            return Schedule(*hit)
        if use_model:
            m = load_default_model(device or "cuda")
            if m is not None:
                s = m.best(group, graph, threads)
                return F.fill_small_grid(s, group.grid) if s is not None else None
        return None

    return pick


# ------------------------------------------------------------------------------ batched launches
# per-structure schedules of a batched launch (one block size per batch module)
BATCH_CANDIDATES: List[Schedule] = [Schedule(tx, ty, zc, 256) for tx, ty in ((64, 8), (64, 4), (32, 16), (32, 8))
                                    for zc in (64, 128, 256)]


def _batch_key(structure: str, threads: int) -> str:
    return f"batch|threads={threads}|{structure}"


def batch_schedule_fn(threads: int = 256):
    """Schedule source for ``BatchRunner``: the per-structure winners ``batch_autotune`` stored
    in the tuning database, else (by returning None) the batch's shape heuristic."""
    from .batch import structure_key
    db = load_db()

    def pick(group, graph) -> Optional[Schedule]:
        hit = db.get(_batch_key(structure_key(group, graph), threads))
        return Schedule(*hit) if hit else None

    return pick


def batch_step_ms(runner, steps: int = 20, warmup: int = 3) -> float:
    """Milliseconds per replay of the runner's captured batch (HIP events)."""
    g = runner.capture()
    for _ in range(warmup):
        g.replay()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        g.replay()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) / steps


def batch_autotune(items, device, cands: Sequence[Schedule] = None, min_gain: float = 0.01,
                   update_db: bool = True, log=None) -> Dict[str, Any]:
    """Greedy per-structure search for a batched launch, timed on the whole batch.

    Starts from the batch heuristic for every structure; for each structure in turn, tries
    every candidate with the others fixed and keeps one that shortens the replayed step by more
    than ``min_gain``.  The result can only match or beat the heuristic on this batch.  Winners
    are stored in the tuning database under the structure's key (``batch_schedule_fn``), the
    structures the search left on the heuristic's schedule included."""
    from .batch import BATCH_THREADS, BatchRunner
    cands = list(cands or [c for c in BATCH_CANDIDATES if c.threads == BATCH_THREADS])
    chosen: Dict[str, Schedule] = {}

    def measure() -> Tuple[float, Any]:
        fn = lambda grp, gr: chosen.get(_structure(grp, gr))
        br = BatchRunner(items, device=device, schedule_fn=fn)
        br.compile(workers=8)
        ms = batch_step_ms(br)
        return ms, br

    from .batch import structure_key as _structure
    best, br = measure()
    base = best
    keys = [k for m in br.mods for k in m["keys"]]
    start = {k: tuple(g.sched.as_tuple()) for m in br.mods for k, g in zip(m["keys"], m["gens"])}
    del br
    if log:
        log(f"heuristic: {base:.4f} ms, {len(keys)} structures")
    trials = []
    for k in keys:
        for c in cands:
            if c.as_tuple() == start[k] and k not in chosen:
                continue
            prev = chosen.get(k)
            chosen[k] = c
            try:
                ms, br = measure()
                del br
            except (RuntimeError, ValueError) as exc:        # a schedule the structure cannot take
                ms = float("inf")
                if log:
                    log(f"  {c.as_tuple()} skipped: {exc}")
            trials.append({"structure": k[:80], "schedule": list(c.as_tuple()), "ms": ms})
            if ms < best * (1.0 - min_gain):
                best = ms
                if log:
                    log(f"  {k[:60]}: {c.as_tuple()} -> {ms:.4f} ms (kept)")
            else:
                if prev is None:
                    del chosen[k]
                else:
                    chosen[k] = prev
        torch.cuda.empty_cache()
    final = {k: (list(chosen[k].as_tuple()) if k in chosen else list(start[k])) for k in keys}
    if update_db:
        db = load_db()
        for k, c in final.items():
            db[_batch_key(k, BATCH_THREADS)] = c
        save_db(db)
    return {"heuristic_ms": base, "best_ms": best, "changed": {k[:80]: final[k] for k in chosen},
            "schedules": {k[:80]: c for k, c in final.items()}, "trials": trials}


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
        "transc": sum(text.count(f) for f in ("spm_expf(", "spm_logf(", "spm_tanhf(", "spm_sinf(", "spm_cosf(", "acos(")),
        "sqrt": text.count("sqrt"),
        "src_ch": src_ch, "out_ch": out_ch, "nred": len(g.reductions),
        "hz": gen.hz, "redge": gen.redge,
        "halo": max([l["H"] for l in gen.src_layout] + [0]),
        "bf16_out": sum(1 for o in g.outputs if gen.out_dtypes.get(o) == torch.bfloat16),
    }


def features(stats: Dict[str, float], grid: Sequence[int], s: Schedule, lds_bytes: int,
             face_wave: int = 0) -> List[float]:
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
        (n2 % s.tx != 0) + (n1 % s.ty != 0), min(n0, s.zc) / s.zc, face_wave / 64.0,
    ]


N_FEATURES = 29


# ------------------------------------------------------------------------------ search
_POOL = None


def _compile_all(sources: List[str], workers: int) -> List[bytes]:
    """Compile in a pool of spawned processes (hipRTC serialises compiles within a process).

    The workers only run hipRTC and never touch a device; each compiled code object also
    lands in the on-disk cache, so the parent's later ``load_module`` is a cache hit."""
    global _POOL
    if workers <= 1 or len(sources) <= 1:
        return [F.compile_source(s)[0] for s in sources]
    if _POOL is None:
        import multiprocessing as mp
        import atexit
        _POOL = cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn"))
        atexit.register(_POOL.shutdown)
    return [c for c, _ in _POOL.map(F.compile_source, sources)]


def _fit_all(gens: List[F.KernelGen], workers: int) -> Tuple[List[str], List[bytes]]:
    """The executor's register rule (``FusedExecutor._fit_registers``) applied to every
    candidate, each stage's recompiles batched through the compile pool: kernels above
    ``MERGED_TAIL_MAX_VGPRS`` first move the edge band to its own kernel, then drop staging.
    What is timed is what would run."""
    srcs = [g.source() for g in gens]
    codes = _compile_all(srcs, workers)
    over = lambda c: F.kernel_registers(c).get("fused_kernel", 0) > F.MERGED_TAIL_MAX_VGPRS
    for stage in ("split", "direct"):
        redo = []
        for j, (gen, code) in enumerate(zip(gens, codes)):
            if not over(code):
                continue
            if stage == "split" and F.SPLIT_TAIL == "auto" and gen.use_lds and not gen.batched \
                    and gen.boundary_blocks() > 0 and not gen.split_tail:
                gen.want_split = True
                redo.append(j)
            elif stage == "direct" and gen.stage:
                gen.stage, gen.stage_off, gen.stage_c = [], {}, 0
                redo.append(j)
        if redo:
            new = _compile_all([gens[j].source() for j in redo], workers)
            for j, c in zip(redo, new):
                srcs[j], codes[j] = gens[j].source(), c
    return srcs, codes


def candidates_for(graph, group, info, out_dtypes, cands: Sequence[Schedule] = None):
    gens = []
    cands = list(cands or CANDIDATES)
    h = F.heuristic_schedule(group.grid)
    if h not in cands:
        cands.append(h)
    for s in cands:
        try:
            gen = F.KernelGen(graph, group, info, s, out_dtypes, {})
        except ValueError:
            continue
        if gen.lds_bytes() <= F.LDS_LIMIT:
            gens.append(gen)
    return gens


def autotune(graph, runtime, wanted: Sequence[str], seed: Optional[Dict[str, Any]] = None,
             cands: Sequence[Schedule] = None, reps: int = 10, workers: int = 12,
             update_db: bool = True, out_dtype: Optional[torch.dtype] = None,
             log=None) -> List[Dict[str, Any]]:
    """Schedule search for every fused group of ``graph``; returns timing rows.

    ``cands``: candidate list (default: every valid candidate, i.e. exhaustive search), or a
    callable (group, graph) -> candidates, e.g. a trained model's shortlist.  The shape
    heuristic's schedule is always timed too."""
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
        group_cands = cands(st.group, graph) if callable(cands) else cands
        gens = candidates_for(graph, st.group, info, base.gen.out_dtypes, group_cands)
        srcs, codes = _fit_all(gens, workers)
        if log:
            log(f"group {si}: {len(gens)} candidates compiled")
        stats = group_stats(base.gen)
        sig = group_signature(st.group, graph)
        best = None
        for j, (gen, src, code) in enumerate(zip(gens, srcs, codes)):
            cg = F.CompiledGroup(gen, src)
            cg.module = F.load_module(src)
            ln = ex.prepare_launch(cg, graph, results)
            ms = ln.time_ms(reps=reps)
            nvox = int(np.prod(st.group.grid))
            rows.append({"sig": sig, "grid": list(st.group.grid), "schedule": list(gen.sched.as_tuple()),
                         "ms": ms, "ns_per_voxel": ms * 1e6 / nvox, "lds": gen.lds_bytes(),
                         "stats": stats, "features": features(stats, st.group.grid, gen.sched, gen.lds_bytes(), gen.fw)})
            if best is None or ms < best[0]:
                best = (ms, gen.sched)
            if log:
                log(f"  {gen.sched.as_tuple()} {ms:.4f} ms ({j + 1}/{len(gens)})")
        if update_db and best is not None:
            db[sig] = list(best[1].as_tuple())
    if update_db:
        save_db(db)
    return rows


def default_time(rows: List[Dict[str, Any]], sched: Optional[Schedule] = None) -> Optional[float]:
    """Time of ``sched`` (default: the shape heuristic of the rows' grid) among ``rows``."""
    if not rows:
        return None
    sched = sched or F.heuristic_schedule(rows[0]["grid"])
    for r in rows:
        if tuple(r["schedule"]) == sched.as_tuple():
            return r["ms"]
    return None


# ------------------------------------------------------------------------------ learned model
def random_group_rows(n_graphs: int, seed: int, runtime, cands: Sequence[Schedule] = None,
                      shape_range: Tuple[int, int] = (64, 256), reps: int = 5, workers: int = 12,
                      log=None, slab_frac: float = 0.0) -> List[Dict[str, Any]]:
    """Exhaustive timings of random op graphs (3-12 ops) on random grids (training data).

    ``slab_frac``: share of the grids drawn slab-like instead (128-320 planes of a 1024-2048
    square-ish cross-section), the shape of an out-of-core slab."""
    import random
    from . import randgraph
    from .compiler import compile_expression
    from . import gen
    rng = random.Random(seed)
    rows: List[Dict[str, Any]] = []
    for gi in range(n_graphs):
        expr, nops = randgraph.random_expression(rng, rng.randint(3, 12))
        shape = randgraph.random_shape(rng, *shape_range)
        if slab_frac > 0 and rng.random() < slab_frac:
            shape = (rng.randint(128, 320), rng.randint(1024, 2048), rng.randint(1024, 2048))
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
    """MLP over (group, schedule) features predicting a group-relative log time.

    Only the ranking of a group's candidates matters, so the target is log(time) minus the
    mean over that group's timed candidates: the network spends no capacity on how fast the
    group is overall.  ``best`` returns the argmin; ``shortlist`` the k best predictions, which
    the autotuner times (a k-candidate search instead of the exhaustive one)."""

    def __init__(self, mlp=None, cands: Sequence[Schedule] = None):
        self.mlp = mlp
        self.cands = list(cands or MODEL_CANDIDATES)

    def fit(self, rows: List[Dict[str, Any]], device="cuda", epochs: int = 40, seed: int = 0,
            weight_decay: float = 1e-3, verbose: bool = False):
        from .mlp import ScheduleMLP
        X = np.asarray([r["features"] for r in rows], dtype=np.float32)
        y = np.log(np.asarray([r["ns_per_voxel"] for r in rows], dtype=np.float64))
        groups: Dict[Tuple, List[int]] = {}
        for i, r in enumerate(rows):
            groups.setdefault((r.get("graph"), r["sig"]), []).append(i)
        for idx in groups.values():
            y[idx] -= y[idx].mean()
        self.mlp = ScheduleMLP(X.shape[1], 256, 4, device, seed)
        return self.mlp.fit(X, y.astype(np.float32), epochs=epochs, batch=256, lr=2e-3, seed=seed,
                            weight_decay=weight_decay, verbose=verbose)

    def predict_rows(self, rows: List[Dict[str, Any]]) -> np.ndarray:
        X = np.asarray([r["features"] for r in rows], dtype=np.float32)
        return self.mlp.predict(X)

    def best(self, group, graph, threads: Optional[int] = None) -> Optional[Schedule]:
        s = self.shortlist(group, graph, 1, threads)
        return s[0] if s else None

    def shortlist(self, group, graph, k: int = 3, threads: Optional[int] = None) -> List[Schedule]:
        info = getattr(group, "_info", None)
        try:
            base = F.KernelGen(graph, group, info, F.DEFAULT_SCHEDULE, {}, {})
        except Exception:
            return None
        stats = group_stats(base)
        feats, scheds = [], []
        for s in self.cands:
            if threads is not None and s.threads != threads:
                continue
            try:
                gen = F.KernelGen(graph, group, info, s, {}, {})
            except ValueError:
                continue
            if gen.lds_bytes() > F.LDS_LIMIT:
                continue
            feats.append(features(stats, group.grid, s, gen.lds_bytes(), gen.fw))
            scheds.append(s)
        if not scheds:
            return []
        pred = self.mlp.predict(np.asarray(feats, dtype=np.float32)).reshape(-1)
        return [scheds[i] for i in np.argsort(pred)[:k]]

    def save(self, path) -> None:
        st = self.mlp.state_dict()
        arrs = {"dims": np.asarray(st["dims"], dtype=np.int64),
                "cands": np.asarray([c.as_tuple() for c in self.cands], dtype=np.int64),
                "x_mean": st["x_mean"].numpy(), "x_std": st["x_std"].numpy(),
                "y_norm": np.asarray([st["y_mean"], st["y_std"]], dtype=np.float64)}
        for l, (w, b) in enumerate(zip(st["W"], st["b"])):
            arrs[f"W{l}"] = w.numpy()
            arrs[f"b{l}"] = b.numpy()
        with open(path, "wb") as fh:
            np.savez_compressed(fh, **arrs)

    @classmethod
    def load(cls, path, device="cuda") -> "ScheduleModel":
        from .mlp import ScheduleMLP
        z = np.load(path, allow_pickle=False)
        dims = [int(v) for v in z["dims"]]
        mlp = ScheduleMLP(dims[0], dims[1], len(dims) - 1, device)
        mlp.load_state_dict({"W": [torch.from_numpy(z[f"W{l}"]) for l in range(len(dims) - 1)],
                             "b": [torch.from_numpy(z[f"b{l}"]) for l in range(len(dims) - 1)],
                             "x_mean": torch.from_numpy(z["x_mean"]), "x_std": torch.from_numpy(z["x_std"]),
                             "y_mean": float(z["y_norm"][0]), "y_std": float(z["y_norm"][1])})
        return cls(mlp, [Schedule(*(int(v) for v in c)) for c in z["cands"]])


def evaluate_choice(rows: List[Dict[str, Any]], model: ScheduleModel, k: int = 1) -> Dict[str, float]:
    """Per group: time of the model's pick (the best of its ``k`` top predictions, as the
    shortlist search times them) vs the exhaustive best and vs the shape heuristic."""
    by: Dict[Tuple, List[Dict[str, Any]]] = {}
    for r in rows:
        by.setdefault((r.get("graph"), r["sig"]), []).append(r)
    regrets, gains, ceiling, hits = [], [], [], 0
    for key, rs in by.items():
        pred = np.asarray(model.predict_rows(rs)).reshape(-1)
        pick = min(rs[i]["ms"] for i in np.argsort(pred)[:k])
        best = min(r["ms"] for r in rs)
        dflt = default_time(rs)
        regrets.append(pick / best - 1.0)
        if dflt is not None:
            gains.append(dflt / pick)
            ceiling.append(dflt / best)
        hits += pick == best
    geo = lambda v: float(np.exp(np.mean(np.log(v)))) if v else None
    # ceiling: the exhaustive search's own gain over the heuristic, the most any pick can reach
    return {"k": k, "groups": len(by), "mean_regret": float(np.mean(regrets)), "max_regret": float(np.max(regrets)),
            "top1": hits / max(len(by), 1), "geomean_gain_vs_default": geo(gains),
            "geomean_ceiling_gain_vs_default": geo(ceiling)}


def config_choice(rows: List[Dict[str, Any]], model: "ScheduleModel",
                  db_sched: Optional[Sequence[int]] = None) -> Dict[str, Any]:
    """One group's exhaustive timings against the model's picks.

    The model ranks only the schedules it was trained on (``MODEL_CANDIDATES``); its top-1
    pick and the best of its top 3 (as the shortlist search would time them) are compared
    with the exhaustive best over *every* timed candidate, and with the best inside the
    model's own candidate set."""
    own = {c.as_tuple() for c in MODEL_CANDIDATES}
    mine = [r for r in rows if tuple(r["schedule"]) in own]
    best = min(rows, key=lambda r: r["ms"])
    best_own = min(mine, key=lambda r: r["ms"])
    order = np.argsort(np.asarray(model.predict_rows(mine)).reshape(-1))
    top1 = mine[int(order[0])]
    top3 = min((mine[int(i)] for i in order[:3]), key=lambda r: r["ms"])
    out = {"grid": rows[0]["grid"], "timed": len(rows), "model_candidates": len(mine),
           "best": {"schedule": best["schedule"], "ms": best["ms"]},
           "best_in_model_set": {"schedule": best_own["schedule"], "ms": best_own["ms"]},
           "model_top1": {"schedule": top1["schedule"], "ms": top1["ms"],
                          "regret": top1["ms"] / best["ms"] - 1.0,
                          "regret_in_model_set": top1["ms"] / best_own["ms"] - 1.0},
           "model_top3": {"schedule": top3["schedule"], "ms": top3["ms"],
                          "regret": top3["ms"] / best["ms"] - 1.0,
                          "regret_in_model_set": top3["ms"] / best_own["ms"] - 1.0}}
    h = default_time(rows)
    if h is not None:
        out["heuristic"] = {"schedule": list(F.heuristic_schedule(rows[0]["grid"]).as_tuple()), "ms": h,
                            "regret": h / best["ms"] - 1.0}
    add_shortlist(out)
    if db_sched is not None:
        hit = [r for r in rows if list(r["schedule"]) == list(db_sched)]
        if hit:
            out["database"] = {"schedule": list(db_sched), "ms": hit[0]["ms"],
                               "regret": hit[0]["ms"] / best["ms"] - 1.0}
    return out


def add_shortlist(entry: Dict[str, Any]) -> Dict[str, Any]:
    """The autotuner's shortlist search: the model's top 3 plus the shape heuristic, all
    timed, the fastest kept (``entry`` as ``config_choice`` makes it)."""
    picks = [entry["model_top3"]] + ([entry["heuristic"]] if "heuristic" in entry else [])
    p = min(picks, key=lambda r: r["ms"])
    entry["shortlist_top3_and_heuristic"] = {"schedule": p["schedule"], "ms": p["ms"],
                                             "regret": p["ms"] / entry["best"]["ms"] - 1.0}
    return entry


def summarize_choices(groups: List[Dict[str, Any]]) -> Dict[str, Any]:
    """Mean / max regret of each pick over the groups of a schedule study."""
    geo = lambda v: float(np.exp(np.mean(np.log(v)))) if v else None
    out: Dict[str, Any] = {"groups": len(groups),
                           "top1_exact": sum(e["model_top1"]["schedule"] == e["best"]["schedule"] for e in groups)}
    for key in ("model_top1", "model_top3", "shortlist_top3_and_heuristic", "heuristic", "database"):
        vals = [e[key]["regret"] for e in groups if key in e]
        if vals:
            out[key] = {"groups": len(vals), "mean_regret": float(np.mean(vals)), "max_regret": float(np.max(vals))}
    out["top1_geomean_gain_vs_heuristic"] = geo([e["heuristic"]["ms"] / e["model_top1"]["ms"]
                                                 for e in groups if "heuristic" in e])
    out["exhaustive_geomean_gain_vs_heuristic"] = geo([e["heuristic"]["ms"] / e["best"]["ms"]
                                                       for e in groups if "heuristic" in e])
    return out


def config_groups(dev, which: Sequence[int] = (1, 2, 3, 4, 5)):
    """(label, graph, wanted) of the benchmark configurations' fused evaluations; config 3
    contributes the first field of each of its six graphs, config 4 one padded slab."""
    from . import workloads as W
    if 1 in which:
        for e, g, w in W.plain_graphs(dev):
            yield f"C1 128^3 {e}", g, w
    if 2 in which:
        f, u = W.stencil_inputs(dev)
        g, sink, reds = W.stencil_graph(f, u)
        del f, u
        yield "C2 512^3 stencil graph", g, [sink] + reds
    if 3 in which:
        exprs, items = W.batched_items(dev)
        for k in range(6):
            g, w = items[k]
            yield f"C3 graph {k} on field {k}: {exprs[k]}", g, w
        del items
    if 4 in which:
        g, w = W.stream_slab_graph(dev)
        yield "C4 slab (258, 4096, 4096)", g, w
    if 5 in which:
        g, sinks, _ = W.tensor_graph(dev)
        yield "C5 768^3 tensor graph", g, sinks


def score_rows(rows: List[Dict[str, Any]], model: "ScheduleModel", db: Optional[Dict[str, list]] = None,
               log=None) -> Dict[str, Any]:
    """``config_choice`` for every group of saved study rows (each row carries its
    ``config`` label), with the summary."""
    db = db or {}
    by: Dict[Tuple[str, str], List[Dict[str, Any]]] = {}
    for r in rows:
        by.setdefault((r.get("config", ""), r["sig"]), []).append(r)
    groups = []
    for (label, sig), rs in by.items():
        entry = {"config": label, **config_choice(rs, model, db.get(sig))}
        groups.append(entry)
        if log:
            log(json.dumps({k: entry[k] for k in ("config", "best", "model_top1", "model_top3")}))
    return {"groups": groups, "summary": summarize_choices(groups)}


def config_regret(dev, which: Sequence[int] = (1, 2, 3, 4, 5), reps: int = 10, workers: int = 12,
                  log=None, rows_out: Optional[List[Dict[str, Any]]] = None) -> Dict[str, Any]:
    """Exhaustive search over every candidate for each fused group of the benchmark
    configurations, scored against the committed model's top-1 / top-3 picks.  The timing
    rows are appended to ``rows_out`` when given (``score_rows`` re-scores them)."""
    from .runtime import Runtime
    model = load_default_model(dev)
    if model is None:
        raise RuntimeError(f"no trained schedule model at {MODEL_PATH}")
    db = load_db()
    report: Dict[str, Any] = {"groups": []}
    for label, g, wanted in config_groups(dev, which):
        if log:
            log(f"== {label}")
        rt = Runtime(device=dev, fusion_enabled=True)
        rows = autotune(g, rt, wanted, cands=CANDIDATES, reps=reps, workers=workers, update_db=False,
                        log=log)
        by: Dict[str, List[Dict[str, Any]]] = {}
        for r in rows:
            by.setdefault(r["sig"], []).append(r)
            if rows_out is not None:
                rows_out.append({**r, "config": label})
        for sig, rs in by.items():
            entry = {"config": label, **config_choice(rs, model, db.get(sig))}
            report["groups"].append(entry)
            if log:
                log(json.dumps({k: entry[k] for k in ("config", "best", "model_top1", "model_top3")}))
        del rt, g
        torch.cuda.empty_cache()
    report["summary"] = summarize_choices(report["groups"])
    return report


# ------------------------------------------------------------------------------ command line
def _read_rows(path) -> List[Dict[str, Any]]:
    import gzip
    rows = []
    with gzip.open(path, "rt") as fh:
        for line in fh:
            if line.strip():
                rows.append(json.loads(line))
    return rows


def _append_rows(path, rows: List[Dict[str, Any]]) -> None:
    import gzip
    with gzip.open(path, "at") as fh:
        for r in rows:
            fh.write(json.dumps(r) + "\n")


def split_by_graph(rows: List[Dict[str, Any]], holdout_every: int = 5):
    train = [r for r in rows if r.get("graph", 0) % holdout_every != 0]
    test = [r for r in rows if r.get("graph", 0) % holdout_every == 0]
    return train, test


def train_and_report(rows: List[Dict[str, Any]], device="cuda", epochs: int = 40,
                     seed: int = 0, verbose: bool = False) -> Tuple["ScheduleModel", Dict[str, Any]]:
    """Fit on 4/5 of the graphs, report regret / gain on the held-out fifth (pick = the top
    prediction, and the best of the top three as the shortlist search would time them)."""
    train, test = split_by_graph(rows)
    model = ScheduleModel(cands=MODEL_CANDIDATES)
    losses = model.fit(train, device=device, epochs=epochs, seed=seed, verbose=verbose)
    rep = {"rows_train": len(train), "rows_test": len(test), "epochs": epochs,
           "graphs": len({r.get("graph") for r in rows}),
           "loss_first": losses[0], "loss_last": losses[-1],
           "train": evaluate_choice(train, model), "heldout": evaluate_choice(test, model),
           "heldout_top3": evaluate_choice(test, model, k=3)}
    return model, rep


def main(argv=None) -> int:
    import argparse
    import time
    ap = argparse.ArgumentParser(prog="python -m stenpy_amd.tune")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect", help="exhaustive timings of random graphs (training rows)")
    c.add_argument("--graphs", type=int, default=100)
    c.add_argument("--first", type=int, default=0)
    c.add_argument("--seed", type=int, default=9)
    c.add_argument("--out", default=str(ROWS_PATH))
    c.add_argument("--min-shape", type=int, default=48)
    c.add_argument("--max-shape", type=int, default=256)
    c.add_argument("--time-limit", type=float, default=1e9, help="stop after this many seconds")
    c.add_argument("--workers", type=int, default=12)
    c.add_argument("--slab-frac", type=float, default=0.0, help="share of slab-shaped grids")
    c.add_argument("--all-candidates", action="store_true",
                   help="time every candidate (512-plane z-chunks included), not the data subset")
    t = sub.add_parser("train", help="train the schedule MLP on collected rows")
    t.add_argument("--rows", default=str(ROWS_PATH))
    t.add_argument("--out", default=str(MODEL_PATH))
    t.add_argument("--report", default=str(DATA_DIR / "schedule_mlp_report.json"))
    t.add_argument("--epochs", type=int, default=40)
    u = sub.add_parser("tune", help="exhaustive search for one expression; stores the winner")
    u.add_argument("--expr", required=True, action="append",
                   help="repeat to tune one graph with a sink per expression")
    u.add_argument("--shape", required=True, help="N0xN1xN2")
    u.add_argument("--boundary", default="periodic")
    u.add_argument("--reductions", default="")
    r = sub.add_parser("configs", help="exhaustive search on the benchmark configurations' groups, "
                                       "scored against the trained model (regret report)")
    r.add_argument("--which", default="1,2,3,4,5", help="comma-separated configuration numbers")
    r.add_argument("--out", default=None, help="write the JSON report here")
    r.add_argument("--workers", type=int, default=12)
    r.add_argument("--save-rows", default=None, help="also write the timing rows (gzipped JSON lines)")
    bt = sub.add_parser("batch", help="greedy per-structure schedule search for the 96-grid batch")
    bt.add_argument("--out", default=None)
    q = sub.add_parser("rescore", help="score a model against saved configuration-study rows")
    q.add_argument("--rows", required=True)
    q.add_argument("--model", default=str(MODEL_PATH))
    q.add_argument("--out", default=None)
    a = ap.parse_args(argv)

    from .runtime import Runtime
    dev = torch.device("cuda", 0)
    if a.cmd == "collect":
        import random
        rt = Runtime(device=dev, fusion_enabled=True)
        t0 = time.time()
        Path(a.out).parent.mkdir(parents=True, exist_ok=True)
        n_rows = 0
        for gi in range(a.first, a.first + a.graphs):
            if time.time() - t0 > a.time_limit:
                print(f"time limit reached after graph {gi - 1}", flush=True)
                break
            rows = random_group_rows(1, a.seed * 100003 + gi, rt,
                                     cands=CANDIDATES if a.all_candidates else DATA_CANDIDATES,
                                     shape_range=(a.min_shape, a.max_shape), workers=a.workers,
                                     log=lambda m: print(m, flush=True), slab_frac=a.slab_frac)
            for r in rows:
                r["graph"] = gi
            _append_rows(a.out, rows)
            n_rows += len(rows)
            rt.fused._plans.clear()
            print(f"[{time.time() - t0:.0f}s] graph {gi}: {len(rows)} rows, total {n_rows}", flush=True)
        return 0
    if a.cmd == "configs":
        which = [int(v) for v in a.which.split(",") if v]
        t0 = time.time()
        saved: List[Dict[str, Any]] = []
        rep = config_regret(dev, which, workers=a.workers,
                            log=lambda m: print(f"[{time.time() - t0:.0f}s] {m}", flush=True), rows_out=saved)
        if a.save_rows:
            _append_rows(a.save_rows, saved)
        rep["command"] = "python -m stenpy_amd.tune configs --which " + a.which
        text = json.dumps(rep, indent=1) + "\n"
        if a.out:
            Path(a.out).write_text(text)
        print(json.dumps(rep["summary"]), flush=True)
        return 0
    if a.cmd == "batch":
        from . import workloads as W
        t0 = time.time()
        _, items = W.batched_items(dev)
        rep = batch_autotune(items, dev, log=lambda m: print(f"[{time.time() - t0:.0f}s] {m}", flush=True))
        if a.out:
            Path(a.out).write_text(json.dumps(rep, indent=1) + "\n")
        print(json.dumps({k: rep[k] for k in ("heuristic_ms", "best_ms", "changed")}), flush=True)
        return 0
    if a.cmd == "rescore":
        model = ScheduleModel.load(a.model, dev)
        rep = score_rows(_read_rows(a.rows), model, load_db())
        rep["model"] = a.model
        if a.out:
            Path(a.out).write_text(json.dumps(rep, indent=1) + "\n")
        print(json.dumps(rep["summary"]), flush=True)
        return 0
    if a.cmd == "train":
        rows = _read_rows(a.rows)
        model, rep = train_and_report(rows, device=dev, epochs=a.epochs, verbose=True)
        model.save(a.out)
        Path(a.report).write_text(json.dumps(rep, indent=1) + "\n")
        print(json.dumps(rep), flush=True)
        return 0
    if a.cmd == "tune":
        from . import gen
        from .compiler import compile_expressions
        shape = tuple(int(v) for v in a.shape.lower().split("x"))
        f = gen.generate(shape, 2, dev)
        u_ = gen.generate_vector(shape, (3, 4, 5), dev)
        g, sinks, *_ = compile_expressions(a.expr, boundary=a.boundary, field_map={"f": f, "u": u_})
        reds = [g.add(r, (sinks[0],), {}) for r in a.reductions.split(",") if r]
        rt = Runtime(device=dev, fusion_enabled=True)
        rows = autotune(g, rt, sinks + reds, log=lambda m: print(m, flush=True))
        by = {}
        for r in rows:
            by.setdefault(r["sig"], []).append(r)
        for sig, rs in by.items():
            rs.sort(key=lambda r: r["ms"])
            print(json.dumps({"grid": rs[0]["grid"], "best": rs[0]["schedule"], "best_ms": rs[0]["ms"],
                              "heuristic_ms": default_time(rs), "n": len(rs)}), flush=True)
        return 0
    return 2


if __name__ == "__main__":
    raise SystemExit(main())

"""Schedule selection for fused groups: tuning database lookup (autotuner fills it).

A schedule is (TX, TY, ZC, threads).  ``schedule_fn(device)`` returns the callback the fused
executor consults per group; it answers from the on-disk database of measured best
schedules keyed by the group's structural signature, else ``None`` (default schedule).
"""
from __future__ import annotations

import json
import os
import threading
from pathlib import Path
from typing import Dict, Optional

import torch

from .fused import Schedule

_DB_LOCK = threading.Lock()


def db_path() -> Path:
    d = os.environ.get("STENPY_TUNE_DB")
    if d:
        return Path(d)
    return Path(__file__).resolve().parent / "tuning_db.json"


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
    parts = [f"grid={group.grid}"]
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


def schedule_fn(device: Optional[torch.device] = None):
    db = load_db()

    def pick(group, graph) -> Optional[Schedule]:
        hit = db.get(group_signature(group, graph))
        return Schedule(*hit) if hit else None

    return pick

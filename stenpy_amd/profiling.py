"""Readers for rocprofv3 CSV output (kernel statistics and PMC counters).

``rocprofv3 --kernel-trace --stats --output-format csv`` writes ``*_kernel_stats.csv`` and
``rocprofv3 --pmc FETCH_SIZE --output-format csv`` writes ``*_counter_collection.csv``.  The
counters FETCH_SIZE / WRITE_SIZE report KiB per dispatch; they are calibrated against a
streaming kernel of known size (read n bytes, write n bytes) run in the same process, so
that ``traffic`` is in bytes of HBM per launch of the measured kernel.
"""
from __future__ import annotations

import csv
import json
from pathlib import Path
from typing import Dict, List, Optional


def _find(d: str, suffix: str) -> List[Path]:
    return sorted(Path(d).rglob(f"*{suffix}"))


def _col(row: Dict[str, str], *names: str) -> Optional[str]:
    low = {k.lower().replace("-", "_"): v for k, v in row.items()}
    for n in names:
        if n in low:
            return low[n]
    return None


def kernel_stats(d: str) -> List[Dict[str, object]]:
    """Rows of every ``*kernel_stats.csv`` under ``d`` (name, calls, total/average ns)."""
    out = []
    for p in _find(d, "kernel_stats.csv"):
        with open(p, newline="") as fh:
            for row in csv.DictReader(fh):
                out.append({"name": _col(row, "name", "kernel_name"),
                            "calls": int(float(_col(row, "calls") or 0)),
                            "total_ns": float(_col(row, "totaldurationns", "total_duration_ns") or 0),
                            "avg_ns": float(_col(row, "averagens", "average_ns") or 0),
                            "pct": float(_col(row, "percentage") or 0)})
    return out


def counter_values(d: str, counter: str) -> Dict[str, List[float]]:
    """Per kernel name, the list of per-dispatch values of ``counter``."""
    out: Dict[str, List[float]] = {}
    for p in _find(d, "counter_collection.csv"):
        with open(p, newline="") as fh:
            for row in csv.DictReader(fh):
                if (_col(row, "counter_name") or "") != counter:
                    continue
                name = _col(row, "kernel_name") or "?"
                out.setdefault(name, []).append(float(_col(row, "counter_value") or 0))
    return out


def _pick(vals: Dict[str, List[float]], needle: str) -> List[float]:
    hits = [v for k, v in vals.items() if needle in k]
    if not hits:
        raise KeyError(f"no dispatch of a kernel named like {needle!r} (have {sorted(vals)[:8]})")
    return [x for h in hits for x in h]


def traffic(fetch_dir: str, write_dir: str, kernel: str, copy_kernel: str,
            copy_bytes: int) -> Dict[str, float]:
    """HBM bytes per launch of ``kernel`` from FETCH_SIZE and WRITE_SIZE passes.

    Each counter is scaled by known/measured for ``copy_kernel``, which reads and writes
    ``copy_bytes`` bytes per dispatch.  The median over dispatches is used throughout.
    """
    med = lambda xs: sorted(xs)[len(xs) // 2]
    f = counter_values(fetch_dir, "FETCH_SIZE")
    w = counter_values(write_dir, "WRITE_SIZE")
    kf, kw = med(_pick(f, kernel)), med(_pick(w, kernel))
    cf, cw = med(_pick(f, copy_kernel)), med(_pick(w, copy_kernel))
    sf = copy_bytes / (cf * 1024.0)
    sw = copy_bytes / (cw * 1024.0)
    return {"fetch_kib": kf, "write_kib": kw, "copy_fetch_kib": cf, "copy_write_kib": cw,
            "copy_bytes": copy_bytes, "scale_fetch": sf, "scale_write": sw,
            "fetch_bytes": kf * 1024.0 * sf, "write_bytes": kw * 1024.0 * sw,
            "bytes_per_launch": kf * 1024.0 * sf + kw * 1024.0 * sw}


def write_traffic(path: str, **kw) -> Dict[str, float]:
    t = traffic(**kw)
    Path(path).write_text(json.dumps(t, indent=1) + "\n")
    return t


def stats_markdown(d: str, command: str, top: int = 12) -> str:
    rows = sorted(kernel_stats(d), key=lambda r: -r["total_ns"])[:top]
    lines = [f"Command: `{command}`", "",
             "| kernel | calls | avg (us) | total (ms) | % |", "|---|---|---|---|---|"]
    for r in rows:
        name = str(r["name"])
        name = name if len(name) < 90 else name[:87] + "..."
        lines.append(f"| `{name}` | {r['calls']} | {r['avg_ns'] / 1e3:.2f} | {r['total_ns'] / 1e6:.3f} | {r['pct']:.1f} |")
    return "\n".join(lines) + "\n"


def trace_tail(d: str, kernel: str, n: int) -> Dict[str, float]:
    """Durations of the last ``n`` dispatches (by start time) of the kernel named ``kernel`` in
    every ``*kernel_trace.csv`` under ``d``: the launches a timed loop ends with."""
    rows = []
    for p in _find(d, "kernel_trace.csv"):
        with open(p, newline="") as fh:
            for row in csv.DictReader(fh):
                if (_col(row, "kernel_name") or "") == kernel:
                    t0, t1 = float(_col(row, "start_timestamp")), float(_col(row, "end_timestamp"))
                    rows.append((t0, t1 - t0, int(float(_col(row, "vgpr_count") or 0))))
    rows.sort()
    tail = [r[1] for r in rows[-n:]]
    if not tail:
        raise KeyError(f"no dispatch of {kernel!r} under {d}")
    return {"dispatches": len(rows), "tail": len(tail), "mean_ns": sum(tail) / len(tail),
            "min_ns": min(tail), "max_ns": max(tail), "vgprs": rows[-1][2]}

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


# ------------------------------------------------------------------------------ run reports
HBM_PEAK_GBS = 8000.0
# arithmetic operations of the device helpers the fused code calls (FP32 or FP64 operations
# of their bodies in csrc/sp_sym3.h, csrc/sp_math32.h and csrc/fused_template.h, counted once)
_CALL_OPS = {"tg_mid": 2, "tg_left2": 5, "tg_right2": 5, "unit_rnorm": 8, "div3f": 4,
             "sym3_trace": 2, "sym3_det_regularised": 20, "sym3_deviatoric": 13,
             "sym3_eigvalsh": 120, "mat3_inverse": 30, "spm_expf": 16, "spm_tanhf": 24,
             "spm_logf": 18, "spm_sinf": 14, "spm_cosf": 14}


def code_ops_per_voxel(lines: List[str]) -> int:
    """Arithmetic operations per voxel of generated interior code: each +, -, *, / between
    operands plus the operations of the helper functions it calls (``_CALL_OPS``)."""
    import re
    text = "\n".join(lines)
    n = len(re.findall(r"(?<=[\w\)\]]) [-+*/] (?=[\w\(\-])", text))
    for fn, k in _CALL_OPS.items():
        n += k * len(re.findall(rf"\b{fn}\(", text))
    return n


def fused_run_report(rt, graph, wanted, reps: int = 3) -> Dict[str, object]:
    """Time, traffic and arithmetic of one fused evaluation of ``graph`` on ``rt`` (plans are
    compiled by a first run): HIP-event milliseconds of the whole evaluation, the algorithmic
    bytes of its fused groups (every source channel read once, every materialised output
    written once; fused reductions add nothing), the arithmetic operations of the generated
    interior code, the achieved rates, and the rocprofv3 command that measures the same run's
    HBM counters."""
    import torch
    from . import fused as F
    results = rt.fused.run(graph, rt, wanted=wanted)
    values = {n: results[n] for n in graph._order if graph.node(n).op_name == "_constant"}
    steps, info, comp = rt.fused.prepare(graph, values, wanted)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        rt.fused.run(graph, rt, wanted=wanted)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    read = written = ops = voxels = 0
    groups = []
    for si, cg in comp.items():
        gen = cg.gen
        nvox = 1
        for s in cg.group.grid:
            nvox *= int(s)
        r = sum(max(L_["C"], 1) for L_ in gen.src_layout) * 4 * nvox
        w = sum(max(gen.info[o].ch, 1) * (2 if gen.out_dtypes.get(o) == torch.bfloat16 else 4)
                for o in cg.group.outputs) * nvox
        e = F._Emit(gen, ("lds", False, False)) if gen.use_lds else F._Emit(gen, ("gather",))
        pos = e.origin()
        for nid in cg.group.outputs:
            e.val(nid, pos)
        for nid in cg.group.reductions:
            e.val(graph.node(nid).input_ids[0], pos)
        k = code_ops_per_voxel(e.lines)
        groups.append({"grid": list(cg.group.grid), "schedule": list(gen.sched.as_tuple()),
                       "bytes_read": r, "bytes_written": w, "ops_per_voxel": k})
        read, written, ops, voxels = read + r, written + w, ops + k * nvox, max(voxels, nvox)
    s = ms * 1e-3
    sig = lambda x: float(f"{x:.5g}")          # five significant digits
    return {"gpu_ms": sig(ms), "voxels": voxels, "gvoxel_s": sig(voxels / s / 1e9),
            "bytes_read": read, "bytes_written": written,
            "achieved_gbs": sig((read + written) / s / 1e9),
            "hbm_frac": sig((read + written) / s / 1e9 / HBM_PEAK_GBS),
            "arith_ops": ops, "arith_gops": sig(ops / s / 1e9), "groups": groups,
            "counters": "rocprofv3 --pmc FETCH_SIZE (then WRITE_SIZE) --output-format csv -d DIR -o run -- "
                        "<this command>: HBM bytes per kernel dispatch (KiB)"}

"""Command line: ``python -m stenpy_amd --field f=data.npy --expr "curl(f)" --device cuda``.

The reference documents this entry point (README.md:21) without shipping it.  Fields are
``name=path`` (``.npy``; HDF5 when h5py is importable) or ``name=gen:SEED:N0xN1xN2`` for the
seeded synthetic field.  The result is written with ``--output`` (``.npy``) or summarised.
``--stream`` evaluates out of core in axis-0 slabs.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from typing import Dict, List


def _parse_field(spec: str):
    if "=" not in spec:
        raise SystemExit(f"--field expects name=path, got {spec!r}")
    name, src = spec.split("=", 1)
    return name.strip(), src.strip()


def main(argv: List[str] = None) -> int:
    ap = argparse.ArgumentParser(prog="stenpy", description="Evaluate a Stenpy expression on the GPU")
    ap.add_argument("--field", action="append", required=True, help="name=path.npy or name=gen:SEED:N0xN1xN2")
    ap.add_argument("--expr", required=True)
    ap.add_argument("--device", default="cuda")
    ap.add_argument("--boundary", default="periodic")
    ap.add_argument("--dx", type=float, default=1.0)
    ap.add_argument("--output", default=None)
    ap.add_argument("--stream", action="store_true", help="evaluate out of core in axis-0 slabs")
    ap.add_argument("--chunk-mb", type=float, default=None)
    ap.add_argument("--no-fusion", action="store_true")
    ap.add_argument("--perf-report", action="store_true")
    a = ap.parse_args(argv)

    import numpy as np
    import torch
    import stenpy_amd as sp
    from stenpy_amd import gen
    from stenpy_amd.pipeline import load_tensor, save_tensor
    from stenpy_amd.stream import GeneratorField, NpyLazyField, stream_graph

    t0 = time.perf_counter()
    fields: Dict[str, object] = {}
    for spec in a.field:
        name, src = _parse_field(spec)
        if src.startswith("gen:"):
            _, seed, dims = src.split(":")
            shape = tuple(int(v) for v in dims.lower().split("x"))
            fields[name] = GeneratorField(shape, int(seed)) if a.stream else gen.generate(shape, int(seed), a.device)
        elif a.stream:
            fields[name] = NpyLazyField(src)
        else:
            fields[name] = load_tensor(src, a.device)[0]
    if a.stream:
        if len(fields) != 1:
            raise SystemExit("--stream supports exactly one field")
        (name, lf), = fields.items()
        probe = torch.zeros((1,) + tuple(lf.shape[1:]))
        g, sink, warns, fm, _ = sp.compile_expression(a.expr, dx=a.dx, boundary=a.boundary, field_map={name: probe})
        rt = sp.Runtime(device=a.device, fusion_enabled=not a.no_fusion)
        stats = stream_graph(g, sink, fm[name], lf, a.output, rt, boundary=a.boundary, chunk_mb=a.chunk_mb)
        print(json.dumps(stats, default=str))
        return 0
    g, sink, warns, fm, _ = sp.compile_expression(a.expr, dx=a.dx, boundary=a.boundary, field_map=fields)
    for w in warns:
        print(f"warning: {w}", file=sys.stderr)
    rt = sp.Runtime(device=a.device, fusion_enabled=not a.no_fusion)
    out = rt.run(g)[sink]
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    if a.output:
        save_tensor(out, (a.dx,) * 3, (0.0,) * 3, a.output)
    summary = {"expr": a.expr, "shape": list(out.shape), "path": rt.last_path, "seconds": round(el, 4)}
    o = out.double()
    summary.update({"min": float(o.min()), "max": float(o.max()), "mean": float(o.mean())})
    if a.perf_report:
        summary["gvoxel_s"] = out.numel() / el / 1e9          # end to end, input loading included
        if rt.last_path == "fused":
            # the evaluation alone, re-run warm: HIP-event time, algorithmic bytes of the fused
            # groups and the arithmetic of their generated code (stenpy.py:355-394 reports
            # wall time and bytes; its flop counter is never incremented)
            from stenpy_amd.profiling import fused_run_report
            summary["perf"] = fused_run_report(rt, g, [sink])
    print(json.dumps(summary))
    return 0


if __name__ == "__main__":
    sys.exit(main())

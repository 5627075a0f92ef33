"""Benchmark: the 512^3 stencil graph, fused/autotuned vs plain per-operator HIP.

One step = one evaluation of
    out = divergence(gradient(f)) + dot(u, gradient(f)) + sqnorm(curl(u))
plus sum/mean/variance/L2 of ``out`` on a 512^3 FP32 field f and a 512^3 x 3 channel-last
vector field u (seeded Fourier inputs, periodic, dx = 1), inputs resident in HBM.

Prints one JSON line.  ``value`` is the fused path's Gvoxel/s; ``roofline`` prices the fused
kernel's algorithmic bytes (20 B/voxel: read f and u, write out) over its average duration
measured with HIP events on the launch stream; ``cpu_baseline`` times the FP64 CPU oracle (the
reference's torch eager semantics) on a bounded slab of the same workload.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def build_graph(f, u):
    from stenpy_amd.workloads import stencil_graph
    return stencil_graph(f, u)


def time_steps(fn, steps: int, warmup: int):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def group_launch(rt, g, res, sink, out_dtype=None):
    """The prepared launch (main + boundary kernel) of the fused group producing ``sink``."""
    ex = rt.fused
    for steps, info, comp in ex._plans.values():
        for si, cg in comp.items():
            if sink in cg.group.outputs and cg.gen.out_dtypes.get(sink) == out_dtype:
                return cg, ex.prepare_launch(cg, g, res)
    raise RuntimeError("fused group not found")


def cpu_model() -> str:
    """The host CPU as /proc/cpuinfo names it, with family / model numbers (a virtual machine
    may report only a generic name such as "AMD EPYC") and the visible CPU count."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("model name", "cpu family", "model", "vendor_id") and k not in info:
                info[k] = v.strip()
    except OSError:
        pass
    import platform
    name = info.get("model name") or platform.processor() or "unknown"
    if "cpu family" in info:
        name += f" (family {info['cpu family']}, model {info.get('model', '?')})"
    return f"{name}, {os.cpu_count()} CPUs visible"


def cpu_threads() -> int:
    """Host threads to use: the box's share (OMP_NUM_THREADS), else every visible CPU."""
    try:
        return max(1, int(os.environ.get("OMP_NUM_THREADS", "")))
    except ValueError:
        return os.cpu_count() or 1


def oracle_leg(f, u, res, sink, reds):
    """The FP64 CPU oracle (the reference's torch eager semantics) on the benchmark's own
    inputs, over the whole field: timed as the CPU baseline, and used with its FP32 twin for
    the error rule of the measured run (e_build = max|GPU - oracle64|, e_ref32 = max|ref32 -
    oracle64|, for the output field and each reduction)."""
    import torch
    from oracle import graph_eval
    threads = cpu_threads()
    torch.set_num_threads(threads)
    fc, uc = f.cpu(), u.cpu()
    g, s_, r_ = build_graph(fc, uc)
    t0 = time.perf_counter()
    o64 = graph_eval.evaluate(g, torch.float64)
    t64 = time.perf_counter() - t0
    o32 = graph_eval.evaluate(g, torch.float32)
    errs = {}
    for name, nid_b, nid_o in [("out", sink, s_)] + [(g.node(b).op_name, a, b) for a, b in zip(reds, r_)]:
        b = res[nid_b].detach().double().cpu()
        ref = o64[nid_o].double()
        e_b = (b - ref).abs().max().item()
        e_32 = (o32[nid_o].double() - ref).abs().max().item()
        errs[name] = {"e_build": e_b, "e_ref32": e_32, "rule_2x": bool(e_b <= 2.0 * e_32)}
    vox = f.numel()
    base = {"value": round(vox / t64 / 1e9, 5), "unit": "Gvoxel/s", "cores": threads, "kind": "port",
            "sample": f"whole {tuple(f.shape)} field, FP64 torch CPU oracle, one timed evaluation",
            "cpu": cpu_model(),
            "note": "lower bound on the reference's CPU time: the restatement omits its "
                    "executor's per-allocation gc.collect and zero fill"}
    return base, errs


def hbm_ceilings(f, u, reps: int = 10):
    """What this GPU sustains for three read:write mixes, from the library's streaming probes
    on the benchmark's own buffers (f: n^3 FP32, u: n^3 x 3): copy of f (1:1), the stencil
    graph's mix (read f and u, write one float per voxel: 16 B read, 4 B written, 4:1; u read
    as three planar thirds, and channel-last as the graph holds it), a read-only pass over u and
    a write-only pass.  HIP events on the launch stream; the best of the flat and the
    persistent launch shape, each the mean of ``reps`` back-to-back launches.  GB/s."""
    import torch
    from stenpy_amd import _lib as L
    n = f.numel()
    out = torch.empty_like(f)
    cases = {"copy": (8 * n, dict(a=f, out=out)), "4r1w": (20 * n, dict(a=f, u=u, out=out)),
             "4r1w_aos": (20 * n, dict(a=f, u=u, out=out)),
             "read": (12 * n, dict(a=u)), "write": (4 * n, dict(a=out))}
    res = {}
    for mix, (nbytes, kw) in cases.items():
        best = 0.0
        for persistent in (False, True):
            if mix == "read":
                kw["partials"] = torch.empty(L.probe_blocks(persistent, kw["a"].numel()),
                                             dtype=torch.float64, device=f.device)
            for _ in range(3):
                L.probe_stream(mix, persistent, **kw)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                L.probe_stream(mix, persistent, **kw)
            e1.record()
            e1.synchronize()
            best = max(best, nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9)
        res[mix] = round(best, 1)
    del out
    return res


def schedule_source(cg, g) -> str:
    from stenpy_amd import tune
    if tune.load_db().get(tune.group_signature(cg.group, g)):
        return "autotuned (exhaustive search, tuning database)"
    return "learned model or shape heuristic"


def _line(metric, value, unit, steps, warmup, ms, config, **extra):
    d = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps,
         "warmup": warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
         "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded Fourier fields)",
         "config": config}
    d.update(extra)
    print(json.dumps(d), flush=True)


def c3_items(dev):
    from stenpy_amd.workloads import batched_items
    return batched_items(dev)


def c5_graph(dev, n: int = 768):
    from stenpy_amd.workloads import tensor_graph
    return tensor_graph(dev, n)


def bench_batched(args) -> int:
    """96 fields, axes U{32..384} (seed 6), 6 random graphs of 3-12 ops (seed 6); field i uses
    graph i mod 6 and seed 1000 + i; every level one persistent launch, replayed as a hipGraph."""
    import random
    import numpy as np
    import torch
    import stenpy_amd as sp
    from stenpy_amd import batch as B, gen, randgraph, tune
    dev = torch.device("cuda:0")
    exprs, items = c3_items(dev)
    t0 = time.perf_counter()
    # default: the per-structure winners of the batch autotuner (python -m stenpy_amd.tune batch)
    # from the tuning database, the batch's shape heuristic for structures it has none for.  The
    # model's picks, ranked on single-field timings, favour short z-chunks that cost more inside
    # one shared persistent launch (measured 217 against 247 Gvoxel/s); ``--schedules model``
    # runs them (one block size per module).
    src = {"model": "the learned model", "heuristic": "batch shape heuristic"}.get(
        args.schedules, "batch autotuner (tuning database), else the batch shape heuristic")
    sfn = {"model": lambda: tune.schedule_fn(dev, threads=B.BATCH_THREADS), "heuristic": lambda: None}.get(
        args.schedules, lambda: tune.batch_schedule_fn(B.BATCH_THREADS))()
    br = B.BatchRunner(items, device=dev, schedule_fn=sfn)
    br.compile(workers=8)
    t_plan = time.perf_counter() - t0
    graph = br.capture()
    st = br.stats()

    def step():
        graph.replay()
    t = time_steps(step, args.steps, args.warmup)
    # the same work one graph at a time through the fused executor (no batching, no capture)
    rt = sp.Runtime(device=dev, fusion_enabled=True)
    seq = lambda: [rt.fused.run(g, rt, wanted=w) for g, w in items]
    t_seq = time_steps(seq, max(2, args.steps // 5), 1)
    vox = st["voxels"]
    # HBM bytes per step from the committed FETCH_SIZE / WRITE_SIZE passes of this workload, when
    # they were taken on the same batch (same algorithmic bytes)
    traffic = None
    tp = ROOT / "profiles" / "traffic_batched_96.json"
    if tp.exists():
        tj = json.loads(tp.read_text())
        if tj.get("algorithmic_bytes_per_step") == st["algorithmic_bytes"]:
            traffic = round(tj["bytes_per_step"])
    _line("batched persistent launch: Gvoxel/s over 96 uneven grids x 6 random graphs",
          round(vox / t / 1e9, 2), "Gvoxel/s", args.steps, args.warmup, t * 1e3,
          {"workload": "96 fields, axes U{32..384} seed 6; 6 random graphs (3-12 ops) seed 6; periodic",
           "exprs": exprs, "levels": st["levels"], "units": st["units"], "structures": st["structures"],
           "voxels": vox, "plan_compile_s": round(t_plan, 2),
           "schedule_source": src,
           "schedules": br.schedules(),
           "sequential_fused_ms": round(t_seq * 1e3, 3), "speedup_vs_sequential": round(t_seq / t, 2)},
          roofline={"bound": "hbm", "achieved": round(st["algorithmic_bytes"] / t / 1e9, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(st["algorithmic_bytes"] / t / 1e9 / HBM_PEAK_GBS, 4),
                    "traffic": traffic, "algorithmic_bytes_per_launch": st["algorithmic_bytes"]})
    return 0


def bench_tensor(args) -> int:
    """768^3: lower half seeded Fourier (seed 8), upper half i^2 + j^2 + k^2; hessian ->
    eigenvalues, determinant, trace, deviatoric, mean curvature, fused."""
    import torch
    import stenpy_amd as sp
    from stenpy_amd import gen, tune
    dev = torch.device("cuda:0")
    n = 768 if args.n == 512 else args.n
    from stenpy_amd.workloads import TENSOR_EXPRS as exprs
    g, sinks, f = c5_graph(dev, n)
    graphs = [(g, sinks)]
    rt = sp.Runtime(device=dev, fusion_enabled=True, schedule_fn=tune.schedule_fn(dev))

    def step():
        for gg, w in graphs:
            rt.fused.run(gg, rt, wanted=w)
    t = time_steps(step, max(3, args.steps // 4), 1)
    nvox = n ** 3
    alg = nvox * 4 * (1 + 3 + 1 + 1 + 9 + 1)     # read f once, write 15 output channels (64 B)
    res = rt.fused.run(g, rt, wanted=sinks)
    cg, ln = group_launch(rt, g, res, sinks[0])
    # the group's kernels per evaluation (flag pass, main z-march, edge band), HIP events
    k_ms = ln.time_ms(reps=max(3, args.steps // 4), warmup=1, rounds=1)
    achieved = alg / (k_ms * 1e-3) / 1e9
    traffic = None
    tp = ROOT / "profiles" / "traffic_tensor_768.json"
    if tp.exists():
        tj = json.loads(tp.read_text())
        if tj.get("n") == n and tuple(tj.get("schedule", ())) == tuple(cg.gen.sched.as_tuple()):
            traffic = round(tj["bytes_per_launch"])
    parity = tensor_slab_parity(f, res, sinks) if not args.no_cpu else None
    _line("tensor ops: Gvoxel/s of hessian -> eig/det/trace/dev + mean curvature on 768^3",
          round(nvox / t / 1e9, 3), "Gvoxel/s", max(3, args.steps // 4), 1, t * 1e3,
          {"workload": f"{n}^3 FP32, lower half Fourier seed 8, upper half i^2+j^2+k^2; periodic; "
                       "one graph with five sinks sharing the hessian", "exprs": exprs,
           "groups": len(rt.fused._plans and next(iter(rt.fused._plans.values()))[2]),
           "schedule": list(cg.gen.sched.as_tuple()), "split_tail": cg.gen.split_tail,
           "staged_wide_stores": bool(cg.gen.stage)},
          roofline={"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "kernel_ms": round(k_ms, 4), "algorithmic_bytes_per_launch": alg,
                    "kernels": "fused_flag (+ fused_flag_tail), fused_kernel, fused_tail (side stream)",
                    "note": "VALU issue is the co-limit: counters give ~64% of the kernel's cycles "
                            "issuing VALU (profiles/tensor_768_fused_kernel_counters.md)"},
          parity=parity)
    return 0


def tensor_slab_parity(f, res, sinks, z0: int = 380, z1: int = 388, margin: int = 3):
    """The 2x error rule on planes [z0, z1) of the measured run's five outputs, across the seam
    between the Fourier half and i^2 + j^2 + k^2: the FP64 oracle and its FP32 twin evaluate the
    same graph on planes [z0 - margin, z1 + margin) of the input, whose interior planes equal the
    whole field's (the stencils reach 2 planes).  The oracle decides the determinant's global
    +1e-12 I from the slab alone; the shift it could miss is far below the FP32 errors."""
    import torch
    from oracle import graph_eval
    from stenpy_amd.compiler import compile_expressions
    from stenpy_amd.workloads import TENSOR_EXPRS
    torch.set_num_threads(cpu_threads())
    slab = f[z0 - margin:z1 + margin].cpu()
    g, s_, *_ = compile_expressions(TENSOR_EXPRS, boundary="periodic", field_map={"f": slab})
    o64 = graph_eval.evaluate(g, torch.float64)
    o32 = graph_eval.evaluate(g, torch.float32)
    out = {"planes": [z0, z1], "rule": "e_build <= 2 e_ref32 against the FP64 oracle"}
    for expr, nb, no in zip(TENSOR_EXPRS, sinks, s_):
        b = res[nb][z0:z1].detach().double().cpu()
        ref = o64[no][margin:margin + (z1 - z0)].double()
        e_b = (b - ref).abs().max().item()
        e_32 = (o32[no][margin:margin + (z1 - z0)].double() - ref).abs().max().item()
        out[expr] = {"e_build": e_b, "e_ref32": e_32, "rule_2x": bool(e_b <= 2.0 * e_32)}
    return out


def bench_spectral(args) -> int:
    """512^3 FP32 field (seed 2): spectral Laplacian, Re(ifftn(-|k|^2 fftn(f))), with the
    transforms in FP64 (rocFFT), against torch.fft on the same GPU in FP32 and FP64."""
    import torch
    from stenpy_amd import gen, ops
    dev = torch.device("cuda:0")
    n = args.n
    f = gen.generate((n, n, n), 2, dev)
    steps = max(3, args.steps // 4)
    t = time_steps(lambda: ops.spectral_laplacian(f), steps, 1)
    k = [torch.fft.fftfreq(n, d=1.0, device=dev) * (2 * math.pi) for _ in range(3)]
    k2 = (k[0].view(-1, 1, 1) ** 2 + k[1].view(1, -1, 1) ** 2 + k[2].view(1, 1, -1) ** 2)

    def torch_lap(x, kk):
        return torch.fft.ifftn(-kk * torch.fft.fftn(x)).real
    t32 = time_steps(lambda: torch_lap(f, k2), steps, 1)
    f64, k64 = f.double(), k2.double()
    t64 = time_steps(lambda: torch_lap(f64, k64), steps, 1)
    del f64, k64
    nvox = n ** 3
    alg = 8 * nvox        # read f, write the result
    _line("spectral Laplacian: Gvoxel/s on a 512^3 field (FP64 transforms, FP32 storage)",
          round(nvox / t / 1e9, 3), "Gvoxel/s", steps, 1, t * 1e3,
          {"workload": f"{n}^3 FP32 Fourier field seed 2, dx = 1",
           "torch_fft_fp32_ms": round(t32 * 1e3, 3), "torch_fft_fp64_ms": round(t64 * 1e3, 3)},
          roofline={"bound": "hbm", "achieved": round(alg / t / 1e9, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(alg / t / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                    "algorithmic_bytes_per_launch": alg})
    return 0


def bench_plain(args) -> int:
    """128^3 FP32 field (seed 1): the five single expressions of the plain eager workload, each
    as one plain HIP kernel per node (the reference's eager executor, re-hosted) and fused, with
    the FP64 CPU oracle's time and the error rule per expression."""
    import torch
    import stenpy_amd as sp
    from stenpy_amd import workloads as W
    from oracle import graph_eval
    dev = torch.device("cuda:0")
    from stenpy_amd import tune
    rt = sp.Runtime(device=dev, fusion_enabled=True, schedule_fn=tune.schedule_fn(dev))
    threads = cpu_threads()
    torch.set_num_threads(threads)
    steps = max(10, args.steps)
    rows, t_fused, t_graph, t_eager, t_cpu = {}, 0.0, 0.0, 0.0, 0.0
    for expr, g, wanted in W.plain_graphs(dev):
        sink = wanted[0]
        te = time_steps(lambda: rt.run_eager(g), steps, args.warmup)
        tf = time_steps(lambda: rt.fused.run(g, rt, wanted=wanted), steps, args.warmup)
        cap, _ = rt.fused.capture(g, rt, wanted)
        tg = time_steps(cap.replay, steps, args.warmup)
        del cap
        res = rt.fused.run(g, rt, wanted=wanted)
        torch.cuda.synchronize()
        field = [g.node(n).params["value"] for n in g._order if g.node(n).op_name == "_constant"
                 and isinstance(g.node(n).params["value"], torch.Tensor) and g.node(n).params["value"].ndim == 3][0]
        gc = sp.compile_expression(expr, dx=1.0, boundary="periodic", field_map={"f": field.cpu()})
        g64, s64 = gc[0], gc[1]
        t0 = time.perf_counter()
        o64 = graph_eval.evaluate(g64, torch.float64)[s64]
        tc = time.perf_counter() - t0
        o32 = graph_eval.evaluate(g64, torch.float32)[s64]
        b = res[sink].detach().double().cpu()
        e_b = (b - o64.double()).abs().max().item()
        e_32 = (o32.double() - o64.double()).abs().max().item()
        rows[expr] = {"fused_ms": round(tf * 1e3, 4), "fused_graph_replay_ms": round(tg * 1e3, 4),
                      "eager_ms": round(te * 1e3, 4),
                      "cpu_oracle_ms": round(tc * 1e3, 2), "e_build": e_b, "e_ref32": e_32,
                      "rule_2x": bool(e_b <= 2.0 * e_32)}
        t_fused, t_graph, t_eager, t_cpu = t_fused + tf, t_graph + tg, t_eager + te, t_cpu + tc
    vox = 5 * 128 ** 3
    _line("plain eager workload: Gvoxel/s over the five 128^3 expressions, fused",
          round(vox / t_fused / 1e9, 3), "Gvoxel/s", steps, args.warmup, t_fused * 1e3,
          {"workload": "128^3 FP32 Fourier field seed 1 (+ noise), periodic, dx = 1; "
                       "f**2 + tanh(f)*exp(f), laplacian(f), gradient(f), mean(f), norm_l2(f), one at a time",
           "expressions": rows, "eager_gvoxel_s": round(vox / t_eager / 1e9, 3),
           "fused_graph_replay_gvoxel_s": round(vox / t_graph / 1e9, 3),
           "note": "a latency workload: the 8 MiB field stays in the 256 MB last-level cache"},
          cpu_baseline={"value": round(vox / t_cpu / 1e9, 5), "unit": "Gvoxel/s", "cores": threads,
                        "kind": "port", "sample": "the five expressions on the whole 128^3 field, FP64 torch CPU oracle",
                        "cpu": cpu_model(),
                        "note": "lower bound on the reference's CPU time (no executor gc / zero fill)"})
    return 0


def pinned_h2d_gbs(dev, nbytes: int = 1 << 30, reps: int = 5) -> float:
    """Host-to-device GB/s of one pinned 1 GiB buffer on this box (HIP events)."""
    import torch
    h = torch.empty(nbytes // 4, dtype=torch.float32, pin_memory=True)
    d = torch.empty(nbytes // 4, dtype=torch.float32, device=dev)
    d.copy_(h, non_blocking=True)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        d.copy_(h, non_blocking=True)
    e1.record()
    e1.synchronize()
    return nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9


def bench_stream(args) -> int:
    """4608 x 4096 x 4096 field generated on the host slab by slab (256 planes, 1-plane halo),
    streamed through pinned buffers; every slab by default (``--slabs N`` times the first N)."""
    import torch
    import stenpy_amd as sp
    from stenpy_amd import tune
    from stenpy_amd.stream import GeneratorField, stream_graph
    from stenpy_amd.compiler import compile_expression
    dev = torch.device("cuda:0")
    shape = (4608, 4096, 4096)
    lazy = GeneratorField(shape, 7)
    probe = torch.zeros((1,) + shape[1:])
    g, sink, _, fm, _ = compile_expression("f**2 + tanh(f)*exp(f) + laplacian(f)", boundary="periodic",
                                           field_map={"f": probe})
    use_model = args.schedules != "heuristic"
    rt = sp.Runtime(device=dev, fusion_enabled=True, schedule_fn=tune.schedule_fn(dev) if use_model else None)
    h2d = pinned_h2d_gbs(dev)
    # untimed warm-up: a pass over the first two slabs plans and compiles the slab graph,
    # captures its two slab shapes and leaves the pinned staging buffers in torch's host
    # allocator cache, as a stream running in a service would find them
    warm = 1 if args.warmup > 0 else 0
    if warm:
        stream_graph(g, sink, fm["f"], lazy, None, rt, boundary="periodic", slab_planes=256, max_slabs=2)
    st = stream_graph(g, sink, fm["f"], lazy, None, rt, boundary="periodic", slab_planes=256,
                      max_slabs=args.slabs or None)
    vox = st["voxels"]
    _line("out-of-core stream: Gvoxel/s (host generation + H2D + fused compute)",
          round(st["gvoxel_s"], 3), "Gvoxel/s", 1, warm, st["wall_s"] * 1e3,
          {"workload": "4608x4096x4096 FP32 seed 7, 256-plane slabs, periodic halo 1; "
                       f"{st['n_chunks']} of 18 slabs", "slabs": st["n_chunks"],
           "warmup_pass": "one untimed pass over 2 slabs (plans, code objects, slab-graph capture, "
                          "pinned staging buffers)" if warm else None,
           "host_gen_s": round(st["host_gen_s"], 2), "l2": st["l2"], "mean": st["mean"],
           "min": st["min"], "max": st["max"], "captured_shapes": st["captured_shapes"],
           "schedules": sorted({tuple(cg.gen.sched.as_tuple()) for plan in rt.fused._plans.values()
                                for cg in plan[2].values()}),
           "schedule_source": "tuning database, then the learned model" if use_model else "shape heuristic",
           "h2d_gb": round(st["h2d_bytes"] / 1e9, 2),
           "h2d_gbs_end_to_end": round(st["h2d_gbs_end_to_end"], 2),
           "pinned_h2d_gbs_probe": round(h2d, 2),
           "frac_of_pinned_h2d": round(st["h2d_gbs_end_to_end"] / h2d, 4)},
          roofline={"bound": "pcie", "achieved": round(st["h2d_gbs_end_to_end"], 2),
                    "peak": round(h2d, 2), "unit": "GB/s",
                    "frac": round(st["h2d_gbs_end_to_end"] / h2d, 4), "traffic": None})
    return 0


def copy_bytes(n: int) -> int:
    return 4 * n ** 3


# the profile pass's streaming copy is the only unary point-wise kernel it launches (4-wide
# k_unary_vec4 on aligned input; older traces name the one-element k_unary_contig)
CALIBRATION_COPY = "k_unary_"


def profile_pass(args) -> int:
    """Only the measured kernels (plus one calibration copy): what rocprofv3 measures.
    graph: the 512^3 stencil graph's fused group; batched: every level launch of the 96-grid
    batch; tensor: the 768^3 five-sink group (determinant flag pass and main launch)."""
    import torch
    import stenpy_amd as sp
    from stenpy_amd import _lib as L, gen, tune
    dev = torch.device("cuda:0")
    n = args.n
    info = {"profile_pass": True, "workload": args.workload, "launches": args.steps,
            "copy_bytes": copy_bytes(n)}
    if args.workload == "batched":
        from stenpy_amd import batch as B
        _, items = c3_items(dev)
        br = B.BatchRunner(items, device=dev)
        br.compile(workers=8)
        for _ in range(args.steps):
            br.run()
        info["stats"] = {k: v for k, v in br.stats().items() if isinstance(v, (int, float))}
    elif args.workload == "tensor":
        g, sinks, _ = c5_graph(dev, 768)
        rt = sp.Runtime(device=dev, fusion_enabled=True, schedule_fn=tune.schedule_fn(dev))
        res = rt.fused.run(g, rt, wanted=sinks)
        cg, ln = group_launch(rt, g, res, sinks[0])
        for _ in range(args.steps):
            ln.run_once()
        from stenpy_amd import fused as F
        info.update(schedule=list(cg.gen.sched.as_tuple()), grid=ln.nblocks + cg.gen.boundary_blocks(),
                    split_tail=cg.gen.split_tail, staged=bool(cg.gen.stage),
                    staged_in_source="so_stage" in cg.source,
                    vgprs=F.kernel_registers(F.compile_source(cg.source)[0]))
    else:
        f = gen.generate((n, n, n), 2, dev)
        u = gen.generate_vector((n, n, n), (3, 4, 5), dev)
        g, sink, reds = build_graph(f, u)
        rt = sp.Runtime(device=dev, fusion_enabled=True, schedule_fn=tune.schedule_fn(dev))
        res = rt.fused.run(g, rt, wanted=[sink] + reds)
        cg, ln = group_launch(rt, g, res, sink)
        for _ in range(args.steps):
            ln.launch()
        info.update(schedule=list(cg.gen.sched.as_tuple()), grid=ln.nblocks + cg.gen.boundary_blocks())
    # a kernel of known arithmetic for the flop counters: 2^18 threads, 12 FP32 adds each
    pf = torch.ones(1 << 20, device=dev)
    L.probe_stream("4r1w", False, a=pf, u=torch.ones(3 << 20, device=dev), out=torch.empty_like(pf))
    c = gen.generate((n, n, n), 2, dev)
    out = torch.empty_like(c)
    for _ in range(3):
        L.unary("neg", c, out)          # streaming copy: reads and writes copy_bytes(n)
    torch.cuda.synchronize()
    print(json.dumps(info), flush=True)
    return 0


def offline_reports(args) -> int:
    from stenpy_amd import profiling
    if args.pmc_traffic:
        t = profiling.traffic(args.pmc_traffic[0], args.pmc_traffic[1], kernel=args.kernel,
                              copy_kernel=CALIBRATION_COPY, copy_bytes=copy_bytes(args.n))
        try:        # the edge-band tail kernel of the same group, when it runs separately
            tt = profiling.traffic(args.pmc_traffic[0], args.pmc_traffic[1], kernel="fused_tail",
                                   copy_kernel=CALIBRATION_COPY, copy_bytes=copy_bytes(args.n))
            t["tail_bytes"] = tt["bytes_per_launch"]
            t["main_bytes"] = t["bytes_per_launch"]
            t["bytes_per_launch"] = t["main_bytes"] + t["tail_bytes"]
        except KeyError:
            pass
        t["kernel"] = args.kernel
        t["algorithmic_bytes"] = args.alg_bytes or 20 * args.n ** 3
        t["n"] = args.grid_n or args.n
        # the schedule the profile pass ran (its own JSON line in the fetch pass's log)
        for log in sorted(Path(args.pmc_traffic[0]).parent.glob(Path(args.pmc_traffic[0]).name + "*.log")):
            for ln_ in log.read_text().splitlines():
                if ln_.startswith("{") and "profile_pass" in ln_:
                    t["schedule"] = json.loads(ln_)["schedule"]
        t["ratio_to_algorithmic"] = t["bytes_per_launch"] / t["algorithmic_bytes"]
        text = json.dumps(t, indent=1) + "\n"
    else:
        from stenpy_amd import profiling as P
        text = P.stats_markdown(args.stats_md, args.cmd)
        if args.tail:
            tt = P.trace_tail(args.stats_md, "fused_kernel", args.tail)
            text += (f"\nThe last {tt['tail']} of {tt['dispatches']} `fused_kernel` dispatches (the roofline's "
                     f"timed FP32 launches): mean {tt['mean_ns'] / 1e3:.2f} us, min {tt['min_ns'] / 1e3:.2f}, "
                     f"max {tt['max_ns'] / 1e3:.2f}.\n")
    if args.out:
        Path(args.out).write_text(text)
    print(text)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--traffic", type=str, default=str(ROOT / "profiles" / "traffic_stencil_graph_512.json"),
                    help="JSON file with measured HBM bytes per launch of the fused kernel")
    ap.add_argument("--profile-pass", action="store_true",
                    help="only launch the fused stencil-graph kernels --steps times plus a streaming copy of known "
                         "size (the target of rocprofv3 --pmc / --kernel-trace runs)")
    ap.add_argument("--pmc-traffic", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"), default=None,
                    help="turn two rocprofv3 --pmc CSV outputs into the traffic JSON (no GPU)")
    ap.add_argument("--stats-md", type=str, default=None, metavar="DIR",
                    help="render a rocprofv3 --stats CSV output as markdown (no GPU)")
    ap.add_argument("--cmd", type=str, default="", help="command line quoted in --stats-md output")
    ap.add_argument("--kernel", type=str, default="fused_kernel",
                    help="with --pmc-traffic: the kernel whose bytes per launch are reported")
    ap.add_argument("--grid-n", type=int, default=0,
                    help="with --pmc-traffic: edge of the profiled workload's grid (default --n, "
                         "which sizes the calibration copy)")
    ap.add_argument("--alg-bytes", type=int, default=0,
                    help="with --pmc-traffic: algorithmic bytes per launch (default 20 n^3)")
    ap.add_argument("--tail", type=int, default=0,
                    help="with --stats-md: also report the mean of the last N fused_kernel dispatches")
    ap.add_argument("--out", type=str, default=None)
    ap.add_argument("--workload", choices=["graph", "plain", "batched", "stream", "tensor", "spectral"], default="graph",
                    help="graph: the 512^3 stencil graph (default); plain: the five 128^3 "
                         "expressions, plain and fused; batched: 96 uneven grids x 6 "
                         "random graphs in batched persistent launches; stream: out-of-core "
                         "4608x4096x4096 stream (bounded sample of slabs); tensor: 768^3 hessian "
                         "tensor ops; spectral: 512^3 spectral Laplacian")
    ap.add_argument("--schedules", choices=["auto", "heuristic", "model"], default="auto",
                    help="batched / stream schedule source; auto: the batch autotuner's database entries "
                         "(else the batch heuristic) for batched, the tuning database then the learned "
                         "model for stream")
    ap.add_argument("--slabs", type=int, default=0, help="stream: slabs of 256 planes to stream (0: all 18)")
    args = ap.parse_args()
    if args.pmc_traffic or args.stats_md:
        return offline_reports(args)
    if args.gpus != 1:
        print(json.dumps({"error": "this backend runs on one GPU; --gpus must be 1"}))
        return 2
    if args.profile_pass:
        return profile_pass(args)
    if args.workload != "graph":
        return {"plain": bench_plain, "batched": bench_batched, "stream": bench_stream,
                "tensor": bench_tensor, "spectral": bench_spectral}[args.workload](args)

    import torch
    import stenpy_amd as sp
    from stenpy_amd import gen, tune

    dev = torch.device("cuda:0")
    n = args.n
    shape = (n, n, n)
    f = gen.generate(shape, 2, dev)
    u = gen.generate_vector(shape, (3, 4, 5), dev)
    del_ = None
    g, sink, reds = build_graph(f, u)
    wanted = [sink] + reds
    nvox = n ** 3

    rt = sp.Runtime(device=dev, fusion_enabled=True, schedule_fn=tune.schedule_fn(dev))
    res = {}

    def fused_fp32():
        res.update(rt.fused.run(g, rt, wanted=wanted))

    def fused_bf16():
        rt.fused.run(g, rt, wanted=wanted, out_dtype=torch.bfloat16)

    def plain():
        rt.run_eager(g, outputs=wanted)

    # the same evaluation replayed from a hipGraph (no per-step host work); on this runtime a
    # replay costs more per step than launching the two kernels directly, so the step timed
    # for the headline is the direct launch, measured last (plans compiled, clocks settled)
    cap, _ = rt.fused.capture(g, rt, wanted)
    cap16, _ = rt.fused.capture(g, rt, wanted, out_dtype=torch.bfloat16)
    t_replay = time_steps(cap.replay, args.steps, args.warmup)
    t_bf16_replay = time_steps(cap16.replay, args.steps, args.warmup)
    t_plain = time_steps(plain, max(3, args.steps // 4), 1)
    t_bf16 = time_steps(fused_bf16, args.steps, args.warmup)
    t_fused = time_steps(fused_fp32, args.steps, args.warmup)

    cg, ln = group_launch(rt, g, res, sink)
    # average duration of the group's kernels over the timed launches (HIP events on their
    # stream), the same launch sequence the kernel-statistics profile records
    k_ms = ln.time_ms(reps=args.steps, warmup=args.warmup, rounds=1)
    ceil = hbm_ceilings(f, u)
    alg_bytes = 20 * nvox
    achieved = alg_bytes / (k_ms * 1e-3) / 1e9
    traffic = None
    if args.traffic and Path(args.traffic).exists():
        tj = json.loads(Path(args.traffic).read_text())
        if tj.get("n") == n and tuple(tj.get("schedule", ())) in ((), tuple(cg.gen.sched.as_tuple())):
            traffic = round(tj.get("bytes_per_launch"))

    line = {
        "metric": "Gvoxel/s + % HBM peak: fused/autotuned vs plain per-op HIP, 512^3 graph",
        "value": round(nvox / t_fused / 1e9, 2),
        "unit": "Gvoxel/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t_fused * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded 8-mode Fourier fields + counter-based noise)",
        "config": {"workload": f"stencil graph lap(f) + u.grad(f) + |curl u|^2, {n}^3 FP32 f + {n}^3x3 u, periodic, dx=1; "
                               "out + sum/mean/variance/L2",
                   "schedule": list(cg.gen.sched.as_tuple()),
                   "hbm_frac_step": round(20 * nvox / t_fused / 1e9 / HBM_PEAK_GBS, 4),
                   "plain_gvoxel_s": round(nvox / t_plain / 1e9, 3),
                   "plain_ms_per_step": round(t_plain * 1e3, 3),
                   "speedup_vs_plain": round(t_plain / t_fused, 2),
                   "bf16_out_gvoxel_s": round(nvox / t_bf16 / 1e9, 2),
                   "schedule_source": schedule_source(cg, g),
                   "step": "the runtime's fused evaluation, launched directly: fused_kernel, then one "
                           "launch of the 4 FP64 final combines; output buffers from torch's caching allocator",
                   "graph_replay_gvoxel_s": round(nvox / t_replay / 1e9, 2),
                   "graph_replay_ms_per_step": round(t_replay * 1e3, 4),
                   "bf16_graph_replay_gvoxel_s": round(nvox / t_bf16_replay / 1e9, 2),
                   "probe_gbs": {"copy_1r1w": ceil["copy"], "stencil_mix_4r1w": ceil["4r1w"],
                                 "stencil_mix_4r1w_channel_last": ceil["4r1w_aos"],
                                 "read_only": ceil["read"], "write_only": ceil["write"]},
                   "bf16_out_hbm_frac_step": round(18 * nvox / t_bf16 / 1e9 / HBM_PEAK_GBS, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel_ms": round(k_ms, 4),
                     "frac_of_probe_4r1w": round(achieved / ceil["4r1w"], 4),
                     "frac_of_probe_4r1w_channel_last": round(achieved / ceil["4r1w_aos"], 4),
                     "kernels": ("fused_kernel (LDS z-march, face wave) + fused_tail (edge band, side stream)"
                                 if cg.gen.split_tail else
                                 "fused_kernel (LDS z-march, face wave, edge-band blocks at the end of the grid)"),
                     "algorithmic_bytes_per_launch": alg_bytes},
    }
    if not args.no_cpu:
        res = rt.fused.run(g, rt, wanted=wanted)
        torch.cuda.synchronize()
        base, errs = oracle_leg(f, u, res, sink, reds)
        line["cpu_baseline"] = base
        line["parity"] = {"rule": "e_build <= 2 e_ref32 against the FP64 oracle, whole field",
                          **errs}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

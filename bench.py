"""Benchmark: config-2 graph on a 512^3 grid, fused/autotuned vs plain per-operator HIP.

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
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

C2_EXPR = "divergence(gradient(f)) + dot(u, gradient(f)) + sqnorm(curl(u))"
REDUCTIONS = ("sum", "mean", "variance", "norm_l2")
HBM_PEAK_GBS = 8000.0


def build_graph(f, u):
    import stenpy_amd as sp
    g, sink, _, _, _ = sp.compile_expression(C2_EXPR, dx=1.0, boundary="periodic",
                                             field_map={"f": f, "u": u})
    reds = [g.add(r, (sink,), {}) for r in REDUCTIONS]
    return g, sink, reds


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


def cpu_baseline(n: int, planes: int):
    """FP64 CPU oracle on a (planes, n, n) slab of the same graph; voxels/s."""
    import torch
    from stenpy_amd import gen
    from oracle import graph_eval
    threads = os.cpu_count() or 1
    torch.set_num_threads(threads)
    shape = (planes, n, n)
    f = torch.from_numpy(gen.generate_host(shape, 2)).double()
    u = torch.stack([torch.from_numpy(gen.generate_host(shape, s)).double() for s in (3, 4, 5)], -1)
    g, sink, reds = build_graph(f, u)
    graph_eval.evaluate(g, torch.float64)           # warm-up
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        graph_eval.evaluate(g, torch.float64)
        ts.append(time.perf_counter() - t0)
    ts.sort()
    vox = planes * n * n
    return vox / ts[1], threads, f"{planes}x{n}x{n} slab, FP64 torch CPU oracle, median of 3"


def copy_bytes(n: int) -> int:
    return 4 * n ** 3


def profile_pass(args) -> int:
    """Fused C2 launches only (plus one calibration copy): what rocprofv3 measures."""
    import torch
    import stenpy_amd as sp
    from stenpy_amd import _lib as L, gen, tune
    dev = torch.device("cuda:0")
    n = args.n
    f = gen.generate((n, n, n), 2, dev)
    u = gen.generate_vector((n, n, n), (3, 4, 5), dev)
    g, sink, reds = build_graph(f, u)
    rt = sp.Runtime(device=dev, fusion_enabled=True, schedule_fn=tune.schedule_fn(dev))
    res = rt.fused.run(g, rt, wanted=[sink] + reds)
    cg, ln = group_launch(rt, g, res, sink)
    for _ in range(args.steps):
        ln.launch()
    out = torch.empty_like(f)
    for _ in range(3):
        L.unary("neg", f, out)          # streaming copy: reads and writes copy_bytes(n)
    torch.cuda.synchronize()
    print(json.dumps({"profile_pass": True, "schedule": list(cg.gen.sched.as_tuple()),
                      "launches": args.steps, "grid": ln.nblocks + cg.gen.boundary_blocks(),
                      "copy_bytes": copy_bytes(n)}), flush=True)
    return 0


def offline_reports(args) -> int:
    from stenpy_amd import profiling
    if args.pmc_traffic:
        t = profiling.traffic(args.pmc_traffic[0], args.pmc_traffic[1], kernel="fused_kernel",
                              copy_kernel="k_unary_contig", copy_bytes=copy_bytes(args.n))
        t["algorithmic_bytes"] = 20 * args.n ** 3
        t["n"] = args.n
        text = json.dumps(t, indent=1) + "\n"
    else:
        from stenpy_amd import profiling as P
        text = P.stats_markdown(args.stats_md, args.cmd)
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
    ap.add_argument("--cpu-planes", type=int, default=32)
    ap.add_argument("--traffic", type=str, default=str(ROOT / "profiles" / "traffic_c2_fused_512.json"),
                    help="JSON file with measured HBM bytes per launch of the fused kernel")
    ap.add_argument("--profile-pass", action="store_true",
                    help="only launch the fused C2 kernel --steps times plus a streaming copy of known "
                         "size (the target of rocprofv3 --pmc / --kernel-trace runs)")
    ap.add_argument("--pmc-traffic", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"), default=None,
                    help="turn two rocprofv3 --pmc CSV outputs into the traffic JSON (no GPU)")
    ap.add_argument("--stats-md", type=str, default=None, metavar="DIR",
                    help="render a rocprofv3 --stats CSV output as markdown (no GPU)")
    ap.add_argument("--cmd", type=str, default="", help="command line quoted in --stats-md output")
    ap.add_argument("--out", type=str, default=None)
    args = ap.parse_args()
    if args.pmc_traffic or args.stats_md:
        return offline_reports(args)
    if args.gpus != 1:
        print(json.dumps({"error": "this backend runs on one GPU; --gpus must be 1"}))
        return 2
    if args.profile_pass:
        return profile_pass(args)

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

    t_fused = time_steps(fused_fp32, args.steps, args.warmup)
    t_bf16 = time_steps(fused_bf16, args.steps, args.warmup)
    t_plain = time_steps(plain, max(3, args.steps // 4), 1)

    cg, ln = group_launch(rt, g, res, sink)
    k_ms = ln.time_ms(reps=max(10, args.steps))
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
        "config": {"workload": f"config-2 graph, {n}^3 FP32 f + {n}^3x3 u, periodic, dx=1; "
                               "out + sum/mean/variance/L2",
                   "schedule": list(cg.gen.sched.as_tuple()),
                   "hbm_frac_step": round(20 * nvox / t_fused / 1e9 / HBM_PEAK_GBS, 4),
                   "plain_gvoxel_s": round(nvox / t_plain / 1e9, 3),
                   "plain_ms_per_step": round(t_plain * 1e3, 3),
                   "speedup_vs_plain": round(t_plain / t_fused, 2),
                   "bf16_out_gvoxel_s": round(nvox / t_bf16 / 1e9, 2),
                   "bf16_out_hbm_frac_step": round(18 * nvox / t_bf16 / 1e9 / HBM_PEAK_GBS, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel_ms": round(k_ms, 4),
                     "kernels": "fused_kernel (LDS z-march over full tiles + tail blocks for the face band)",
                     "algorithmic_bytes_per_launch": alg_bytes},
    }
    if not args.no_cpu:
        rate, cores, sample = cpu_baseline(n, args.cpu_planes)
        line["cpu_baseline"] = {"value": round(rate / 1e9, 5), "unit": "Gvoxel/s", "cores": cores,
                                "kind": "port", "sample": sample}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

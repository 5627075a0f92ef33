"""Out-of-core evaluation: axis-0 slabs streamed host -> device through pinned buffers.

Reference counterparts: ``Runtime.run_chunked`` (stenpy_engine.py:697-888), halo reads
``_read_halo_chunk`` (647-687) and crop (689-695), ``Pipeline._run_streamed_graph``
(stenpy.py:762-884) and ``stream_execute`` (474-615).  Behaviour kept: slabs along axis 0,
each padded with halo planes read with periodic wrap-around (or truncated at the ends
otherwise), the graph evaluated on the padded slab and cropped, min/max/mean of the sink
reported.  Design:

* the halo is the graph's *composite* stencil radius ``r`` (the reference uses the largest
  single-op radius, wrong for chained stencils);
* only planes that are not already on the device cross PCIe: the leading ``2r`` planes of a
  slab are the trailing planes of the previous slab's device buffer and are copied device to
  device, and the periodic wrap at the end reuses a device copy of the first ``r`` planes.
  A producer thread fills ``n_buffers`` pinned staging chunks (generator or file); each is
  copied on a side stream into one of two device slab buffers, so host generation and PCIe
  copies of the next slab overlap the compute of the current one;
* the sink's L2 norm, sum and min/max are reduced inside the slab's fused kernel over the
  interior planes only (a reduction z-window), into device FP64 results; nothing is read
  back until the last slab has run, so the host never waits for a slab;
* slabs of a repeated shape replay a hipGraph of the fused evaluation captured once per
  device buffer;
* a producer or writer exception is re-raised in the caller (the reference's reader thread
  can deadlock the pipeline on an exception, stenpy_engine.py:773-816).
"""
from __future__ import annotations

import math
import queue
import threading
import time
import warnings
from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib as L
from . import gen
from .graph import Graph
from .registry import OP_METADATA
from .runtime import LazyFieldProtocol, Runtime


class GeneratorField(LazyFieldProtocol):
    """The seeded Fourier field, produced on the host plane range by plane range."""

    def __init__(self, shape: Sequence[int], seed: int, threads: int = 0):
        self.shape = tuple(int(s) for s in shape)
        self.seed = int(seed)
        self.threads = threads
        self._modes = gen.fourier_params(self.seed)

    @property
    def dtype(self):
        return np.float32

    @property
    def nbytes(self) -> int:
        return int(np.prod(self.shape)) * 4

    def read_into(self, z0: int, z1: int, out: np.ndarray) -> None:
        """Planes [z0, z1) (0 <= z0 <= z1 <= n0) into the C-contiguous float32 array ``out``."""
        assert out.flags.c_contiguous and out.dtype == np.float32
        ks, a, ph = self._modes
        L.host_gen_fourier(out.ctypes.data, self.shape, z0, z1 - z0, ks, a, ph, self.seed, self.threads)

    def __getitem__(self, idx) -> np.ndarray:
        sl = idx[0] if isinstance(idx, tuple) else idx
        if not isinstance(sl, slice):
            sl = slice(sl, sl + 1)
        z0, z1, _ = sl.indices(self.shape[0])
        out = np.empty((z1 - z0,) + self.shape[1:], dtype=np.float32)
        self.read_into(z0, z1, out)
        return out if not isinstance(idx, tuple) else out[(slice(None),) + tuple(idx[1:])]


class NpyLazyField(LazyFieldProtocol):
    """A ``.npy`` file read plane range by plane range through a memory map."""

    def __init__(self, path: str):
        self.path = path
        self._mm = np.load(path, mmap_mode="r")
        self.shape = tuple(self._mm.shape)

    @property
    def dtype(self):
        return np.float32

    def read_into(self, z0: int, z1: int, out: np.ndarray) -> None:
        out[...] = self._mm[z0:z1]

    def __getitem__(self, idx) -> np.ndarray:
        return np.asarray(self._mm[idx], dtype=np.float32)


def composite_radius(graph: Graph) -> int:
    """Largest sum of stencil radii along any path (the halo a slab needs)."""
    memo: Dict[str, int] = {}
    for node in graph.topological_sort():
        r_in = max([memo[i] for i in node.input_ids] + [0])
        memo[node.id] = r_in + int(OP_METADATA.get(node.op_name, {}).get("stencil_radius", 0) or 0)
    return max(memo.values()) if memo else 0


def halo_ranges(start: int, end: int, total: int, radius: int, boundary: str):
    """Plane pieces [(a, b), ...] that make up the padded slab, plus (lpad, rpad)."""
    if radius <= 0:
        return [(start, end)], 0, 0
    if boundary != "periodic":
        lo, hi = max(0, start - radius), min(total, end + radius)
        return [(lo, hi)], start - lo, hi - end
    pieces = []
    if start - radius >= 0:
        pieces.append((start - radius, start))
    else:
        pieces.append((total - (radius - start), total))
        if start > 0:
            pieces.append((0, start))
    pieces.append((start, end))
    if end + radius <= total:
        pieces.append((end, end + radius))
    else:
        if end < total:
            pieces.append((end, total))
        pieces.append((0, end + radius - total))
    return pieces, radius, radius


# ------------------------------------------------------------------------------ slab planning
class SlabPlan:
    """Where every plane of one padded slab comes from.

    ``lo, hi`` are the padded slab's planes in unwrapped coordinates (periodic slabs may run
    below 0 or past n0); ``moves`` lists (dst, kind, a, b): ``kind`` "host" reads global planes
    [a, b) from the lazy field, "prev" copies planes [a, b) of the previous slab's device
    buffer, "wrap" copies planes [a, b) of the saved first planes of the field."""

    def __init__(self, k: int, s: int, e: int, lo: int, hi: int, moves: List[Tuple[int, str, int, int]]):
        self.k, self.s, self.e, self.lo, self.hi, self.moves = k, s, e, lo, hi, moves

    @property
    def planes(self) -> int:
        return self.hi - self.lo

    @property
    def lpad(self) -> int:
        return self.s - self.lo

    def host_moves(self) -> List[Tuple[int, int, int]]:
        return [(d, a, b) for d, kind, a, b in self.moves if kind == "host"]


def plan_slabs(n0: int, slab: int, r: int, boundary: str, max_slabs: Optional[int] = None,
               reuse: bool = True) -> List[SlabPlan]:
    """Slab k covers [k*slab, min((k+1)*slab, n0)) padded by ``r`` planes on each side: with
    periodic wrap-around (periodic) or truncated at the field's ends (any other boundary), as
    stenpy_engine.py:647-687 reads them.  With ``reuse``, planes that the previous slab already
    holds on the device, and the wrapped first planes, are taken from the device."""
    periodic = boundary == "periodic"
    plans: List[SlabPlan] = []
    starts = list(range(0, n0, slab))
    if max_slabs is not None:
        starts = starts[:max_slabs]
    wrap_ok = reuse and periodic and len(starts) > 1 and r <= min(slab, n0)
    for k, s in enumerate(starts):
        e = min(s + slab, n0)
        lo, hi = (s - r, e + r) if periodic else (max(0, s - r), min(n0, e + r))
        moves: List[Tuple[int, str, int, int]] = []
        u = lo
        if reuse and k > 0:
            prev = plans[-1]
            take = max(0, min(prev.hi, hi) - u)          # unwrapped planes the previous slab holds
            if take:
                moves.append((0, "prev", u - prev.lo, u - prev.lo + take))
                u += take
        while u < hi:
            g0 = u % n0 if periodic else u
            if periodic and u >= n0 and wrap_ok and k > 0 and u - n0 + (hi - u) <= r:
                moves.append((u - lo, "wrap", u - n0, u - n0 + (hi - u)))
                u = hi
                continue
            run = min(hi - u, n0 - g0)                   # contiguous global planes from g0
            moves.append((u - lo, "host", g0, g0 + run))
            u += run
        plans.append(SlabPlan(k, s, e, lo, hi, moves))
    return plans


class _Slot:
    """One pinned host staging buffer."""

    def __init__(self, planes: int, plane_shape: Tuple[int, ...]):
        self.host = torch.empty((planes,) + plane_shape, dtype=torch.float32, pin_memory=True)
        self.free = threading.Event()
        self.free.set()


def _release_after(ev: "torch.cuda.Event", slot: _Slot) -> None:
    ev.synchronize()
    slot.free.set()


def _put(q: "queue.Queue", item, stop: threading.Event) -> bool:
    """Bounded put that gives up once ``stop`` is set (the consumer has gone)."""
    while not stop.is_set():
        try:
            q.put(item, timeout=0.1)
            return True
        except queue.Full:
            continue
    return False


STREAM_REDUCTIONS = ("norm_l2", "sum", "min_max")


def stream_graph(graph: Graph, sink: str, src_id: str, lazy, out_path: Optional[str] = None,
                 rt: Optional[Runtime] = None, boundary: str = "periodic",
                 chunk_mb: Optional[float] = None, slab_planes: Optional[int] = None,
                 n_buffers: int = 3, max_slabs: Optional[int] = None,
                 stage_mb: float = 1024.0, capture: bool = True) -> Dict[str, Any]:
    """Evaluate ``graph`` over ``lazy`` (the source node ``src_id``) slab by slab.

    A padded slab (``slab_planes`` planes plus the composite-radius halo on each side) is
    assembled in one of two device buffers: planes the previous slab already holds are copied
    device to device, the rest arrive from ``n_buffers`` pinned staging chunks of at most
    ``stage_mb`` MiB that a producer thread fills.  The fused graph runs on the slab as soon
    as its last plane has landed, with the sink's L2 norm, sum and min/max reduced over the
    slab's interior inside the same kernel; results stay on the device until the end."""
    rt = rt or Runtime(device="cuda")
    dev = rt.device
    n0 = lazy.shape[0]
    plane_shape = tuple(lazy.shape[1:])
    plane_bytes = int(np.prod(plane_shape)) * 4
    r = composite_radius(graph)
    if slab_planes is None:
        budget = int((chunk_mb or 4096) * 2 ** 20)
        slab_planes = max(2 * r + 1, min(n0, budget // plane_bytes))
    slab_planes = min(max(1, slab_planes), n0)
    plans = plan_slabs(n0, slab_planes, r, boundary, max_slabs)
    cap = max(p.planes for p in plans)
    stage = max(1, min(cap, int(stage_mb * 2 ** 20) // plane_bytes))
    slots = [_Slot(stage, plane_shape) for _ in range(max(2, n_buffers))]
    dev_bufs = [torch.empty((cap,) + plane_shape, dtype=torch.float32, device=dev) for _ in range(2)]
    first_planes = None                   # device copy of global planes [0, r) (periodic wrap)
    buf_done: List[Optional[torch.cuda.Event]] = [None, None]
    copy_stream = torch.cuda.Stream(device=dev)
    ready: "queue.Queue" = queue.Queue(maxsize=len(slots))
    stop = threading.Event()
    t_gen = [0.0]
    h2d_bytes = [0]

    def producer():
        try:
            n = 0
            for pl in plans:
                chunks = []
                for dst, a, b in pl.host_moves():
                    for z in range(a, b, stage):
                        chunks.append((dst + (z - a), z, min(stage, b - z)))
                for ci, (dst, z, cnt) in enumerate(chunks):
                    slot = slots[n % len(slots)]
                    n += 1
                    while not slot.free.wait(0.1):
                        if stop.is_set():
                            return
                    slot.free.clear()
                    t0 = time.perf_counter()
                    lazy.read_into(z, z + cnt, slot.host[:cnt].numpy())
                    t_gen[0] += time.perf_counter() - t0
                    if not _put(ready, ("chunk", pl.k, slot, dst, cnt, ci == len(chunks) - 1), stop):
                        return
                if not chunks and not _put(ready, ("empty", pl.k), stop):
                    return
        except BaseException as exc:  # propagate to the consumer instead of hanging it
            _put(ready, ("error", exc), stop)

    # the slab graph: the source replaced by the device slab, plus the streamed reductions
    reds: Dict[str, str] = {}
    slab_graphs: Dict[Tuple[int, int], Tuple[Graph, Dict[str, str]]] = {}

    def slab_graph(b: int, planes: int):
        key = (b, planes)
        if key not in slab_graphs:
            g = graph.clone_with_replacement(src_id, dev_bufs[b][:planes])
            ids = {kind: g.add(kind, (sink,), {}) for kind in STREAM_REDUCTIONS}
            slab_graphs[key] = (g, ids)
        return slab_graphs[key]

    captured: Dict[Tuple[int, int, int], Any] = {}
    seen: Dict[Tuple[int, int, int], int] = {}
    per_slab: List[Dict[str, torch.Tensor]] = []
    writer = _Writer(out_path, n0, dev) if out_path is not None else None
    fused_ok = rt.fusion_enabled

    sink_ch = [1]
    use_capture = [bool(capture)]

    def evaluate(pl: SlabPlan, b: int) -> None:
        nonlocal fused_ok
        if writer is not None:
            # result buffers are reused by the next evaluation: wait for the last D2H copy
            torch.cuda.current_stream(dev).wait_stream(writer.stream)
        g, ids = slab_graph(b, pl.planes)
        win = (pl.lpad, pl.lpad + (pl.e - pl.s))
        wanted = [sink] + list(ids.values())
        res = None
        if fused_ok:
            key = (b, pl.planes, pl.lpad)
            try:
                if key not in captured and use_capture[0] and seen.get(key, 0) >= 1:
                    try:
                        captured[key] = _capture_window(rt, g, wanted, win)
                    except Exception as e:          # replay is an optimisation: run eagerly
                        warnings.warn(f"slab capture failed, evaluating slabs eagerly: {e}")
                        use_capture[0] = False
                if key in captured:
                    graph_, res = captured[key]
                    graph_.replay()
                else:
                    res = rt.fused.run(g, rt, wanted=wanted, red_window=win)
                seen[key] = seen.get(key, 0) + 1
            except Exception as e:  # same policy as Runtime.run: warn and use the plain path
                import os
                if os.environ.get("STENPY_FUSED_STRICT", "0").lower() in ("1", "true", "yes"):
                    raise
                warnings.warn(f"Fusion failed, falling back to original: {e}")
                fused_ok = False
                res = None
        if res is None:
            res = rt.run_eager(g, outputs=[sink])
            crop = res[sink][win[0]:win[1]]
            res = dict(res)
            for kind, nid in ids.items():
                res[nid] = L.reduce(kind, crop.contiguous())
        # copy the slab's results out of buffers that the next replay reuses
        per_slab.append({kind: res[nid].clone() for kind, nid in ids.items()})
        sink_ch[0] = int(np.prod(res[sink].shape[3:])) if res[sink].ndim > 3 else 1
        if writer is not None:
            writer.put(res[sink][win[0]:win[1]], pl.s)

    th = threading.Thread(target=producer, daemon=True)
    th.start()
    t_start = time.perf_counter()
    compute = torch.cuda.current_stream(dev)
    try:
        for pl in plans:
            b = pl.k % 2
            dbuf = dev_bufs[b]
            with torch.cuda.stream(copy_stream):
                if buf_done[b] is not None:
                    copy_stream.wait_event(buf_done[b])      # slab k-2 is done with this buffer
                for dst, kind, a, bb in pl.moves:
                    if kind == "prev":
                        dbuf[dst:dst + bb - a].copy_(dev_bufs[1 - b][a:bb], non_blocking=True)
                    elif kind == "wrap":
                        dbuf[dst:dst + bb - a].copy_(first_planes[a:bb], non_blocking=True)
            if pl.host_moves():
                while True:
                    msg = ready.get()
                    if msg[0] == "error":
                        raise RuntimeError(f"slab producer failed: {msg[1]}") from msg[1]
                    _, k, slot, dst, cnt, last = msg
                    assert k == pl.k
                    with torch.cuda.stream(copy_stream):
                        dbuf[dst:dst + cnt].copy_(slot.host[:cnt], non_blocking=True)
                        ev = torch.cuda.Event()
                        ev.record(copy_stream)
                    h2d_bytes[0] += cnt * plane_bytes
                    # the pinned chunk can be refilled once its copy has finished
                    threading.Thread(target=_release_after, args=(ev, slot), daemon=True).start()
                    if last:
                        break
            else:
                msg = ready.get()
                if msg[0] == "error":
                    raise RuntimeError(f"slab producer failed: {msg[1]}") from msg[1]
            if pl.k == 0 and boundary == "periodic" and len(plans) > 1 and r > 0:
                with torch.cuda.stream(copy_stream):
                    first_planes = dbuf[pl.lpad:pl.lpad + r].clone()
            landed = torch.cuda.Event()
            landed.record(copy_stream)
            compute.wait_event(landed)
            evaluate(pl, b)
            done = torch.cuda.Event()
            done.record(compute)
            buf_done[b] = done
            if writer is not None:
                writer.check()
    finally:
        stop.set()
        th.join(timeout=60)
        if writer is not None:
            writer.close()
    torch.cuda.synchronize(dev)
    wall = time.perf_counter() - t_start
    # FP64 combine across slabs, once, after the last slab
    l2sq = math.fsum(float(d["norm_l2"]) ** 2 for d in per_slab)
    total = math.fsum(float(d["sum"]) for d in per_slab)
    mms = [d["min_max"].tolist() for d in per_slab]
    mn, mx = min(m[0] for m in mms), max(m[1] for m in mms)
    voxels = sum(p.e - p.s for p in plans) * int(np.prod(plane_shape))
    count = voxels * sink_ch[0]
    return {"shape_out": [sum(p.e - p.s for p in plans)] + list(plane_shape), "n_chunks": len(plans),
            "halo": r, "slab_planes": slab_planes, "stage_planes": stage, "l2": math.sqrt(l2sq),
            "sum": total, "mean": total / count if count else 0.0, "min": mn, "max": mx,
            "voxels": voxels, "wall_s": wall, "gvoxel_s": voxels / wall / 1e9,
            "host_gen_s": t_gen[0], "h2d_bytes": h2d_bytes[0],
            "h2d_gbs_end_to_end": h2d_bytes[0] / wall / 1e9, "captured_shapes": len(captured),
            "out_path": out_path}


def _capture_window(rt: Runtime, g: Graph, wanted: Sequence[str], win: Tuple[int, int]):
    """hipGraph of one fused slab evaluation with the reductions' z-window (replays reuse the
    returned result tensors).  The plan was compiled by an earlier eager run of this shape."""
    from .fused import CapturedGraph
    ex = rt.fused
    cg = torch.cuda.CUDAGraph()
    # thread-local: the staging threads keep synchronising on copy events meanwhile
    with torch.cuda.graph(cg, capture_error_mode="thread_local"):
        res = ex.run(g, rt, wanted=wanted, red_window=win)
    keep = [c for plan in ex._plans.values() for c in plan[2].values()]
    return CapturedGraph(cg, keep), res


class _Writer:
    """Cropped slab outputs -> pinned host buffers (D2H on a side stream) -> a writer thread
    that copies them into an ``.npy`` memory map.  Exceptions surface in ``check``/``close``."""

    def __init__(self, path: str, n0: int, dev):
        self.path, self.n0, self.dev = path, n0, dev
        self.mm = None
        self.item_shape: Optional[Tuple[int, ...]] = None
        self.stream = torch.cuda.Stream(device=dev)
        self.q: "queue.Queue" = queue.Queue(maxsize=2)
        self.err: List[BaseException] = []
        self.bufs: List[torch.Tensor] = []
        self.free: "queue.Queue" = queue.Queue()
        self.th = threading.Thread(target=self._run, daemon=True)
        self.th.start()

    def put(self, out: torch.Tensor, z0: int) -> None:
        self.check()
        if self.mm is None:
            self.item_shape = tuple(out.shape[3:])
            self.mm = np.lib.format.open_memmap(self.path, mode="w+", dtype=np.float32,
                                                shape=(self.n0,) + tuple(out.shape[1:]))
            for _ in range(2):
                self.free.put(torch.empty(out.shape, dtype=torch.float32, pin_memory=True))
        host = self.free.get()
        if host.shape != out.shape:
            host = torch.empty(out.shape, dtype=torch.float32, pin_memory=True)
        self.stream.wait_stream(torch.cuda.current_stream(self.dev))
        with torch.cuda.stream(self.stream):
            host.copy_(out, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(self.stream)
        out.record_stream(self.stream)
        self.q.put((host, ev, z0))

    def _run(self) -> None:
        while True:
            item = self.q.get()
            if item is None:
                return
            host, ev, z0 = item
            try:
                ev.synchronize()
                self.mm[z0:z0 + host.shape[0]] = host.numpy()
            except BaseException as exc:
                self.err.append(exc)
            self.free.put(host)

    def check(self) -> None:
        if self.err:
            raise RuntimeError(f"slab writer failed: {self.err[0]}") from self.err[0]

    def close(self) -> None:
        self.q.put(None)
        self.th.join(timeout=120)
        if self.mm is not None:
            self.mm.flush()
        self.check()


def stream_execute(expr: str, lazy, out_path: Optional[str] = None, dx: float = 1.0,
                   boundary: str = "periodic", field_name: str = "f", **kw) -> Dict[str, Any]:
    """Compile ``expr`` over one lazily read field and stream it (stenpy.py:474-615)."""
    from .compiler import compile_expression
    probe = torch.zeros((1,) + tuple(lazy.shape[1:]), dtype=torch.float32)
    g, sink, _, fm, _ = compile_expression(expr, dx=dx, boundary=boundary, field_map={field_name: probe})
    return stream_graph(g, sink, fm[field_name], lazy, out_path, boundary=boundary, **kw)

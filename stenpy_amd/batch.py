"""Batched evaluation: many graphs over many grids of uneven shape in persistent launches.

Every graph is planned into fused groups as for a single evaluation.  Groups are placed in
dependency levels (a group that reads a reduction result or a field made by another group
comes one level later).  For each level one hipRTC module holds the code of every distinct
group structure (a namespace per structure, grid sizes read from per-grid parameters, the
LDS arrays views of one dynamic shared buffer) and one persistent kernel walks a work list
of (structure, grid, tile-or-tail-block) units; a second launch does every FP64 final
combine of that level's reductions.  Buffers, parameters and work lists are built once, so
``run`` only launches and the whole batch can be replayed from a captured hipGraph.

Reference counterpart: none -- the reference evaluates one graph per ``Runtime.run`` call
(stenpy_engine.py:945-1146); batching is this backend's addition.  Per-grid results follow
the same per-voxel semantics as the fused executor (and its parity tests).
"""
from __future__ import annotations

import math
import os
import struct
from dataclasses import dataclass, field as dc_field
from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib as L
from . import fused as F
from .graph import Graph

# block size shared by every structure of a batch module (one kernel, one block shape)
BATCH_THREADS = int(os.environ.get("STENPY_BATCH_THREADS", "256"))
_N_INT_FIELDS = 58            # ints after the pointers in a batched Params (see KernelGen._body)


def batch_schedule(grid: Sequence[int], threads: int = BATCH_THREADS) -> F.Schedule:
    """Heuristic tile for one structure of a batch (block size fixed to ``threads``)."""
    h = F.heuristic_schedule(grid)
    tx = min(h.tx, threads)
    ty = 2 * (threads // tx)
    n0, n1, n2 = grid
    tiles = -(-n2 // tx) * -(-n1 // ty)
    zc = 256
    while zc > 32 and (zc >= n0 or tiles * -(-n0 // zc) < 64):
        zc //= 2
    return F.Schedule(tx, ty, zc, threads)


def structure_key(group, graph: Graph) -> str:
    """Group structure without the grid (groups with equal keys share one code section)."""
    from .tune import group_signature
    sig = group_signature(group, graph)
    return sig.split("|", 1)[1] if "|" in sig else sig


def queue_split(costs: np.ndarray, n_queues: int = 8) -> np.ndarray:
    """Boundaries [0, b1, ..., n] splitting units (in their order) into ``n_queues``
    contiguous runs of about equal summed cost."""
    c = np.cumsum(np.asarray(costs, dtype=np.float64))
    tot = c[-1] if len(c) else 0.0
    cuts = [0] + [int(np.searchsorted(c, tot * q / n_queues, side="right")) for q in range(1, n_queues)]
    cuts = [min(max(v, 0), len(c)) for v in cuts] + [len(c)]
    for i in range(1, len(cuts)):
        cuts[i] = max(cuts[i], cuts[i - 1])
    return np.asarray(cuts, dtype=np.int32)


@dataclass
class _Entry:
    """One group of one grid inside a level."""
    item: int
    step: int
    group: Any
    key: str
    srcs: List[torch.Tensor] = dc_field(default_factory=list)
    outs: List[torch.Tensor] = dc_field(default_factory=list)
    reds: List[torch.Tensor] = dc_field(default_factory=list)
    scal: List[torch.Tensor] = dc_field(default_factory=list)
    nb: int = 0
    nbt: int = 0


class BatchRunner:
    """Plan, buffers and launches for a list of (graph, wanted node ids)."""

    def __init__(self, graphs: Sequence[Tuple[Graph, Sequence[str]]], device="cuda",
                 threads: int = BATCH_THREADS, schedule_fn=None):
        self.device = torch.device(device)
        self.threads = threads
        self.schedule_fn = schedule_fn
        self.items = []
        self.results: List[Dict[str, Any]] = []
        for gi, (graph, wanted) in enumerate(graphs):
            wanted = list(wanted)
            values = {n: graph.node(n).params["value"] for n in graph._order
                      if graph.node(n).op_name == "_constant"}
            steps, info = F.plan_graph(graph, values, wanted)
            eager = [graph.node(s.node).op_name for s in steps
                     if s.kind != "group" and info[s.node].ch != 0]
            if eager:
                raise NotImplementedError(f"graph {gi}: ops {sorted(set(eager))} are not fusable; "
                                          "a batch needs fully fused graphs")
            res: Dict[str, Any] = {}
            for nid, v in values.items():
                if isinstance(v, torch.Tensor) and v.ndim >= 3:
                    t = v.to(device=self.device, dtype=torch.float32).contiguous()
                    res[nid] = t
                else:
                    res[nid] = v
            self.items.append((graph, wanted, steps, info))
            self.results.append(res)
        self._plan_levels()
        self._build_modules()
        self._allocate()
        self._pack()

    # -- planning --------------------------------------------------------------------------------
    def _plan_levels(self) -> None:
        """Levels of groups; scalar steps (ops on reduction results) run at the start of the
        first level whose launch needs them."""
        self.levels: List[List[_Entry]] = []
        self.pre_steps: Dict[int, List[Tuple[int, str]]] = {}
        for ii, (graph, wanted, steps, info) in enumerate(self.items):
            made: Dict[str, int] = {}          # node id -> level of the group that makes it
            for si, st in enumerate(steps):
                if st.kind != "group":       # scalar op on reduction results / constants
                    node = graph.node(st.node)
                    ready = max([made[i] + 1 for i in node.input_ids if i in made] + [0])
                    self.pre_steps.setdefault(ready, []).append((ii, st.node))
                    made[st.node] = ready - 1
                    while len(self.levels) <= ready:
                        self.levels.append([])
                    continue
                g = st.group
                lvl = 0
                for dep in list(g.sources) + list(g.dscalars):
                    if dep in made:
                        lvl = max(lvl, made[dep] + 1)
                for o in list(g.outputs) + list(g.reductions):
                    made[o] = lvl
                while len(self.levels) <= lvl:
                    self.levels.append([])
                self.levels[lvl].append(_Entry(ii, si, g, structure_key(g, graph)))

    def _schedule(self, entries: List[_Entry]) -> F.Schedule:
        grids = sorted(e.group.grid for e in entries)
        med = grids[len(grids) // 2]
        e0 = entries[0]
        graph = self.items[e0.item][0]
        if self.schedule_fn is not None:
            s = self.schedule_fn(e0.group, graph)
            if s is not None and s.threads == self.threads:
                return s
        return batch_schedule(med, self.threads)

    def _build_modules(self) -> None:
        self.mods = []
        for lvl, entries in enumerate(self.levels):
            if not entries:
                self.mods.append({"keys": [], "gens": [], "source": None, "dyn_shared": 0, "module": None})
                continue
            keys: List[str] = []
            for e in entries:
                if e.key not in keys:
                    keys.append(e.key)
            gens = []
            for k in keys:
                members = [e for e in entries if e.key == k]
                # compile-time choices use the smallest grid (they agree for axes >= 3)
                rep = min(members, key=lambda e: int(np.prod(e.group.grid)))
                graph, _, _, info = self.items[rep.item]
                sched = self._schedule(members)
                gen = None
                for cand in [sched] + [F.Schedule(32, 16, sched.zc, self.threads),
                                       F.Schedule(32, 8, sched.zc, self.threads)]:
                    try:
                        g_ = F.KernelGen(graph, rep.group, info, cand, {}, {}, batched=True)
                    except ValueError:
                        continue
                    if g_.lds_bytes() <= F.LDS_LIMIT:
                        gen = g_
                        break
                if gen is None:
                    raise RuntimeError(f"no batched schedule fits LDS for structure {k[:80]}")
                gens.append(gen)
            idx64 = any(int(np.prod(e.group.grid)) * 9 >= 2 ** 31 - 1 for e in entries)
            parts = F.KernelGen._header_sources() + ["extern __shared__ float sp_smem[];"]
            for ki, gen in enumerate(gens):
                parts.append(gen.section(f"g{ki}", idx64))
            parts.append(self._batch_kernel(len(gens)))
            src = "\n".join(parts)
            dyn = max(g.lds_array_bytes() for g in gens)
            self.mods.append({"keys": keys, "gens": gens, "source": src, "dyn_shared": dyn,
                              "module": None})

    def _batch_kernel(self, k: int) -> str:
        """Persistent kernel over 8 work queues (contiguous unit ranges of equal estimated
        cost).  A block starts on queue blockIdx % 8 -- its XCD's, so neighbouring tiles share
        that XCD's L2 -- and then steals from the others; each block scans every queue once,
        so every block exits.  The last block to finish resets the counters for the next
        launch (no memset, so the launch can be captured and replayed)."""
        nt = self.threads
        lines = ["struct BatchArgs {", "    const int4* units;", "    const int* qstart;", "    int* counters;",
                 f"    const void* params[{k}];", "};",
                 f'extern "C" __global__ void __launch_bounds__({nt}) batch_kernel(const BatchArgs a) {{',
                 f"    __shared__ float red_scratch[4 * {nt} / 64 + 4];",
                 "    __shared__ int s_w;",
                 "    const int xq = blockIdx.x & 7;",
                 "    for (int qi = 0; qi < 8; ++qi) {",
                 "        const int q = (xq + qi) & 7, q0 = a.qstart[q], q1 = a.qstart[q + 1];",
                 "        while (true) {",
                 "            if (threadIdx.x == 0) s_w = q0 + atomicAdd(a.counters + q, 1);",
                 "            __syncthreads();",
                 "            const int w = s_w;",
                 "            __syncthreads();   // every thread has read s_w",
                 "            if (w >= q1) break;",
                 "            const int4 u = a.units[w];",
                 "            switch (u.x) {"]
        for i in range(k):
            # a local copy: no output store can alias it, so its fields stay in registers
            lines.append(f"            case {i}: {{ const g{i}::Params prm = reinterpret_cast<const g{i}::Params*>(a.params[{i}])[u.y];"
                         f" g{i}::run_unit(prm, u.z, red_scratch); }} break;")
        lines += ["            default: break;", "            }",
                  "            __syncthreads();   // the next unit reuses the LDS", "        }", "    }",
                  "    if (threadIdx.x == 0) {",
                  "        __threadfence();",
                  "        if (atomicAdd(a.counters + 8, 1) == (int)gridDim.x - 1) {",
                  "            for (int q = 0; q < 9; ++q) atomicExch(a.counters + q, 0);",
                  "            __threadfence();",
                  "        }",
                  "    }",
                  "}"]
        return "\n".join(lines)

    def schedules(self) -> List[List[List[int]]]:
        """The schedule of each kernel structure, per level (TX, TY, ZC, threads)."""
        return [[list(g.sched.as_tuple()) for g in m["gens"]] for m in self.mods]

    # -- buffers and parameters ------------------------------------------------------------------
    def _allocate(self) -> None:
        dev = self.device
        for lvl, entries in enumerate(self.levels):
            for ii, nid in self.pre_steps.get(lvl, []):
                self.results[ii][nid] = torch.empty(1, dtype=torch.float64, device=dev)
            mod = self.mods[lvl]
            for e in entries:
                graph, _, _, info = self.items[e.item]
                gen = mod["gens"][mod["keys"].index(e.key)]
                res = self.results[e.item]
                n0, n1, n2 = e.group.grid
                s = gen.sched
                e.nb = -(-n2 // s.tx) * -(-n1 // s.ty) * -(-n0 // s.zc)
                e.nbt = gen.boundary_blocks(e.group.grid)
                e.srcs = []
                for sid in e.group.sources:
                    t = res[sid]
                    if not isinstance(t, torch.Tensor):
                        raise RuntimeError(f"source {sid} of item {e.item} is not a field")
                    if t.dtype != torch.float32 or not t.is_contiguous():
                        t = t.to(torch.float32).contiguous()
                        res[sid] = t
                    e.srcs.append(t)
                e.outs = []
                for o in e.group.outputs:
                    ch = info[o].ch
                    t = torch.empty((n0, n1, n2) + {1: (), 3: (3,), 9: (3, 3)}[ch], dtype=torch.float32, device=dev)
                    res[o] = t
                    e.outs.append(t)
                e.reds = [torch.empty((e.nb + e.nbt) * 4, dtype=torch.float64, device=dev)
                          for _ in e.group.reductions]
                for r in e.group.reductions:
                    op = graph.node(r).op_name
                    res[r] = torch.empty(2 if op == "min_max" else 1, dtype=torch.float64, device=dev)
                e.scal = []
                for d in e.group.dscalars:
                    t = res[d]
                    if not isinstance(t, torch.Tensor):
                        t = torch.tensor([float(t)], dtype=torch.float64, device=dev)
                        res[d] = t
                    e.scal.append(t)

    def _params_bytes(self, e: _Entry, gen: "F.KernelGen") -> bytes:
        s = gen.sched
        n0, n1, n2 = e.group.grid
        ptrs = [t.data_ptr() for t in e.srcs] + [t.data_ptr() for t in e.outs] + \
               [t.data_ptr() for t in e.reds] + [t.data_ptr() for t in e.scal]
        tx_, ty_, zcn = -(-n2 // s.tx), -(-n1 // s.ty), -(-n0 // s.zc)
        _, _, ylo, yhi, xlo, xhi = gen.interior_bounds(e.group.grid)
        boxes = gen.boundary_boxes(e.group.grid)
        flat, starts, acc = [], [0], 0
        for k in range(6):
            if k < len(boxes):
                b = boxes[k]
                flat += list(b)
                acc += (b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4])
            else:
                flat += [0, 0, 0, 0, 0, 0]
            starts.append(acc)
        starts.append(acc)
        ints = [tx_, ty_, zcn, e.nb] + [0] + [n0, n1, n2, ylo, yhi, xlo, xhi, e.nbt] + starts + flat
        assert len(ints) == _N_INT_FIELDS - 1
        return struct.pack(f"<{len(ptrs)}Q4i2i8i8i36i", *ptrs, tx_, ty_, zcn, e.nb, *F.Z_ALL,
                           n0, n1, n2, ylo, yhi, xlo, xhi, e.nbt, *starts, *flat)

    def _pack(self) -> None:
        dev = self.device
        ncu = torch.cuda.get_device_properties(dev).multi_processor_count if dev.type == "cuda" else 256
        self.launches = []
        for lvl, entries in enumerate(self.levels):
            mod = self.mods[lvl]
            keys = mod["keys"]
            per_key: List[List[_Entry]] = [[e for e in entries if e.key == k] for k in keys]
            params = []
            for ki, es in enumerate(per_key):
                blob = b"".join(self._params_bytes(e, mod["gens"][ki]) for e in es)
                params.append(torch.frombuffer(bytearray(blob), dtype=torch.uint8).to(dev))
            units, costs = [], []
            for ki, es in enumerate(per_key):
                gen = mod["gens"][ki]
                s = gen.sched
                w = self._cost_weight(gen)
                for ii, e in enumerate(es):
                    n0 = e.group.grid[0]
                    tile = s.tx * s.ty * min(s.zc, n0)
                    tail = self.threads * gen.TAIL_VPT
                    units += [(ki, ii, b, 0) for b in range(e.nb + e.nbt)]
                    costs += [w * tile] * e.nb + [w * tail] * e.nbt
            u = np.asarray(units, dtype=np.int32).reshape(-1, 4)
            qstart = queue_split(np.asarray(costs), 8)
            units_t = torch.from_numpy(np.ascontiguousarray(u)).to(dev)
            qstart_t = torch.from_numpy(qstart).to(dev)
            counters = torch.zeros(9, dtype=torch.int32, device=dev)
            descs = []
            for e in entries:
                graph = self.items[e.item][0]
                info = self.items[e.item][3]
                n = int(np.prod(e.group.grid))
                for r, ws in zip(e.group.reductions, e.reds):
                    node = graph.node(r)
                    cnt = n * max(info[node.input_ids[0]].ch, 1)
                    descs.append(struct.pack("<QQqii", ws.data_ptr(), self.results[e.item][r].data_ptr(),
                                             cnt, F.RED_KIND[node.op_name], e.nb + e.nbt))
            descs_t = (torch.frombuffer(bytearray(b"".join(descs)), dtype=torch.uint8).to(dev)
                       if descs else None)
            args = struct.pack(f"<3Q{max(len(params), 1)}Q", units_t.data_ptr(), qstart_t.data_ptr(),
                               counters.data_ptr(), *([p.data_ptr() for p in params] or [0]))
            self.launches.append({"params": params, "units": units_t, "nunits": len(units),
                                  "qstart": qstart_t, "counters": counters,
                                  "descs": descs_t, "ndescs": len(descs), "args": args,
                                  "ncu": ncu, "grid": None})

    @staticmethod
    def _cost_weight(gen: "F.KernelGen") -> float:
        """Relative cost per voxel of a structure: transcendentals (FP64, per tap) dominate."""
        from .tune import group_stats
        st = group_stats(gen)
        return 1.0 + 0.25 * st["taps"] + 8.0 * st["transc"]

    # -- execution -------------------------------------------------------------------------------
    def compile(self, workers: int = 1) -> None:
        srcs = [m["source"] for m in self.mods if m["module"] is None and m["source"]]
        if workers > 1 and len(srcs) > 1:
            from .tune import _compile_all
            _compile_all(srcs, workers)      # fills the code-object cache
        for m in self.mods:
            if m["module"] is None and m["source"]:
                code, _ = F.compile_source(m["source"], "batch_kernel")
                m["module"] = L.Module(code, "batch_kernel")

    def _grid(self, lvl: int) -> int:
        ln, m = self.launches[lvl], self.mods[lvl]
        if ln["grid"] is None:
            per_cu = max(1, m["module"].occupancy(self.threads, m["dyn_shared"]))
            g = per_cu * ln["ncu"]
            g = max(8, (g // 8) * 8)
            ln["grid"] = min(g, max(1, ln["nunits"]))
        return ln["grid"]

    def run(self, stream: Optional[int] = None) -> None:
        """All levels: one persistent launch each, then the level's batched FP64 combines."""
        self.compile()
        st = stream if stream is not None else L.stream_ptr(self.device)
        from .registry import OP_REGISTRY
        for lvl, ln in enumerate(self.launches):
            # the level's scalar steps (ops on earlier levels' reduction results) as one device
            # program; a step it cannot express runs through the registry, in order
            pending = []
            for ii, nid in self.pre_steps.get(lvl, []):
                graph = self.items[ii][0]
                node = graph.node(nid)
                res = self.results[ii]
                ins = [res[i] for i in node.input_ids]
                d = F.scalar_op(node, ins, res[nid])
                if d is not None:
                    pending.append(d)
                    continue
                if pending:
                    L.scalar_program(pending, st)
                    pending = []
                v = OP_REGISTRY[node.op_name](*ins, **node.params)
                res[nid].copy_(torch.as_tensor(v, device=self.device).reshape(-1)[:1].to(torch.float64))
            if pending:
                L.scalar_program(pending, st)
            if ln["nunits"] == 0:
                continue
            m = self.mods[lvl]
            m["module"].launch((self._grid(lvl), 1, 1), (self.threads, 1, 1), ln["args"],
                               m["dyn_shared"], st)
            if ln["ndescs"]:
                L.check(L.lib().sp_reduce_combine_batch(ln["descs"].data_ptr(), ln["ndescs"], st),
                        "combine batch")

    def outputs(self, item: int) -> Dict[str, Any]:
        graph, wanted, _, _ = self.items[item]
        res = self.results[item]
        out = {}
        for w in wanted:
            v = res[w]
            if isinstance(v, torch.Tensor) and v.dtype == torch.float64 and v.numel() == 1:
                v = v.reshape(())
            out[w] = v
        return out

    def capture(self, warmup: int = 1) -> "F.CapturedGraph":
        """A hipGraph replaying every launch of the batch (buffers are fixed at plan time); it
        keeps this runner, and with it the modules and buffers it launches, alive."""
        self.compile()
        for lvl in range(len(self.launches)):
            if self.launches[lvl]["nunits"]:
                self._grid(lvl)
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            for _ in range(warmup):
                self.run()
        torch.cuda.current_stream(self.device).wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run()
        return F.CapturedGraph(g, [self])

    def stats(self) -> Dict[str, Any]:
        """Units, grids and algorithmic bytes (each source channel read once, each output
        channel written once) of the batch."""
        tot_bytes, tot_vox = 0, 0
        for lvl, entries in enumerate(self.levels):
            for e in entries:
                n = int(np.prod(e.group.grid))
                tot_bytes += n * 4 * (sum(max(t.shape[3:].numel(), 1) for t in e.srcs) +
                                      sum(max(t.shape[3:].numel(), 1) for t in e.outs))
        for graph, wanted, steps, info in self.items:
            tot_vox += int(np.prod(steps[0].group.grid))
        return {"levels": len(self.levels), "units": [ln["nunits"] for ln in self.launches],
                "structures": [len(m["keys"]) for m in self.mods], "voxels": tot_vox,
                "algorithmic_bytes": tot_bytes}

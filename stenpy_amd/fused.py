"""Fused executor: graph -> CDNA4 HIP stencil kernels compiled at run time with hipRTC.

Replaces the reference's fusion pipeline (fused_compiler.py: ``plan_fusion`` 273-321,
``lower_graph`` 329-402, ``TritonCodegen`` 404-529, ``KernelCache`` 531-557,
``compile_and_execute`` 632-753; entry ``Runtime._run_fused`` stenpy_engine.py:966-981).  The
reference path is unreachable in practice (circular import) and only ever lowered point-wise
ops plus the 7-point Laplacian; here every stencil, point-wise, per-voxel tensor and
reduction operator of the hot path fuses, with the CPU eager semantics as the parity target.

Planning.  Nodes are assigned *levels*: a fusable node joins the level of its inputs unless
an input is a reduction or an eager (non-fusable) node, or the composite stencil reach
would exceed the halo limit; fusable nodes of one level and one grid form a group -> one
kernel.  Inside a group nothing is materialised except requested outputs and values read by
later steps; stencils of stencils are evaluated by recomputation at offsets.

Kernel shape (one block = TY x TX voxels of a plane, marching along axis 0 over ZC planes):
  * every field source has an LDS ring of D = pow2(2*Rint + 2) planes of (TY+2H) x (TX+2H)
    cells per channel; the next plane is prefetched into registers while the current plane
    is computed, and one barrier per plane suffices because the ring has a spare slot;
  * loads are flat and coalesced over (row, x, channel), so channel-last vector fields are
    read as contiguous rows and de-interleaved into per-channel LDS planes;
  * the per-voxel expression is generated once per *variant*: LDS reads with compile-time
    offsets for tiles away from the x/y boundaries, runtime boundary transforms on the axes
    whose tile touches a boundary, and a global-gather variant for the few planes near the
    z boundaries (where one-sided torch.gradient formulas reach beyond the ring);
  * block ids are remapped so that each of the 8 XCDs (own L2) walks a contiguous range of
    tiles, keeping halo re-reads inside one L2;
  * reductions keep compensated FP32 state per thread (Neumaier / Welford), merge across the
    block with shuffles and LDS, and write FP64 block records that ``sp_reduce_combine``
    merges in FP64.
"""
from __future__ import annotations

import ctypes
import hashlib
import math
import os
import struct
import threading
import time
import warnings
from dataclasses import dataclass, field as dc_field
from pathlib import Path
from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib as L
from .graph import Graph, GraphNode

CSRC = Path(__file__).resolve().parent / "csrc"

# ------------------------------------------------------------------------------ op tables
UNARY_OPS = {"neg", "exp", "log", "sqrt", "sin", "cos", "tanh", "clamp", "cast_bf16"}
BINARY_OPS = {"add", "sub", "mul", "div", "velocity_dot_grad", "div_last"}
STENCIL_OPS = {"laplacian", "gradient", "divergence", "curl", "gradient_nd", "hessian",
               "jacobian", "mean_curvature", "material_derivative"}
TENSOR_OPS = {"trace", "determinant", "eigenvalues", "deviatoric", "inverse"}
REDUCTION_OPS = {"sum", "mean", "norm_l2", "variance", "min_max", "integrate"}
CHANNEL_OPS = {"dot", "sqnorm", "select_component"}

# (interior reach, edge reach) per stencil op: the interior reach sizes the plane ring, the
# edge reach (one-sided torch.gradient formulas look two cells inward) sizes the in-plane halo.
STENCIL_REACH = {"laplacian": (1, 1), "gradient": (1, 2), "divergence": (1, 2), "curl": (1, 2),
                 "gradient_nd": (1, 2), "hessian": (2, 4), "jacobian": (1, 2),
                 "mean_curvature": (2, 4), "material_derivative": (1, 2)}
MAX_RINT = 2
MAX_REDGE = 4
RED_KIND = {"sum": 0, "mean": 1, "norm_l2": 2, "variance": 3, "min_max": 5, "integrate": 6}


@dataclass(frozen=True)
class Schedule:
    """Tile width (x, contiguous), tile height (y), planes per block (z) and block size."""
    tx: int = 64
    ty: int = 8
    zc: int = 64
    threads: int = 256

    def as_tuple(self):
        return (self.tx, self.ty, self.zc, self.threads)


DEFAULT_SCHEDULE = Schedule()


def heuristic_schedule(grid: Sequence[int]) -> Schedule:
    """Shape-only schedule used when no tuned or learned schedule is available.

    Wide tiles cut the x-halo share, VPT = 2 rows per thread, and z-chunks long enough that
    the halo planes of a chunk are a small share while leaving >= ~2 blocks per CU."""
    n0, n1, n2 = (int(v) for v in grid)
    tx = 128 if n2 >= 384 else 64 if n2 >= 96 else 32 if n2 >= 32 else 16
    ty = 8 if n1 >= 16 else 4
    nt = min(512, max(64, tx * ty // 2))
    tiles = -(-n2 // tx) * -(-n1 // ty)
    zc = 256
    while zc > 32 and (zc >= n0 or tiles * -(-n0 // zc) < 512):
        zc //= 2
    return Schedule(tx, ty, zc, nt)
def fill_small_grid(s: Schedule, grid: Sequence[int], min_blocks: int = 512, min_zc: int = 8) -> Schedule:
    """``s`` with its z-chunk halved until the launch has ``min_blocks`` blocks (two per CU) or
    the chunk reaches ``min_zc`` planes.  The schedule search and the model rank chunks of 32
    planes and more, which leave a small grid (128^3: 256 blocks of a 64 x 8 tile) to a
    fraction of the GPU's 256 CUs."""
    n0, n1, n2 = (int(v) for v in grid)
    tiles = -(-n2 // s.tx) * -(-n1 // s.ty)
    zc = s.zc
    while zc > min_zc and tiles * -(-n0 // zc) < min_blocks:
        zc //= 2
    return s if zc == s.zc else Schedule(s.tx, s.ty, max(zc, min_zc), s.threads)


# LDS per block kept within 80 KiB so that at least two blocks share a CU (160 KiB per CU)
LDS_LIMIT = 80 * 1024
FALLBACK_SCHEDULES = [Schedule(64, 4, 64, 256), Schedule(32, 8, 64, 256), Schedule(32, 4, 64, 128),
                      Schedule(16, 8, 64, 128), Schedule(16, 4, 64, 64)]
# Edge handling of the main kernel.  "band": only voxels at least the edge reach away from
# every face run in the LDS kernel (one code variant, low register pressure) and the thin
# face band goes to the gather kernel.  "variants": tiles touching an x/y face run
# edge-specialised LDS variants.  "auto" picks "band" when the band is a small share.
EDGE_MODE = os.environ.get("STENPY_EDGE_MODE", "auto")
# The gather evaluation of the face band (z faces, y faces, corners, partial tiles) runs as a
# separate small kernel (no LDS arrays, few registers) on a side stream, beside the LDS
# z-marching kernel; off: the same work as extra blocks at the end of the main grid.
# where the edge band's gather blocks run: "auto" puts them at the end of the main grid unless
# that code pushes the main kernel past MERGED_TAIL_MAX_VGPRS (below 4 waves per SIMD), in which
# case they run as a separate kernel on a side stream; "1" / "0" force either
SPLIT_TAIL = os.environ.get("STENPY_SPLIT_TAIL", "auto")
MERGED_TAIL_MAX_VGPRS = 128
# stage wide channel-last outputs through LDS so every store instruction writes whole lines
STAGE_STORES = os.environ.get("STENPY_STAGE_STORES", "1") == "1"
# the fewest channels an output needs for staging (a lane's C values are C * 4 bytes apart)
STAGE_MIN_CH = int(os.environ.get("STENPY_STAGE_MIN_CH", "4"))
COMBINE_MULTI_MAX = 16        # descriptors per sp_reduce_combine_multi launch
BAND_MAX_FRACTION = 0.08


# ------------------------------------------------------------------------------ node typing
@dataclass
class Info:
    kind: str                    # "field" | "const" | "dscalar" | "fused" | "eager"
    ch: int = 1                  # channels per voxel (1, 3, 9); 0 for scalars
    grid: Optional[Tuple[int, int, int]] = None
    value: Optional[float] = None
    level: int = 0
    rint: int = 0
    redge: int = 0


def _is_field_tensor(v: Any) -> bool:
    return isinstance(v, torch.Tensor) and v.ndim in (3, 4, 5) and v.dtype in (
        torch.float32, torch.float64, torch.bfloat16)


def _field_channels(t: torch.Tensor) -> Optional[int]:
    if t.ndim == 3:
        return 1
    if t.ndim == 4 and t.shape[3] == 3:
        return 3
    if t.ndim == 5 and t.shape[3:] == (3, 3):
        return 9
    return None


def _scalar_value(v: Any) -> Optional[float]:
    if isinstance(v, (int, float)):
        return float(v)
    if isinstance(v, torch.Tensor) and v.numel() == 1 and v.ndim <= 1 and not v.is_cuda:
        return float(v.reshape(-1)[0])
    return None


def _value_signature(v: Any) -> Tuple:
    """What a plan depends on in a seeded or materialised value: a host scalar is compiled into
    the kernel as a literal, so its value is part of the plan; a device scalar is read at run
    time and a field only contributes its shape and dtype."""
    s = _scalar_value(v)
    if s is not None:
        return ("const", s)
    if isinstance(v, torch.Tensor):
        return ("tensor", tuple(v.shape), str(v.dtype), v.is_cuda)
    return ("other", type(v).__name__)


def _dx3(dx: Any) -> Tuple[float, float, float]:
    if isinstance(dx, (int, float)):
        return (float(dx),) * 3
    dx = tuple(float(v) for v in dx)
    if len(dx) != 3:
        raise ValueError("dx must have 3 entries")
    return dx


def _fusable(node: GraphNode, ins: List[Info]) -> Optional[Tuple[int, Tuple[int, int, int]]]:
    """(channels, grid) of a fusable node, or None."""
    op = node.op_name
    grids = {i.grid for i in ins if i.grid is not None}
    if len(grids) != 1:
        return None
    grid = next(iter(grids))
    if min(grid) < 3:
        return None
    chs = [i.ch for i in ins]
    p = node.params
    if op in UNARY_OPS:
        return (chs[0], grid) if len(ins) == 1 else None
    if op in BINARY_OPS:
        if op == "mul" and len(ins) == 1 and p.get("scalar") is not None:
            return chs[0], grid
        if len(ins) != 2 or op == "div_last":
            return None
        a, b = ins
        if a.ch == b.ch or a.ch == 0 or b.ch == 0:
            return max(a.ch, b.ch), grid
        return None
    if op == "dot":
        return (1, grid) if len(ins) == 2 and chs == [3, 3] else None
    if op == "sqnorm":
        return (1, grid) if chs == [3] else None
    if op == "select_component":
        return (1, grid) if len(ins) == 1 and chs[0] in (3, 9) and "index" in p else None
    if op in STENCIL_OPS:
        dx = p.get("dx", 1.0)
        try:
            _dx3(dx)
        except (ValueError, TypeError):
            return None
        if op == "laplacian":
            return (1, grid) if chs == [1] else None
        if op == "gradient":
            if chs != [1] or not isinstance(dx, (int, float)):
                return None
            return (1 if p.get("dim") is not None else 3), grid
        if op in ("divergence", "curl", "jacobian"):
            if chs != [3]:
                return None
            return {"divergence": 1, "curl": 3, "jacobian": 9}[op], grid
        if op in ("gradient_nd", "hessian", "mean_curvature"):
            if chs != [1]:
                return None
            return {"gradient_nd": 3, "hessian": 9, "mean_curvature": 1}[op], grid
        if op == "material_derivative":
            return (1, grid) if chs == [1, 3] else None
    if op in TENSOR_OPS:
        if chs != [9]:
            return None
        return {"trace": 1, "determinant": 1, "eigenvalues": 3, "deviatoric": 9, "inverse": 9}[op], grid
    if op in REDUCTION_OPS:
        if len(ins) != 1 or chs[0] == 0 or any(k in p and p[k] is not None for k in ("dim",)) or p.get("keepdim"):
            return None
        if op == "integrate":
            if chs != [1] or p.get("dims") is not None:
                return None
        if op == "variance" and p.get("unbiased", True) is False:
            return None
        return 0, grid
    return None


# ------------------------------------------------------------------------------ planning
@dataclass
class Group:
    level: int
    grid: Tuple[int, int, int]
    nodes: List[str] = dc_field(default_factory=list)          # fused compute nodes (topo)
    sources: List[str] = dc_field(default_factory=list)        # field inputs
    dscalars: List[str] = dc_field(default_factory=list)       # device 0-d inputs
    outputs: List[str] = dc_field(default_factory=list)        # materialised nodes
    reductions: List[str] = dc_field(default_factory=list)
    rint: int = 0
    redge: int = 0


@dataclass
class Step:
    kind: str            # "eager" | "group"
    node: Optional[str] = None
    group: Optional[Group] = None
    level: int = 0


def plan_graph(graph: Graph, values: Dict[str, Any], wanted: Sequence[str]) -> Tuple[List[Step], Dict[str, Info]]:
    """Partition ``graph`` into eager steps and fused groups."""
    info: Dict[str, Info] = {}
    order = graph.topological_sort()
    for node in order:
        nid = node.id
        if nid in values:
            v = values[nid]
            if _is_field_tensor(v) and _field_channels(v) is not None:
                info[nid] = Info("field", _field_channels(v), tuple(v.shape[:3]))
            elif isinstance(v, torch.Tensor) and v.numel() == 1 and v.ndim <= 1 and v.is_cuda:
                info[nid] = Info("dscalar", 0)
            elif _scalar_value(v) is not None:
                info[nid] = Info("const", 0, value=_scalar_value(v))
            else:
                info[nid] = Info("eager", 1)
            continue
        if node.op_name == "_constant":
            v = node.params.get("value")
            if _is_field_tensor(v) and _field_channels(v) is not None:
                info[nid] = Info("field", _field_channels(v), tuple(v.shape[:3]))
            elif _scalar_value(v) is not None:
                info[nid] = Info("const", 0, value=_scalar_value(v))
            elif isinstance(v, torch.Tensor) and v.numel() == 1 and v.ndim <= 1:
                info[nid] = Info("dscalar", 0)
            else:
                info[nid] = Info("eager", 1)
            continue
        ins = [info[i] for i in node.input_ids]
        fz = None
        if all(i.ch >= 0 for i in ins) and any(i.ch > 0 for i in ins):
            fz = _fusable(node, ins)
        if fz is None:
            lvl = max([i.level for i in ins], default=0) + 1
            # scalar in -> scalar out; anything else has a shape known only at run time
            ch = 0 if all(i.ch == 0 for i in ins) else -1
            info[nid] = Info("eager", ch, None, level=lvl)
            continue
        ch, grid = fz
        # level and composite reach
        lvl, rint, redge = 0, 0, 0
        for i_id, i in zip(node.input_ids, ins):
            if i.kind == "fused":
                in_op = graph.node(i_id).op_name
                if in_op in REDUCTION_OPS:
                    lvl = max(lvl, i.level + 1)
                else:
                    lvl = max(lvl, i.level)
            elif i.kind in ("eager",):
                lvl = max(lvl, i.level + 1)
            elif i.kind == "dscalar":
                lvl = max(lvl, 1)
        op_int, op_edge = STENCIL_REACH.get(node.op_name, (0, 0))
        for i_id, i in zip(node.input_ids, ins):
            if i.kind == "fused" and i.level == lvl and graph.node(i_id).op_name not in REDUCTION_OPS:
                rint = max(rint, i.rint)
                redge = max(redge, i.redge)
        rint += op_int
        redge += op_edge
        if rint > MAX_RINT or redge > MAX_REDGE:
            lvl += 1
            rint, redge = op_int, op_edge
        info[nid] = Info("fused", ch, grid, level=lvl, rint=rint, redge=redge)

    consumers = graph.build_consumer_map()
    groups: Dict[Tuple[int, Tuple[int, int, int]], Group] = {}
    steps: List[Step] = []
    for node in order:
        nid = node.id
        inf = info[nid]
        if nid in values or node.op_name == "_constant":
            continue
        if inf.kind == "eager":
            steps.append(Step("eager", node=nid, level=inf.level))
            continue
        key = (inf.level, inf.grid)
        g = groups.get(key)
        if g is None:
            g = Group(level=inf.level, grid=inf.grid)
            groups[key] = g
            steps.append(Step("group", group=g, level=inf.level))
        g.nodes.append(nid)
        g._info = info
        if node.op_name in REDUCTION_OPS:
            g.reductions.append(nid)
    # sources, scalars, outputs, reach
    wanted = set(wanted)
    for st in steps:
        if st.kind != "group":
            continue
        g = st.group
        members = set(g.nodes)
        for nid in g.nodes:
            node = graph.node(nid)
            for i in node.input_ids:
                ii = info[i]
                inside = i in members and graph.node(i).op_name not in REDUCTION_OPS
                if inside:
                    continue
                if ii.kind == "field" or (ii.kind == "fused" and ii.ch > 0):
                    if i not in g.sources:
                        g.sources.append(i)
                elif ii.kind == "dscalar" or (ii.kind in ("fused", "eager") and ii.ch == 0):
                    if i not in g.dscalars:
                        g.dscalars.append(i)
                elif ii.kind != "const":
                    raise RuntimeError(f"fused group input {i} has unknown shape")
            if node.op_name in REDUCTION_OPS:
                continue
            outside = [c for c in consumers[nid] if c not in members]
            if nid in wanted or outside:
                g.outputs.append(nid)
            g.rint = max(g.rint, info[nid].rint)
            g.redge = max(g.redge, info[nid].redge)
    steps.sort(key=lambda s: s.level)
    return steps, info


# ------------------------------------------------------------------------------ codegen
def _ct_index(fn: str, c: int, n: int) -> int:
    """Compile-time counterparts of the device coordinate helpers wrapc/clampc/reflc."""
    if fn == "wrapc":
        return c + n if c < 0 else (c - n if c >= n else c)
    if fn == "clampc":
        return 0 if c < 0 else (n - 1 if c >= n else c)
    if fn == "reflc":
        if c < 0:
            return -c if -c < n - 1 else n - 1
        if c >= n:
            return 2 * (n - 1) - c if 2 * (n - 1) - c > 0 else 0
        return c
    raise KeyError(fn)


def _flt(x: float) -> str:
    f = float(np.float32(x))
    if math.isnan(f):
        return "NAN"
    if math.isinf(f):
        return "INFINITY" if f > 0 else "(-INFINITY)"
    r = repr(f)
    if "e" not in r and "." not in r:
        r += ".0"
    return r + "f"


class _NeedRing(Exception):
    """A queue-mode source was tapped off its own column at another plane."""

    def __init__(self, si: int):
        super().__init__(si)
        self.si = si


class _Emit:
    """Per-variant expression emitter with (node, position) memoisation.

    ``mode`` is ("lds", ex, ey), ("xface", side, d) or ("gather",).  A position is a 3-tuple
    of axis coordinates, each ("I", k) -- the voxel's own coordinate plus a compile-time
    offset --, ("E", expr) -- a runtime expression for an absolute coordinate -- or ("A", c),
    an absolute coordinate known at compile time (the x coordinate in the x-face variants,
    which evaluate the voxels d cells from the low or high x face of a full tile).
    """

    AX_VAR = ("z", "y", "x")
    AX_N = ("N0", "N1", "N2")

    def __init__(self, gen: "KernelGen", mode: Tuple, record: Optional[Dict[int, set]] = None):
        self.g = gen
        self.mode = mode
        self.record = record
        self.lines: List[str] = []
        self.memo: Dict[Any, List[str]] = {}
        self.cmemo: Dict[str, str] = {}
        self.n = 0

    # -- helpers -----------------------------------------------------------------------
    def var(self, expr: str, ty: str = "float") -> str:
        name = f"{'v' if ty == 'float' else 'c'}{self.n}"
        self.n += 1
        self.lines.append(f"const {ty} {name} = {expr};")
        return name

    def edge_axis(self, a: int) -> bool:
        if self.mode[0] == "gather":
            return True
        if self.mode[0] == "xface" or a == 0:
            return False
        return bool(self.mode[1] if a == 2 else self.mode[2])

    def face_x(self) -> int:
        _, side, d = self.mode
        return d if side == "lo" else self.g.grid[2] - 1 - d

    def origin(self) -> Tuple:
        if self.mode[0] == "xface":
            return (("I", 0), ("I", 0), ("A", self.face_x()))
        return tuple(("E", self.AX_VAR[a]) if self.edge_axis(a) else ("I", 0) for a in range(3))

    def cexpr(self, c, a: int) -> str:
        if c[0] == "I":
            k = c[1]
            return f"({self.AX_VAR[a]}{k:+d})" if k else self.AX_VAR[a]
        if c[0] == "A":
            return str(c[1])
        return c[1]

    def shift(self, pos, a: int, d: int):
        c = pos[a]
        if c[0] in ("I", "A"):
            nc = (c[0], c[1] + d)
        else:
            e = f"({c[1]}{d:+d})"
            nc = ("E", self._cvar(e))
        return tuple(nc if i == a else pos[i] for i in range(3))

    def _cvar(self, e: str) -> str:
        if e in self.cmemo:
            return self.cmemo[e]
        name = self.var(e, "int")
        self.cmemo[e] = name
        return name

    def transform(self, pos, a: int, fn: str):
        c = pos[a]
        if c[0] == "I":
            return pos
        if c[0] == "A":
            nc = ("A", _ct_index(fn, c[1], self.g.grid[a]))
        else:
            nc = ("E", self._cvar(f"{fn}({c[1]}, {self.AX_N[a]})"))
        return tuple(nc if i == a else pos[i] for i in range(3))

    def inside(self, pos, a: int) -> str:
        c = pos[a]
        if c[0] == "I":
            return "true"
        if c[0] == "A":
            return "true" if 0 <= c[1] < self.g.grid[a] else "false"
        return f"inside({c[1]}, {self.AX_N[a]})"

    def at(self, pos, a: int, which: str) -> str:
        c = pos[a]
        if c[0] == "I":
            return "false"
        if c[0] == "A":
            hit = c[1] == 0 if which == "lo" else c[1] == self.g.grid[a] - 1
            return "true" if hit else "false"
        return f"({c[1]} == 0)" if which == "lo" else f"({c[1]} == {self.AX_N[a]} - 1)"

    # -- taps --------------------------------------------------------------------------
    def _is_origin(self, c, a: int) -> bool:
        if self.mode[0] == "xface" and a == 2:
            return c == ("A", self.face_x())
        return c == ("I", 0) or (c[0] == "E" and c[1] == self.AX_VAR[a])

    def tap(self, si: int, ch: int, pos) -> str:
        key = ("tap", si, ch, pos)
        if key in self.memo:
            return self.memo[key][0]
        g = self.g
        src = g.src_layout[si]
        if self.mode[0] == "gather":
            cz = self._cvar(f"clampc({self.cexpr(pos[0], 0)}, N0)")
            cy = self._cvar(f"clampc({self.cexpr(pos[1], 1)}, N1)")
            cx = self._cvar(f"clampc({self.cexpr(pos[2], 2)}, N2)")
            idx = f"((((IDX){cz} * N1 + {cy}) * N2 + {cx}) * {src['C']} + {ch})"
            v = self.var(f"p.src{si}[{idx}]")
            self.memo[key] = [v]
            return v
        kz = pos[0][1]
        if self.record is not None and all(c[0] == "I" for c in pos):
            self.record.setdefault(si, set()).add((pos[0][1], pos[1][1], pos[2][1]))
        own = self._is_origin(pos[1], 1) and self._is_origin(pos[2], 2)
        if src["mode"] == "queue":
            if own and kz in (-1, 0, 1):
                v = {-1: f"qm{si}[{ch}]", 0: f"q0{si}[{ch}]", 1: f"qp{si}[{ch}]"}[kz]
                self.memo[key] = [v]
                return v
            if kz != 0 or not src["lds"]:
                raise _NeedRing(si)
            slot = "sq"
        else:
            slot = f"((unsigned)(s{kz:+d}) % {src['D']}u)" if kz else f"((unsigned)s % {src['D']}u)"
        rows = []
        for a, (lvar, t0, T, N) in ((1, ("ly", "y0", "TY", "N1")), (2, ("lx", "x0", "TX", "N2"))):
            H = src["HY"] if a == 1 else src["H"]
            c = pos[a]
            if c[0] == "I":
                if abs(c[1]) > H:
                    raise RuntimeError(f"LDS tap {c[1]:+d} on axis {a} outside the halo of {H}")
                rows.append(f"({lvar}+{H + c[1]})")
            elif c[0] == "A":
                # the x-face tile starts at 0 (low face) or N2 - TX (high face); same mapping
                # as lmap, evaluated here
                tx, n2 = g.sched.tx, g.grid[2]
                l = c[1] - (0 if self.mode[1] == "lo" else n2 - tx)
                if l >= tx + H:
                    l -= n2
                if l < -H:
                    l += n2
                l = min(max(l, -H), tx + H - 1)
                rows.append(str(l + H))
            else:
                rows.append(self._cvar(f"lmap({c[1]}, {t0}, {T}, {H}, {N})"))
        v = self.var(f"lds_s{si}[{slot}][{ch}][{rows[0]}][{rows[1]}]")
        self.memo[key] = [v]
        return v

    # -- values ------------------------------------------------------------------------
    def val(self, nid, pos) -> List[str]:
        key = (nid, pos)
        if key in self.memo:
            return self.memo[key]
        res = self._val(nid, pos)
        self.memo[key] = res
        return res

    def _val(self, nid, pos) -> List[str]:
        g = self.g
        if isinstance(nid, tuple):      # virtual nodes
            return self._virtual(nid, pos)
        if nid in g.xform_node and self.mode[0] != "gather":
            return [self.tap(g.xform_node[nid], 0, pos)]       # LDS holds the transformed value
        if nid in g.src_index:
            si = g.src_index[nid]
            if si in g.xform and self.mode[0] != "gather":
                raise RuntimeError("raw tap of a load-transformed source")
            return [self.tap(si, c, pos) for c in range(g.src_layout[si]["C"])]
        inf = g.info[nid]
        if inf.kind == "const":
            return [_flt(inf.value)]
        if nid in g.dsc_index:
            return [f"sc{g.dsc_index[nid]}"]
        node = g.graph.node(nid)
        op, p = node.op_name, node.params
        ins = node.input_ids
        if op in UNARY_OPS:
            a = self.val(ins[0], pos)
            return [self.var(self._unary(op, x, p)) for x in a]
        if op in BINARY_OPS:
            if op == "mul" and len(ins) == 1:
                a = self.val(ins[0], pos)
                s = _flt(p["scalar"])
                return [self.var(f"{x} * {s}") for x in a]
            a, b = self.val(ins[0], pos), self.val(ins[1], pos)
            n = max(len(a), len(b))
            a = a * n if len(a) == 1 else a
            b = b * n if len(b) == 1 else b
            return [self.var(self._binary(op, x, y, p)) for x, y in zip(a, b)]
        if op == "dot":
            u, w = self.val(ins[0], pos), self.val(ins[1], pos)
            acc = f"{u[0]} * {w[0]}"
            for c in range(1, len(u)):
                acc = f"({acc}) + {u[c]} * {w[c]}"
            return [self.var(acc)]
        if op == "sqnorm":
            u = self.val(ins[0], pos)
            acc = f"{u[0]} * {u[0]}"
            for c in range(1, len(u)):
                acc = f"({acc}) + {u[c]} * {u[c]}"
            return [self.var(acc)]
        if op == "select_component":
            u = self.val(ins[0], pos)
            return [u[int(p["index"]) % len(u)]]
        if op in STENCIL_OPS:
            return self._stencil(node, pos)
        if op in TENSOR_OPS:
            return self._tensor(op, self.val(ins[0], pos))
        raise KeyError(f"cannot fuse {op}")

    @staticmethod
    def _unary(op: str, x: str, p: Dict) -> str:
        if op == "neg":
            return f"-{x}"
        # transcendentals round an FP64 value once (FP32 library versions are 1-2 ulp, the
        # reference CPU path is near correctly rounded), through sp_math32.h's short FP64
        # evaluations (0.5 ulp + 2^-16)
        if op == "exp":
            return f"spm_expf({x})"
        if op == "log":
            e = _flt(p.get("eps", 1e-15))
            return f"spm_logf({x} < {e} ? {e} : {x})"
        if op == "tanh":
            return f"spm_tanhf({x})"
        if op == "sqrt":
            e = _flt(p.get("eps", 0.0))
            return f"sqrtf({x} < {e} ? {e} : {x})"
        if op in ("sin", "cos"):
            return f"spm_{op}f({x})"
        if op == "clamp":
            lo, hi = _flt(p.get("lo", 0.0)), _flt(p.get("hi", 1.0))
            return f"({x} < {lo} ? {lo} : ({x} > {hi} ? {hi} : {x}))"
        return x  # cast_bf16: value unchanged, store narrows

    @staticmethod
    def _binary(op: str, a: str, b: str, p: Dict) -> str:
        if op == "add":
            return f"{a} + {b}"
        if op == "sub":
            return f"{a} - {b}"
        if op in ("mul", "velocity_dot_grad"):
            return f"{a} * {b}"
        e = _flt(p.get("eps", 1e-15))
        return f"{a} / (fabsf({b}) < {e} ? ({b} < 0.0f ? -{e} : {e}) : {b})"

    # -- derivatives ---------------------------------------------------------------------
    def d1(self, inner, ch: int, pos, a: int, h: float, order: int) -> str:
        """torch.gradient derivative of channel ``ch`` of ``inner`` along axis ``a`` at pos."""
        key = ("d1", inner, ch, pos, a, h, order)
        if key in self.memo:
            return self.memo[key][0]
        hs = _flt(h)
        if pos[a][0] == "A" and pos[a][1] in (0, self.g.grid[a] - 1):
            # compile-time face coordinate: only the one-sided formula
            v0 = self.val(inner, pos)[ch]
            lo = pos[a][1] == 0
            s1, s2 = (1, 2) if lo else (-1, -2)
            v1 = self.val(inner, self.shift(pos, a, s1))[ch]
            if order == 2:
                v2 = self.val(inner, self.shift(pos, a, s2))[ch]
                e = f"tg_left2({v0}, {v1}, {v2}, {hs})" if lo else f"tg_right2({v2}, {v1}, {v0}, {hs})"
            else:
                e = f"tg_left1({v0}, {v1}, {hs})" if lo else f"tg_right1({v1}, {v0}, {hs})"
            r = self.var(e)
            self.memo[key] = [r]
            return r
        vp = self.val(inner, self.shift(pos, a, +1))[ch]
        vm = self.val(inner, self.shift(pos, a, -1))[ch]
        mid = f"tg_mid({vp}, {vm}, {hs})"
        if pos[a][0] in ("I", "A"):
            r = self.var(mid)
        else:
            v0 = self.val(inner, pos)[ch]
            if order == 2:
                vp2 = self.val(inner, self.shift(pos, a, +2))[ch]
                vm2 = self.val(inner, self.shift(pos, a, -2))[ch]
                left = f"tg_left2({v0}, {vp}, {vp2}, {hs})"
                right = f"tg_right2({vm2}, {vm}, {v0}, {hs})"
            else:
                left = f"tg_left1({v0}, {vp}, {hs})"
                right = f"tg_right1({vm}, {v0}, {hs})"
            r = self.var(f"{self.at(pos, a, 'lo')} ? {left} : ({self.at(pos, a, 'hi')} ? {right} : {mid})")
        self.memo[key] = [r]
        return r

    def roll_d1(self, inner, ch: int, pos, a: int, h: float) -> str:
        """Periodic central difference (x[i+1] - x[i-1]) / (2h) via wrapped neighbours."""
        pp = self.transform(self.shift(pos, a, +1), a, "wrapc")
        pm = self.transform(self.shift(pos, a, -1), a, "wrapc")
        vp, vm = self.val(inner, pp)[ch], self.val(inner, pm)[ch]
        return self.var(f"({vp} - {vm}) / {_flt(2.0 * h)}")

    def _order_nd(self) -> int:
        return 2 if min(self.g.grid) >= 3 else 1

    def _stencil(self, node: GraphNode, pos) -> List[str]:
        op, p, ins = node.op_name, node.params, node.input_ids
        bnd = p.get("boundary", "neumann")
        periodic = bnd == "periodic"
        dxs = _dx3(p.get("dx", 1.0))
        grid = self.g.grid
        if op == "laplacian":
            c = self.val(ins[0], pos)[0]
            terms = []
            for a in range(3):
                h2 = _flt(float(np.float32(dxs[a]) * np.float32(dxs[a])))
                pp, pm = self.shift(pos, a, +1), self.shift(pos, a, -1)
                if periodic:
                    xp = self.val(ins[0], self.transform(pp, a, "wrapc"))[0]
                    xm = self.val(ins[0], self.transform(pm, a, "wrapc"))[0]
                elif bnd == "neumann":
                    xp = self.val(ins[0], self.transform(pp, a, "clampc"))[0]
                    xm = self.val(ins[0], self.transform(pm, a, "clampc"))[0]
                elif bnd == "reflect" and grid[a] > 1:
                    xp = self.val(ins[0], self.transform(pp, a, "reflc"))[0]
                    xm = self.val(ins[0], self.transform(pm, a, "reflc"))[0]
                else:
                    vp, vm = self.val(ins[0], pp)[0], self.val(ins[0], pm)[0]
                    xp = self.var(f"{self.inside(pp, a)} ? {vp} : 0.0f")
                    xm = self.var(f"{self.inside(pm, a)} ? {vm} : 0.0f")
                terms.append(f"((({xp} + {xm}) - 2.0f * {c}) / {h2})")
            return [self.var(f"({terms[0]} + {terms[1]}) + {terms[2]}")]
        if op == "gradient":
            h = float(p.get("dx", 1.0))
            axes = [int(p["dim"]) % 3] if p.get("dim") is not None else [0, 1, 2]
            out = []
            for a in axes:
                if periodic:
                    out.append(self.roll_d1(ins[0], 0, pos, a, h))
                else:
                    out.append(self.d1(ins[0], 0, pos, a, h, 2 if grid[a] >= 3 else 1))
            return out
        if op == "divergence":
            terms = []
            for a in range(3):
                if periodic:
                    terms.append(self.roll_d1(ins[0], a, pos, a, dxs[a]))
                else:
                    terms.append(self.d1(ins[0], a, pos, a, dxs[a], 2))
            return [self.var(f"({terms[0]} + {terms[1]}) + {terms[2]}")]
        if op == "curl":
            d = lambda a, c: self.d1(ins[0], c, pos, a, dxs[a], 2)
            return [self.var(f"{d(1, 2)} - {d(2, 1)}"), self.var(f"{d(2, 0)} - {d(0, 2)}"),
                    self.var(f"{d(0, 1)} - {d(1, 0)}")]
        if op == "gradient_nd":
            o = self._order_nd()
            return [self.d1(ins[0], 0, pos, a, dxs[a], o) for a in range(3)]
        if op == "jacobian":
            o = self._order_nd()
            return [self.d1(ins[0], c, pos, a, dxs[a], o) for c in range(3) for a in range(3)]
        if op == "hessian":
            o = self._order_nd()
            out = []
            for i in range(3):
                gi = ("d1v", ins[0], 0, i, dxs[i], o)
                for j in range(3):
                    out.append(self.d1(gi, 0, pos, j, dxs[j], o))
            return out
        if op == "mean_curvature":
            eps = float(p.get("eps", 1e-15))
            nv = ("nrm", ins[0], dxs, self._order_nd(), eps)
            terms = []
            for a in range(3):
                if periodic:
                    terms.append(self.roll_d1(nv, a, pos, a, dxs[a]))
                else:
                    terms.append(self.d1(nv, a, pos, a, dxs[a], 2))
            return [self.var(f"({terms[0]} + {terms[1]}) + {terms[2]}")]
        if op == "material_derivative":
            o = self._order_nd()
            v = self.val(ins[1], pos)
            gs = [self.d1(ins[0], 0, pos, a, dxs[a], o) for a in range(3)]
            acc = f"{v[0]} * {gs[0]}"
            for a in (1, 2):
                acc = f"({acc}) + {v[a]} * {gs[a]}"
            return [self.var(acc)]
        raise KeyError(op)

    def _virtual(self, vid, pos) -> List[str]:
        kind = vid[0]
        if kind == "d1v":               # one torch.gradient derivative as a value
            _, inner, ch, a, h, o = vid
            return [self.d1(inner, ch, pos, a, h, o)]
        if kind == "nrm":               # unit normal of gradient_nd
            _, inner, dxs, o, eps = vid
            gs = [self.d1(inner, 0, pos, a, dxs[a], o) for a in range(3)]
            # g / max(|g|, eps): the FP32 reciprocal norm of the FP32 sum of squares (hardware
            # reciprocal square root and one Newton step) times each component, instead of a
            # rounded square root and a division per component
            r = self.var(f"unit_rnorm(({gs[0]} * {gs[0]} + {gs[1]} * {gs[1]}) + {gs[2]} * {gs[2]}, {_flt(eps)})")
            return [self.var(f"{x} * {r}") for x in gs]
        raise KeyError(kind)

    def _tensor(self, op: str, m: List[str]) -> List[str]:
        arr = f"t{self.n}"
        self.n += 1
        self.lines.append(f"const float {arr}[9] = {{{', '.join(m)}}};")
        if op == "trace":
            return [self.var(f"sym3_trace({arr})")]
        if op == "determinant":
            return [self.var(f"sym3_det_regularised({arr}, A.detreg)")]
        if op == "eigenvalues":
            e = f"e{self.n}"
            self.n += 1
            self.lines.append(f"float {e}[3]; sym3_eigvalsh({arr}, {e});")
            return [self.var(f"{e}[{i}]") for i in range(3)]
        out = f"o{self.n}"
        self.n += 1
        fn = "sym3_deviatoric" if op == "deviatoric" else "mat3_inverse"
        self.lines.append(f"float {out}[9]; {fn}({arr}, {out});")
        return [self.var(f"{out}[{i}]") for i in range(9)]


def _op_reach_kind(node: GraphNode) -> List[str]:
    """Composition of reach steps an op applies to its (first) input, innermost first."""
    op, p = node.op_name, node.params
    periodic = p.get("boundary", "neumann") == "periodic"
    if op == "laplacian":
        return ["wrap" if periodic else "pad"]
    if op in ("gradient", "divergence"):
        return ["wrap" if periodic else "onesided"]
    if op in ("curl", "gradient_nd", "jacobian", "material_derivative"):
        return ["onesided"]
    if op == "hessian":
        return ["onesided", "onesided"]
    if op == "mean_curvature":
        return ["onesided", "wrap" if periodic else "onesided"]
    return []


def source_reach(graph: Graph, group: Group) -> Dict[str, Tuple[int, int, int]]:
    """Per-source (interior reach, edge reach, LDS halo) of a fused group.

    interior reach: cells reached by interior formulas (sizes the plane ring);
    edge reach: cells reached when one-sided edge formulas fire (2 per torch.gradient level;
    selects the boundary variants); halo: in-plane LDS halo -- equal to the interior reach
    except where a periodic op wraps over a one-sided op, which then reads the far edge's
    inward taps from the near side's halo.
    """
    members = set(group.nodes)
    sources = set(group.sources)
    memo: Dict[str, Dict[str, Tuple[int, int, int]]] = {}

    def walk(nid: str) -> Dict[str, Tuple[int, int, int]]:
        if nid in memo:
            return memo[nid]
        if nid in sources:
            res = {nid: (0, 0, 0)}
        elif nid not in members:
            res = {}
        else:
            node = graph.node(nid)
            res = {}
            steps = _op_reach_kind(node)
            for k, i in enumerate(node.input_ids):
                if graph.node(i).op_name in REDUCTION_OPS and i not in sources:
                    continue
                sub = walk(i)
                # material_derivative differentiates only its first input
                my_steps = steps if not (node.op_name == "material_derivative" and k == 1) else []
                for src, (r, e, h) in sub.items():
                    for st in my_steps:
                        if st == "wrap":
                            r, e, h = r + 1, e + 1, max(h, e) + 1
                        elif st == "onesided":
                            r, e, h = r + 1, e + 2, h + 1
                        else:
                            r, e, h = r + 1, e + 1, h + 1
                    old = res.get(src, (0, 0, 0))
                    res[src] = (max(old[0], r), max(old[1], e), max(old[2], h))
        memo[nid] = res
        return res

    out: Dict[str, Tuple[int, int, int]] = {s: (0, 0, 0) for s in group.sources}
    for nid in group.nodes:
        for src, v in walk(nid).items():
            o = out[src]
            out[src] = (max(o[0], v[0]), max(o[1], v[1]), max(o[2], v[2]))
    return out


def edge_band(graph: Graph, group: Group) -> int:
    """Smallest distance from a face at which interior formulas are exact for the group.

    The LDS kernel evaluates every op with its interior formula on wrapped reads (halo and
    plane loads wrap around), which is exact for periodic ops at any distance.  A one-sided
    or padded op differs only on the face layer itself (distance 0) and needs its input exact
    within its reach; a periodic op over such an input needs that input's band plus its reach.
    Voxels closer to a face than the band are evaluated by the gather path.
    """
    members = set(group.nodes)
    sources = set(group.sources)
    memo: Dict[str, int] = {}

    def need(nid: str) -> int:
        if nid in memo:
            return memo[nid]
        if nid in sources or nid not in members:
            memo[nid] = 0
            return 0
        node = graph.node(nid)
        steps = _op_reach_kind(node)
        out = 0
        for k, i in enumerate(node.input_ids):
            if graph.node(i).op_name in REDUCTION_OPS and i not in sources:
                continue
            n = need(i)
            my_steps = steps if not (node.op_name == "material_derivative" and k == 1) else []
            for st in my_steps:
                if st == "wrap":
                    n = n + 1 if n > 0 else 0
                else:                      # one-sided torch.gradient edges, padded laplacian
                    n = max(1, n + 1)
            out = max(out, n)
        memo[nid] = out
        return out

    return max([need(n) for n in group.nodes] + [0])


class KernelGen:
    """Builds the HIP source of one fused group for one schedule.

    Per field source the generator picks a mode from the taps the interior code makes:
      * "queue": every tap off the current plane sits on the voxel's own column (and the
        interior reach is 1) -- the z-neighbours live in a per-thread register queue and LDS
        holds only the current plane (two slots: current + next); centre taps need no LDS;
      * "ring": an LDS ring of 2*hz+2 planes (mixed z/in-plane taps, e.g. hessians).
    Each thread owns V voxels per plane (column tid % TX, rows tid / TX + k * NT / TX); it
    loads its own cells plus a share of the halo, with plane-relative addresses computed
    once per block, into one of two register buffers.  The z-loop is unrolled by two so
    that each buffer's loads have two iterations to land.
    """

    def __init__(self, graph: Graph, group: Group, info: Dict[str, Info], sched: Schedule,
                 out_dtypes: Dict[str, torch.dtype], src_meta: Dict[str, Tuple[int, ...]],
                 edge_mode: Optional[str] = None, batched: bool = False):
        self.graph, self.group, self.info, self.sched = graph, group, info, sched
        self.grid = group.grid
        self.out_dtypes = out_dtypes
        self.src_index = {s: i for i, s in enumerate(group.sources)}
        self.dsc_index = {s: i for i, s in enumerate(group.dscalars)}
        reach = source_reach(graph, group)
        self.hz = max([r[0] for r in reach.values()] + [0])
        self.redge = edge_band(graph, group)
        self.use_lds = self.hz > 0 and len(group.sources) > 0
        if sched.threads % sched.tx != 0 or sched.ty % (sched.threads // sched.tx) != 0:
            raise ValueError(f"schedule {sched} does not map threads onto whole tile rows")
        self.rp = sched.threads // sched.tx          # rows per pass
        self.v = sched.ty // self.rp                 # voxels per thread per plane
        self.src_layout = [{"C": info[s].ch, "H": reach[s][2], "HY": reach[s][2], "D": 2 * self.hz + 2,
                            "mode": "ring", "lds": True} for s in group.sources]
        self.xform = self._load_transforms()
        self.xform_node = {v[1]: si for si, v in self.xform.items()}
        # batched: one code for many grids (shapes are run-time values, see section())
        self.batched = batched
        if batched:
            if min(self.grid) < 2 * self.redge + 1 or min(self.grid) < 3:
                raise ValueError("batched groups need every axis >= max(3, 2 * band + 1)")
            if self._det_inputs():
                raise ValueError("determinant groups are not batched (their relaunch is per grid)")
            edge_mode = "band"
        self.edge_mode = self._pick_edge_mode(edge_mode or EDGE_MODE)
        self._choose_modes()
        # In band mode every LDS evaluation of an in-plane tap sits at an interior row (the y
        # faces go to the gather blocks), so the rows above and below the tile need only the
        # interior reach; the columns keep the full halo, which the face wave's wrapped
        # one-sided x formulas read (a periodic op over a one-sided one widens it).
        for s_, L_ in zip(group.sources, self.src_layout):
            L_["HY"] = min(L_["H"], reach[s_][0]) if self.edge_mode == "band" else L_["H"]
        # outputs of >= 4 channels written through a per-wave LDS transpose (coalesced lines)
        self.stage = [oi for oi, o in enumerate(group.outputs)
                      if STAGE_STORES and self.use_lds and not batched and (sched.tx % 64 == 0 or 64 % sched.tx == 0)
                      and info[o].ch >= STAGE_MIN_CH and out_dtypes.get(o) != torch.bfloat16]
        # per wave: 64 * C floats per staged output, deposited by the lanes inside the variant
        self.stage_off: Dict[int, int] = {}
        self.stage_c = 0
        for oi in self.stage:
            self.stage_off[oi] = 64 * self.stage_c
            self.stage_c += info[group.outputs[oi]].ch

    _COSTLY = {"exp", "log", "sqrt", "sin", "cos", "tanh"}

    def _load_transforms(self) -> Dict[int, Tuple[List[str], str]]:
        """Scalar sources used only through a point-wise chain with a transcendental that feeds
        a stencil (e.g. f in laplacian(tanh(f))): {source index: (chain, last node)}.

        The chain is applied once per loaded cell (own and halo) before the LDS store, so the
        stencil's taps read the transformed value instead of re-evaluating the chain -- in
        FP64 for transcendentals -- at every tap.  Chain steps: unary ops, scaling, and binary
        ops with a constant operand; each intermediate node has the next step as its only
        consumer and is not an output."""
        g, graph = self.group, self.graph
        members = set(g.nodes)
        cons: Dict[str, List[str]] = {}
        for nid in g.nodes:
            for i in graph.node(nid).input_ids:
                if nid not in cons.setdefault(i, []):
                    cons[i].append(nid)
        ends = set(g.outputs) | {graph.node(r).input_ids[0] for r in g.reductions} | set(self._det_inputs())
        out: Dict[int, Tuple[List[str], str]] = {}
        for si, sid in enumerate(g.sources):
            if self.info[sid].ch != 1:
                continue
            chain, cur = [], sid
            while True:
                cs = cons.get(cur, [])
                if len(cs) != 1 or (cur != sid and cur in ends):
                    break
                node = graph.node(cs[0])
                op, ins = node.op_name, node.input_ids
                if op in UNARY_OPS and op != "cast_bf16":
                    ok = True
                elif op == "mul" and len(ins) == 1:
                    ok = True
                elif op in ("add", "sub", "mul", "div") and len(ins) == 2:
                    other = ins[1] if ins[0] == cur else ins[0]
                    ok = ins.count(cur) == 2 or self.info[other].kind == "const"
                else:
                    ok = False
                if not ok or cs[0] not in members:
                    break
                chain.append(cs[0])
                cur = cs[0]
            if not chain or not any(graph.node(n).op_name in self._COSTLY for n in chain):
                continue
            last = chain[-1]
            if not any(graph.node(c).op_name in STENCIL_OPS for c in cons.get(last, [])):
                continue
            out[si] = (chain, last)
        return out

    def _xform_functions(self) -> List[str]:
        """``xf<i>(x)``: the load transform of source i (see _load_transforms)."""
        fns = []
        for si, (chain, last) in self.xform.items():
            lines, cur, prev = [], "x", self.group.sources[si]
            for k, nid in enumerate(chain):
                node = self.graph.node(nid)
                op, p_, ins = node.op_name, node.params, node.input_ids
                if op in UNARY_OPS:
                    e = _Emit._unary(op, cur, p_)
                elif len(ins) == 1:
                    e = f"{cur} * {_flt(p_['scalar'])}"
                elif ins[0] == ins[1]:                       # x op x (e.g. f*f)
                    e = _Emit._binary(op, cur, cur, p_)
                else:
                    cval = _flt(self.info[ins[1] if ins[0] == prev else ins[0]].value)
                    e = _Emit._binary(op, cur, cval, p_) if ins[0] == prev else _Emit._binary(op, cval, cur, p_)
                lines.append(f"    const float t{k} = {e};")
                cur, prev = f"t{k}", nid
            fns.append(f"__device__ __forceinline__ float xf{si}(const float x) {{\n" + "\n".join(lines) +
                       f"\n    return {cur};\n}}\n")
        return fns

    def _pick_edge_mode(self, mode: str) -> str:
        if mode in ("band", "variants"):
            return mode
        n0, n1, n2 = self.grid
        r = self.redge
        if n1 <= 2 * r or n2 <= 2 * r:
            return "variants"
        frac = 1.0 - (n1 - 2 * r) * (n2 - 2 * r) / float(n1 * n2)
        return "band" if frac <= BAND_MAX_FRACTION else "variants"

    def interior_bounds(self, grid=None) -> Tuple[int, int, int, int, int, int]:
        """Box (z0, z1, y0, y1, x0, x1) of voxels the main kernel evaluates."""
        n0, n1, n2 = grid or self.grid
        r, s = self.redge, self.sched
        xp = (n2 // s.tx) * s.tx          # start of a partial x tile (== n2 if none)
        yp = (n1 // s.ty) * s.ty
        zlo = min(r, n0)
        zhi = max(n0 - r, zlo)
        if self.edge_mode == "band":
            # partial tiles included: own and halo loads wrap, so cells past the grid edge
            # hold the periodic images and interior formulas stay exact up to the band
            ylo, xlo = min(r, n1), min(r, n2)
            yhi, xhi = max(n1 - r, ylo), max(n2 - r, xlo)
        else:
            ylo, yhi, xlo, xhi = 0, yp, 0, xp
        return zlo, zhi, ylo, yhi, xlo, xhi

    # -- source modes ----------------------------------------------------------------------
    def _choose_modes(self) -> None:
        if not self.use_lds:
            return
        rec: Dict[int, set] = {}
        e = _Emit(self, ("lds", False, False), record=rec)
        self._walk(e)
        for si, L_ in enumerate(self.src_layout):
            taps = rec.get(si, set())
            zoff = all(ky == 0 and kx == 0 for kz, ky, kx in taps if kz != 0)
            inplane = any((ky, kx) != (0, 0) for kz, ky, kx in taps)
            if self.hz == 1 and zoff:
                L_.update(mode="queue", D=2, lds=inplane)
                if not inplane:
                    L_["H"] = L_["HY"] = 0
        # the edge variants may still need a ring (e.g. a wrapped coordinate off the column)
        while True:
            try:
                for m in self._variants():
                    self._walk(_Emit(self, m))
                break
            except _NeedRing as nr:
                L_ = self.src_layout[nr.si]
                L_.update(mode="ring", D=2 * self.hz + 2, lds=True)

    def _memo(self, name: str, fn):
        """Value of a geometry query on the generator's own grid, computed once: the grid,
        schedule, edge mode and band are fixed at construction, and a run asks several times."""
        memo = self.__dict__.setdefault("_geom_memo", {})
        if name not in memo:
            memo[name] = fn()
        return memo[name]

    def xface_sides(self) -> List[str]:
        """x faces whose band the main kernel evaluates with compile-time face variants.

        A face voxel sits in its own cache line per row and plane, so leaving the x faces to
        the gather path re-reads whole lines long after the LDS kernel fetched them."""
        return self._memo("xface_sides", self._xface_sides)

    def _xface_sides(self) -> List[str]:
        if not self.use_lds or self.edge_mode != "band" or self.redge == 0 or self.batched:
            return []
        n2, tx, r = self.grid[2], self.sched.tx, self.redge
        if n2 < tx or n2 < 2 * r + 1 or r > tx:
            return []
        # one extra wave per block evaluates them: RP rows x r columns x 2 faces must fit
        if self.rp * r * 2 > 64 or self.sched.threads + 64 > 1024:
            return []
        sides = ["lo"]
        if n2 % tx == 0:
            sides.append("hi")
        return sides

    @property
    def fw(self) -> int:
        """Threads of the face wave (0 when the x faces are left to the tail blocks)."""
        return 64 if self.xface_sides() else 0

    def block_threads(self) -> int:
        return self.sched.threads + self.fw

    def _variants(self):
        v = [("gather",)]
        if self.use_lds:
            if self.edge_mode == "band":
                v.append(("lds", False, False))
                v += [("xface", side, d) for side in self.xface_sides() for d in range(self.redge)]
            else:
                v += [("lds", ex, ey) for ex in (False, True) for ey in (False, True)]
        return v

    # -- boundary kernel -------------------------------------------------------------------------
    def boundary_boxes(self, grid=None) -> List[Tuple[int, int, int, int, int, int]]:
        """Disjoint boxes (z0, z1, y0, y1, x0, x1) the main kernel leaves out."""
        if grid is None:
            return self._memo("boundary_boxes", lambda: self._boundary_boxes(None))
        return self._boundary_boxes(grid)

    def _boundary_boxes(self, grid) -> List[Tuple[int, int, int, int, int, int]]:
        if not self.use_lds:
            return []
        n0, n1, n2 = grid or self.grid
        zlo, zhi, ylo, yhi, xlo, xhi = self.interior_bounds(grid)
        sides = self.xface_sides()
        boxes = [(0, zlo, 0, n1, 0, n2), (zhi, n0, 0, n1, 0, n2),
                 (zlo, zhi, 0, ylo, 0, n2), (zlo, zhi, yhi, n1, 0, n2)]
        if "lo" not in sides:
            boxes.append((zlo, zhi, ylo, yhi, 0, xlo))
        if "hi" not in sides:
            boxes.append((zlo, zhi, ylo, yhi, xhi, n2))
        return [b for b in boxes if b[1] > b[0] and b[3] > b[2] and b[5] > b[4]]

    TAIL_VPT = 4     # voxels per thread of a tail block

    def boundary_blocks(self, grid=None) -> int:
        if grid is None:
            return self._memo("boundary_blocks", lambda: self._boundary_blocks(None))
        return self._boundary_blocks(grid)

    def _boundary_blocks(self, grid) -> int:
        tot = sum((b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4]) for b in self.boundary_boxes(grid))
        per = self.block_threads() * self.TAIL_VPT
        return (tot + per - 1) // per

    @property
    def split_tail(self) -> bool:
        if not (self.use_lds and not self.batched and self.boundary_blocks() > 0):
            return False
        want = getattr(self, "want_split", None)
        return want if want is not None else SPLIT_TAIL == "1"

    def _boundary_part(self, A, ndet: int) -> None:
        """Tail blocks (block index b >= p.nblocks): gather evaluation of the boundary boxes."""
        g = self.group
        boxes = self.boundary_boxes()
        A("    if (b >= nb) {")
        A(f"      for (int it_ = 0; it_ < {self.TAIL_VPT}; ++it_) {{")
        A(f"        const i64 t = ((i64)(b - nb) * {self.TAIL_VPT} + it_) * NTB + tid;")
        A("        int z = -1, y = 0, x = 0;")
        if self.batched:
            # boxes of this grid, from the parameters: bstart = prefix sums of box sizes
            A("        #pragma unroll")
            A("        for (int k = 0; k < 6; ++k) {")
            A("            if (z < 0 && t >= p.bstart[k] && t < p.bstart[k + 1]) {")
            A("                const int* bx = p.boxes + 6 * k;")
            A("                const i64 q = t - p.bstart[k];")
            A("                const int ny = bx[3] - bx[2], nx = bx[5] - bx[4];")
            A("                x = bx[4] + (int)(q % nx); y = bx[2] + (int)((q / nx) % ny); z = bx[0] + (int)(q / ((i64)nx * ny));")
            A("            }")
            A("        }")
            boxes = []
        off = 0
        for (z0, z1, y0, y1, x0, x1) in boxes:
            ny, nx = y1 - y0, x1 - x0
            size = (z1 - z0) * ny * nx
            A(f"        if (z < 0 && t >= {off} && t < {off + size}) {{ const i64 q = t - {off}; "
              f"x = {x0} + (int)(q % {nx}); y = {y0} + (int)((q / {nx}) % {ny}); z = {z0} + (int)(q / {nx * ny}); }}")
            off += size
        qa = "".join(", nullptr, nullptr, nullptr" for _ in self.queue_sources())
        A("        if (z >= 0) {")
        A("            const IDX vox = ((IDX)z * N1 + y) * N2 + x;")
        A(f"            {self._variant_name(('gather',))}(p, A, 0, 0, z, y, x, 0, 0, 0, 0, vox{qa});")
        A("        }")
        A("      }")
        self._write_records(A, ndet, "        ")
        A("        return;")
        A("    }")

    def accumulators(self) -> Tuple[List[Tuple[str, str, Optional[int]]], Dict[int, List[int]]]:
        """Per-thread accumulators shared between the group's reductions, and which ones each
        reduction reads.  Kinds: "s1" compensated sum of x (sum, mean, variance), "s2"
        compensated sum of x^2 with the exact FMA split of each square (variance, norm_l2),
        "mm" min/max, "int" quadrature-weighted sum (one per integrate).  Sharing keeps the
        values each reduction's record receives unchanged while a reduction set such as
        sum + mean + variance + norm_l2 of one value costs two accumulators instead of six."""
        accs: List[Tuple[str, str, Optional[int]]] = []
        index: Dict[Tuple[str, str], int] = {}
        uses: Dict[int, List[int]] = {}

        def get(kind: str, src: str, ri: Optional[int] = None) -> int:
            key = (kind, src if kind != "int" else f"{src}#{ri}")
            if key not in index:
                index[key] = len(accs)
                accs.append((kind, src, ri))
            return index[key]

        for ri, nid in enumerate(self.group.reductions):
            node = self.graph.node(nid)
            src, op = node.input_ids[0], node.op_name
            if op == "variance":
                uses[ri] = [get("s1", src), get("s2", src)]
            elif op == "norm_l2":
                uses[ri] = [get("s2", src)]
            elif op == "min_max":
                uses[ri] = [get("mm", src)]
            elif op == "integrate":
                uses[ri] = [get("int", src, ri)]
            else:
                uses[ri] = [get("s1", src)]
        return accs, uses

    def _write_records(self, A, ndet: int, ind: str) -> None:
        flag = getattr(self, "_flag", False)
        if flag:
            # any thread that saw a small diagonal sets the device flag (all write the same value)
            cond = " || ".join(f"A.dflag{i} > 0.0f" for i in range(ndet))
            A(f"{ind}if ({cond}) __hip_atomic_store((double*)p.detflag, 1.0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);")
            return
        _, uses = self.accumulators()
        for ri, nid in enumerate(self.group.reductions):
            op = self.graph.node(nid).op_name
            rec = f"p.red{ri} + (i64)rb * 4"
            if op == "variance":
                k1, k2 = uses[ri]
                A(f"{ind}{{ RVar v_; v_.s1 = A.a{k1}; v_.s2 = A.a{k2}; block_var(v_, red_scratch, {rec}); }}")
            elif op == "integrate":
                dxs = _dx3(self.graph.node(nid).params.get("dx", 1.0))
                scale = float(dxs[0]) * float(dxs[1]) * float(dxs[2]) / 216.0
                A(f"{ind}block_sum(A.a{uses[ri][0]}, red_scratch, {rec}, {scale!r});")
            else:
                fn = "block_mm" if op == "min_max" else "block_sum"
                A(f"{ind}{fn}(A.a{uses[ri][0]}, red_scratch, {rec});")

    def _walk(self, e: "_Emit") -> None:
        pos = e.origin()
        for nid in self.group.outputs:
            e.val(nid, pos)
        for nid in self.group.reductions:
            e.val(self.graph.node(nid).input_ids[0], pos)
        for nid in self._det_inputs():
            e.val(nid, pos)

    def queue_sources(self) -> List[int]:
        return [i for i, L_ in enumerate(self.src_layout) if self.use_lds and L_["mode"] == "queue"]

    def lds_sources(self) -> List[int]:
        return [i for i, L_ in enumerate(self.src_layout) if self.use_lds and L_["lds"]]

    def lds_bytes(self) -> int:
        stage = (self.sched.threads // 64) * 64 * getattr(self, "stage_c", 0) * 4
        return self.lds_array_bytes() + stage + 4 * (self.block_threads() // 64) * 4 + 16

    # -- per-voxel code ------------------------------------------------------------------------
    def _emit_variant(self, mode) -> Tuple[str, str]:
        e = _Emit(self, mode)
        pos = e.origin()
        body = []
        flag = getattr(self, "_flag", False)
        staged = mode[0] == "lds"
        for oi, nid in enumerate([] if flag else self.group.outputs):
            vals = e.val(nid, pos)
            C = len(vals)
            bf = self.out_dtypes.get(nid) == torch.bfloat16
            if staged and oi in self.stage:
                # deposit into this lane's slot of the wave's staging area; the wave stores the
                # whole area with coalesced lines after the variant
                body.append(f"{{ float* st_ = &so_stage[threadIdx.x >> 6][{self.stage_off[oi]} + (threadIdx.x & 63) * {C}];")
                body += [f"  st_[{c}] = {v};" for c, v in enumerate(vals)]
                body += ["}", "A.vflag = 1;"]
                continue
            for c, v in enumerate(vals):
                st = "st_bf16" if bf else "st_f32"
                cast = "(unsigned short*)" if bf else "(float*)"
                body.append(f"{st}({cast}p.out{oi}, vox * {C} + {c}, {v});")
        accs = [] if flag else self.accumulators()[0]
        if accs:
            body.append("if (z >= p.zr0 && z < p.zr1) {")
        for k, (kind, src, ri) in enumerate(accs):
            vals = e.val(src, pos)
            if kind == "int":
                node = self.graph.node(self.group.reductions[ri])
                meth = 0 if node.params.get("method", "simpson") == "simpson" else 1
                # integer weight numerators (exact product); the FP64 record takes dx^3 / 216
                w = f"(quad_k(z, N0, {meth}) * quad_k(y, N1, {meth})) * quad_k(x, N2, {meth})"
                body.append(f"rsum_add_prod(A.a{k}, {vals[0]}, {w});")
            elif kind == "s2":
                for v in vals:
                    body.append(f"rsum_add_sq(A.a{k}, {v});")
            else:
                for v in vals:
                    body.append(f"A.a{k}.add({v});")
        if accs:
            body.append("}")
        for ki, nid in enumerate(self._det_inputs() if flag else []):
            vals = e.val(nid, pos)
            body.append(f"if (fabsf({vals[0]}) < 1e-12f || fabsf({vals[4]}) < 1e-12f || "
                        f"fabsf({vals[8]}) < 1e-12f) A.dflag{ki} = 1.0f;")
        name = self._variant_name(mode)
        code = "\n        ".join(e.lines + body)
        qargs = "".join(f", const float* qm{i}, const float* q0{i}, const float* qp{i}"
                        for i in self.queue_sources())
        sc = " ".join(f"const float sc{i} = A.sc{i};" for i in range(len(self.group.dscalars)))
        fn = (f"__device__ __forceinline__ void {name}(const Params& p, Acc& A, int s, int sq, int z,\n"
              f"        int y, int x, int ly, int lx, int y0, int x0, IDX vox{qargs}) {{\n"
              f"    {sc}\n        {code}\n}}\n")
        return name, fn

    def _variant_name(self, mode) -> str:
        sfx = "_flag" if getattr(self, "_flag", False) else ""
        if mode[0] == "gather":
            return "ev_gather" + sfx
        if mode[0] == "xface":
            return f"ev_x{mode[1]}{mode[2]}" + sfx
        return f"ev_lds{int(mode[1])}{int(mode[2])}" + sfx

    def main_boxes(self, grid=None) -> List[Tuple[int, int, int, int, int, int]]:
        """Boxes of voxels the main kernel evaluates (within full tiles)."""
        grid = grid or self.grid
        if not self.use_lds:
            return [(0, grid[0], 0, grid[1], 0, grid[2])]
        zlo, zhi, ylo, yhi, xlo, xhi = self.interior_bounds(grid)
        out = [(zlo, zhi, ylo, yhi, xlo, xhi)]
        n2 = grid[2]
        for side in self.xface_sides():
            out.append((zlo, zhi, ylo, yhi, 0, xlo) if side == "lo" else (zlo, zhi, ylo, yhi, n2 - self.redge, n2))
        return out

    def _det_inputs(self) -> List[str]:
        out = []
        for nid in self.group.nodes:
            node = self.graph.node(nid)
            if node.op_name == "determinant" and node.input_ids[0] not in out:
                out.append(node.input_ids[0])
        return out

    # macros a generated section defines (undefined again at the end of a batched section)
    _SECTION_MACROS = ("N0", "N1", "N2", "TX", "TY", "ZC", "NT", "HZ", "REDGE", "RP", "VPT", "FW", "NTB",
                       "XLO_IN", "XHI_IN", "YLO", "YHI", "XLO", "XHI")

    def source(self) -> str:
        if self.batched:
            raise ValueError("a batched generator emits sections (section()), not a module")
        return "\n".join(self._header_sources() + self._body())

    @staticmethod
    def _header_sources() -> List[str]:
        return [(CSRC / hdr).read_text().replace("#pragma once", "") for hdr in ("sp_sym3.h", "sp_math32.h", "fused_template.h")]

    def section(self, ns: str, idx64: bool = False) -> str:
        """The group's code inside ``namespace ns`` for a batch module: grid sizes and interior
        bounds are read from the per-grid ``Params`` and the LDS arrays are views of the
        module's dynamic shared buffer ``sp_smem``, so sections share one LDS allocation."""
        if not self.batched:
            raise ValueError("section() needs a batched generator")
        lines = [f"namespace {ns} {{"] + self._body(idx64) + ["} // namespace " + ns]
        lines += [f"#undef {m}" for m in self._SECTION_MACROS]
        lines += [f"#undef lds_s{i}" for i in self.lds_sources()]
        return "\n".join(lines)

    def lds_array_bytes(self) -> int:
        s = self.sched
        return sum(L_["D"] * L_["C"] * (s.ty + 2 * L_["HY"] + 1) * (s.tx + 2 * L_["H"]) * 4
                   for L_ in (self.src_layout[i] for i in self.lds_sources()))

    def _body(self, idx64: bool = False) -> List[str]:
        s, g = self.sched, self.group
        n0, n1, n2 = self.grid
        total = n0 * n1 * n2 * max([L_["C"] for L_ in self.src_layout] + [9])
        idx_t = "i64" if (idx64 or total >= 2 ** 31 - 1) else "int"
        ns, no, nr, nd = len(g.sources), len(g.outputs), len(g.reductions), len(g.dscalars)
        ndet = len(self._det_inputs())
        lines = []
        if self.batched:
            lines.append("#define N0 (p.n0)\n#define N1 (p.n1)\n#define N2 (p.n2)\n"
                         f"#define TX {s.tx}\n#define TY {s.ty}\n#define ZC {s.zc}\n#define NT {s.threads}\n"
                         f"#define HZ {self.hz}\n#define REDGE {self.redge}\n#define RP {self.rp}\n"
                         f"#define VPT {self.v}\ntypedef {idx_t} IDX;\n"
                         "#define FW 0\n#define NTB NT\n"
                         "#define XLO_IN 0\n#define XHI_IN 0\n"
                         "#define YLO (p.ylo)\n#define YHI (p.yhi)\n#define XLO (p.xlo)\n#define XHI (p.xhi)\n")
        else:
            lines.append(f"#define N0 {n0}\n#define N1 {n1}\n#define N2 {n2}\n"
                     f"#define TX {s.tx}\n#define TY {s.ty}\n#define ZC {s.zc}\n#define NT {s.threads}\n"
                     f"#define HZ {self.hz}\n#define REDGE {self.redge}\n#define RP {self.rp}\n"
                     f"#define VPT {self.v}\ntypedef {idx_t} IDX;\n"
                     f"#define FW {self.fw}\n#define NTB (NT + FW)\n"

                     f"#define XLO_IN {int('lo' in self.xface_sides())}\n#define XHI_IN {int('hi' in self.xface_sides())}\n")
        if self.use_lds and not self.batched:
            _, _, ylo, yhi, xlo, xhi = self.interior_bounds()
            lines.append(f"#define YLO {ylo}\n#define YHI {yhi}\n#define XLO {xlo}\n#define XHI {xhi}\n")
        pf = [f"    const float* __restrict__ src{i};" for i in range(ns)]
        pf += [f"    void* out{i};" for i in range(no)]
        pf += [f"    double* red{i};" for i in range(nr)]
        pf += [f"    double* dflag{i};" for i in range(ndet)]
        pf += [f"    const double* scal{i};" for i in range(nd)]
        if ndet:
            pf.append("    const double* detflag;      // 1.0: any |H_ii| < 1e-12 (from the flag pass)")
        pf.append("    int tiles_x, tiles_y, zchunks, nblocks;\n"
                  "    int zr0, zr1;               // reductions take the planes [zr0, zr1)")
        if self.batched:
            pf.append("    int n0, n1, n2, ylo, yhi, xlo, xhi, nbt;\n    int bstart[8];\n    int boxes[36];")
        lines.append("struct Params {\n" + "\n".join(pf) + "\n};\n")
        acc = []
        for k, (kind, _, _) in enumerate(self.accumulators()[0]):
            acc.append(f"    {'RMinMax' if kind == 'mm' else 'RSum'} a{k};")
        acc += [f"    float dflag{i};" for i in range(ndet)]
        acc += [f"    float sc{i};" for i in range(nd)]
        acc.append("    float detreg;")
        if self.stage:
            acc.append("    int vflag;")
        acc.append("    int dummy_;")
        lines.append("struct Acc {\n" + "\n".join(acc) + "\n};\n")
        off = 0
        for i in self.lds_sources():
            L_ = self.src_layout[i]
            dims = f"[{L_['D']}][{L_['C']}][TY + {2 * L_['HY'] + 1}][TX + {2 * L_['H']}]"
            if self.batched:
                lines.append(f"#define lds_s{i} (*reinterpret_cast<float (*){dims}>(sp_smem + {off}))")
                off += L_["D"] * L_["C"] * (s.ty + 2 * L_["HY"] + 1) * (s.tx + 2 * L_["H"])
            else:
                # one spare row per channel (the last channel's absorbs the stores of unused
                # halo entries)
                lines.append(f"__shared__ float lds_s{i}{dims};")
        if self.stage:
            lines.append(f"__shared__ float so_stage[NT / 64][{64 * self.stage_c}];")
        lines += self._xform_functions()
        fns = []
        for m in self._variants():
            name, fn = self._emit_variant(m)
            fns.append((m, name))
            lines.append(fn)
        lines.append(self._kernel(fns, ndet))
        if self.split_tail:
            lines.append(self._tail_kernel(ndet))
        if ndet and not self.batched:
            # flag pass: the same z-march evaluating only the determinant inputs' diagonals;
            # its any-small-diagonal result sets the main pass's regularisation on the device
            self._flag = True
            try:
                ffns = []
                for m in self._variants():
                    name, fn = self._emit_variant(m)
                    ffns.append((m, name))
                    lines.append(fn)
                lines.append(self._kernel(ffns, ndet))
                if self.split_tail:
                    lines.append(self._tail_kernel(ndet))
            finally:
                self._flag = False
        return lines

    def _tail_kernel(self, ndet: int) -> str:
        """``fused_tail``: the boundary boxes alone, block b = p.nblocks + blockIdx.x (its
        records follow the main kernel's)."""
        g = self.group
        out: List[str] = []
        A = out.append
        kname = "fused_flag_tail" if getattr(self, "_flag", False) else "fused_tail"
        A(f'extern "C" __global__ void __launch_bounds__(NTB) {kname}(const Params p) {{')
        A("    __shared__ float red_scratch[4 * NTB / 64 + 4];")
        A("    const int tid = threadIdx.x, nb = p.nblocks, b = nb + (int)blockIdx.x, rb = b;")
        if getattr(self, "_flag", False):
            # the flag is an "any": once one block has set it, blocks that start later stop
            A("    if (__hip_atomic_load((double*)p.detflag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > 0.0) return;")
        A("    Acc A;")
        for k in range(len(self.accumulators()[0])):
            A(f"    A.a{k}.init();")
        for i in range(ndet):
            A(f"    A.dflag{i} = 0.0f;")
        # the global +1e-12 I regularisation of determinants, decided by the flag pass
        A("    A.detreg = p.detflag[0] > 0.0 ? 1e-12f : 0.0f;" if ndet and not getattr(self, "_flag", False)
          else "    A.detreg = 0.0f;")
        for i in range(len(g.dscalars)):
            A(f"    A.sc{i} = (float)p.scal{i}[0];")
        self._boundary_part(A, ndet)
        A("}")
        return "\n".join(out)

    # -- kernel skeleton ---------------------------------------------------------------------
    def _kernel(self, fns, ndet: int) -> str:
        s, g = self.sched, self.group
        out: List[str] = []
        A = out.append
        if self.batched:
            # one work unit (tile or tail block) of one grid; called from the batch kernel
            A("__device__ __forceinline__ void run_unit(const Params& p, const int b, float* red_scratch) {")
            A("    const int tid = threadIdx.x, rb = b;")
            A("    const int nb = p.nblocks;")
        else:
            kname = "fused_flag" if getattr(self, "_flag", False) else "fused_kernel"
            A(f'extern "C" __global__ void __launch_bounds__(NTB) {kname}(const Params p) {{')
            A("    __shared__ float red_scratch[4 * NTB / 64 + 4];")
            A("    const int tid = threadIdx.x, rb = blockIdx.x;")
            if getattr(self, "_flag", False):
                # the flag is an "any": once one block has set it, blocks that start later stop
                A("    if (__hip_atomic_load((double*)p.detflag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > 0.0) return;")
            # XCD-aware block remap: XCD k (hardware block id % 8) walks a contiguous tile range
            A("    const int b = blockIdx.x, nb = p.nblocks, per = nb >> 3, rem = nb & 7;")
        A("    Acc A;")
        for k in range(len(self.accumulators()[0])):
            A(f"    A.a{k}.init();")
        for i in range(ndet):
            A(f"    A.dflag{i} = 0.0f;")
        # the global +1e-12 I regularisation of determinants, decided by the flag pass
        A("    A.detreg = p.detflag[0] > 0.0 ? 1e-12f : 0.0f;" if ndet and not getattr(self, "_flag", False)
          else "    A.detreg = 0.0f;")
        for i in range(len(g.dscalars)):
            A(f"    A.sc{i} = (float)p.scal{i}[0];")
        if self.use_lds and not self.split_tail:
            self._boundary_part(A, ndet)
        if self.batched:
            # the unit list is ordered on the host so that each XCD gets contiguous tiles
            A("    const int lb = b;")
        else:
            A("    const int xcd = b & 7, lb = xcd * per + (xcd < rem ? xcd : rem) + (b >> 3);")
        A("    const int tx_i = lb % p.tiles_x, t_ = lb / p.tiles_x;")
        A("    const int ty_i = t_ % p.tiles_y, zc_i = t_ / p.tiles_y;")
        A("    const int x0 = tx_i * TX, y0 = ty_i * TY;")
        A("    const int zb = zc_i * ZC, ze = (zb + ZC < N0) ? zb + ZC : N0;")
        A("    int lx = tid % TX, ly0 = tid / TX;")
        if self.fw:
            # the face wave (threads NT..NT+63): lane -> (face, column d, row phase); it loads
            # its own face cells like any thread so that every wave runs the same memory code
            A("    const bool fwave = tid >= NT;")
            A("    bool fact = false;")
            A("    if (fwave) {")
            A("        const int fl = tid - NT, fs = fl / (RP * REDGE), fd = (fl / RP) % REDGE;")
            A("        const bool hlo = XLO_IN && x0 == 0, hhi = XHI_IN && x0 + TX == N2;")
            A("        fact = fs < (int)hlo + (int)hhi;")
            A("        const bool lo = fs == 0 && hlo;")
            A("        ly0 = fl % RP;")
            A("        lx = !fact ? 0 : (lo ? fd : TX - 1 - fd);")
            A("    }")
            A("    const bool xface_tile = (XLO_IN && x0 == 0) || (XHI_IN && x0 + TX == N2);")
        A("    const bool ex = (x0 < REDGE) || (x0 + TX - 1 + REDGE > N2 - 1);")
        A("    const bool ey = (y0 < REDGE) || (y0 + TY - 1 + REDGE > N1 - 1);")
        # partial tiles (last tile of an axis not divisible by the tile) use the gather
        # variant: one-sided edge formulas there reach inward past the tile start
        A("    const bool partial = (x0 + TX > N2) || (y0 + TY > N1);")
        if self.use_lds and self.edge_mode != "band":
            # partial tiles are evaluated by the boundary kernel; write empty records
            A("    if (partial) {")
            for ri, nid in enumerate(g.reductions):
                A(f"        if (tid == 0) {{ double* r_ = p.red{ri} + (i64)rb * 4; r_[0] = 0.0; r_[1] = 0.0; "
                  f"r_[2] = {'INFINITY' if self.graph.node(nid).op_name == 'min_max' else '0.0'}; r_[3] = 0.0; "
                  f"{'r_[0] = INFINITY; r_[1] = -INFINITY;' if self.graph.node(nid).op_name == 'min_max' else ''} }}")
            for i in range(ndet):
                A(f"        if (tid == 0) {{ double* r_ = p.dflag{i} + (i64)rb * 4; r_[0] = INFINITY; r_[1] = -INFINITY; r_[2] = 0.0; r_[3] = 0.0; }}")
            A("        return;")
            A("    }")
        names = {m: n for m, n in fns}
        qs = self.queue_sources()
        lds = self.lds_sources()
        if self.use_lds:
            self._emit_loaders(A)
        if self.use_lds and self.fw:
            # Wave-uniform split (NT is a multiple of 64): the compute waves run straight-line,
            # unconditional loads, so the compiler waits on load counts, never on zero; the face
            # wave loads only its own face cells in x-face tiles and elsewhere just meets the
            # barriers: one for the prologue and one per plane, as the compute waves do.
            A("    if (__builtin_amdgcn_readfirstlane((int)fwave)) {")
            A("      if (!xface_tile) {")
            A("        for (int zz_ = zb - 1; zz_ < ze; ++zz_) __syncthreads();")
            A("      } else {")
            self._role = "face"
            self._emit_march(A, names)
            A("      }")
            A("    } else {")
            self._role = "compute"
            self._emit_march(A, names)
            A("    }")
            self._role = "all"
        else:
            self._emit_march(A, names)
        self._write_records(A, ndet, "    ")
        A("}")
        return "\n".join(out)

    def _emit_march(self, A, names) -> None:
        """Prologue and z-loop for the current role (unrolled by four: the register ring)."""
        if self.use_lds:
            self._emit_prologue(A)
        A("    for (int z = zb; z < ze; z += 4) {")
        self._emit_step(A, names, "A", 0, t=0)
        for t in (1, 2, 3):
            A(f"        if (z + {t} >= ze) break;")
            A("        {")
            A(f"        const int z{t}_ = z + {t};")
            self._emit_step(A, names, "AB"[t % 2], t % 2, zname=f"z{t}_", t=t)
            A("        }")
        A("    }")

    def _halo_count(self, i: int) -> int:
        L_ = self.src_layout[i]
        if not L_["lds"]:
            return 0
        H, HY, C = L_["H"], L_["HY"], L_["C"]
        W = self.sched.tx + 2 * H
        return (2 * HY * W + 2 * H * self.sched.ty) * C

    def _emit_loaders(self, A) -> None:
        s = self.sched
        for i, L_ in enumerate(self.src_layout):
            C, H, HY = L_["C"], L_["H"], L_["HY"]
            W, RH = s.tx + 2 * H, s.ty + 2 * HY
            HT = self._halo_count(i)
            KH = (HT + s.threads - 1) // s.threads
            L_["KH"] = KH
            # own cells: plane-relative offsets; cells past the grid edge (partial tiles) load
            # their periodic images, like the halo (band mode), or the edge value (variants)
            A(f"    unsigned og{i}[VPT];")
            A("    #pragma unroll")
            A("    for (int k = 0; k < VPT; ++k)")
            fn = "wrapm" if self.edge_mode == "band" else "clampc"
            A(f"        og{i}[k] = ({fn}(y0 + ly0 + k * RP, N1) * N2 + {fn}(x0 + lx, N2)) * {C};")
            if KH:
                A(f"    unsigned hg{i}[{KH}];")
                A(f"    int hl{i}[{KH}];")
                A("    #pragma unroll")
                A(f"    for (int j = 0; j < {KH}; ++j) {{")
                A(f"        const int e = tid < NT ? tid + j * NT : {HT};")
                A(f"        int r, col, ch;")
                top = HY * W * C
                mid = s.ty * 2 * H * C
                A(f"        if (e < {top}) {{ r = e / {W * C}; const int q = e - r * {W * C}; col = q / {C}; ch = q - col * {C}; }}")
                A(f"        else if (e < {top + mid}) {{ const int e2 = e - {top}; r = {HY} + e2 / {2 * H * C};"
                  f" const int q = e2 - (r - {HY}) * {2 * H * C}; const int jj = q / {C}; ch = q - jj * {C};"
                  f" col = jj < {H} ? jj : TX + jj; }}")
                A(f"        else {{ const int e3 = e - {top + mid}; r = TY + {HY} + e3 / {W * C};"
                  f" const int q = e3 - (r - TY - {HY}) * {W * C}; col = q / {C}; ch = q - col * {C}; }}")
                A(f"        hg{i}[j] = (wrapm(y0 - {HY} + r, N1) * N2 + wrapm(x0 - {H} + col, N2)) * {C} + ch;")
                # out-of-range entries write a spare word at the end of the slot array
                A(f"        hl{i}[j] = e < {HT} ? (ch * {RH + 1} + r) * {W} + col : {((C - 1) * (RH + 1) + RH) * W};")
                A("    }")
            for buf in ("A", "B"):
                if L_["mode"] != "queue":
                    A(f"    float {buf}o{i}[VPT][{C}];")
                if KH:
                    A(f"    float {buf}h{i}[{KH}];")
            if L_["mode"] == "queue":
                # planes z-1, z, z+1 and the one in flight live in a ring of four register sets
                # whose roles advance one slot per plane; the z-loop is unrolled by four, so
                # every role is a fixed register and nothing is copied between planes
                A(f"    float " + ", ".join(f"R{i}_{r}[VPT][{C}]" for r in range(4)) + ";")

    def _fetch(self, A, i: int, buf: str, zexpr: str, indent: str = "    ",
               own: Optional[str] = None) -> None:
        """Load plane ``zexpr`` of source i: own cells into ``own`` (default ``{buf}o{i}``),
        halo cells into ``{buf}h{i}``."""
        own = own or f"{buf}o{i}"
        L_ = self.src_layout[i]
        C, KH = L_["C"], L_.get("KH", 0)
        face = getattr(self, "_role", "all") == "face"
        if face and L_["mode"] != "queue":
            return          # the face wave reads LDS-ring sources from LDS only
        if face:
            KH = 0          # ... and needs the own cells of queue sources, never their halo
        # a uniform plane pointer plus loop-invariant unsigned lane offsets: the loads take the
        # scalar-base + 32-bit offset form and need no per-load address arithmetic
        A(f"{indent}{{ const float* __restrict__ pl_ = p.src{i} + (IDX)wrapm({zexpr}, N0) * (N1 * N2 * {C});")
        A(f"{indent}  #pragma unroll")
        A(f"{indent}  for (int k = 0; k < VPT; ++k)")
        A(f"{indent}    #pragma unroll")
        A(f"{indent}    for (int c = 0; c < {C}; ++c) {own}[k][c] = pl_[og{i}[k] + (unsigned)c];")
        if KH:
            A(f"{indent}  #pragma unroll")
            A(f"{indent}  for (int j = 0; j < {KH}; ++j) {buf}h{i}[j] = pl_[hg{i}[j]];")
        A(f"{indent}}}")

    def _xf_own(self, A, i: int, buf: str, indent: str = "    ", own: Optional[str] = None) -> None:
        """Apply source i's load transform to its own-cell buffer in place."""
        own = own or f"{buf}o{i}"
        if i in self.xform:
            A(f"{indent}#pragma unroll")
            A(f"{indent}for (int k = 0; k < VPT; ++k) {own}[k][0] = xf{i}({own}[k][0]);")

    def _store(self, A, i: int, buf: str, slot: str, indent: str = "    ", own_done: bool = False,
               own: Optional[str] = None) -> None:
        own = own or f"{buf}o{i}"
        L_ = self.src_layout[i]
        if getattr(self, "_role", "all") == "face":
            return          # the compute waves fill LDS, the face wave only reads it
        if not own_done:
            self._xf_own(A, i, buf, indent, own)
        if not L_["lds"]:
            return
        C, H, HY, KH = L_["C"], L_["H"], L_["HY"], L_.get("KH", 0)
        W, RH = self.sched.tx + 2 * H, self.sched.ty + 2 * HY
        A(f"{indent}{{ float* ls = &lds_s{i}[{slot}][0][0][0];")
        A(f"{indent}  #pragma unroll")
        A(f"{indent}  for (int k = 0; k < VPT; ++k)")
        A(f"{indent}    #pragma unroll")
        A(f"{indent}    for (int c = 0; c < {C}; ++c) ls[(c * {RH + 1} + ly0 + k * RP + {HY}) * {W} + lx + {H}] = {own}[k][c];")
        if KH:
            xf = (lambda v: f"xf{i}({v})") if i in self.xform else (lambda v: v)
            A(f"{indent}  #pragma unroll")
            A(f"{indent}  for (int j = 0; j < {KH}; ++j) ls[hl{i}[j]] = {xf(f'{buf}h{i}[j]')};")
        A(f"{indent}}}")

    def _lead(self, i: int) -> int:
        return 1 if self.src_layout[i]["mode"] == "queue" else self.hz + 1

    def _emit_prologue(self, A) -> None:
        for i, L_ in enumerate(self.src_layout):
            C = L_["C"]
            if L_["mode"] == "queue":
                # ring slot r holds plane zb - 1 + r
                self._fetch(A, i, "A", "zb - 1", own=f"R{i}_0")
                self._xf_own(A, i, "A", own=f"R{i}_0")
                self._fetch(A, i, "A", "zb", own=f"R{i}_1")
                self._xf_own(A, i, "A", own=f"R{i}_1")
                self._store(A, i, "A", "0", own_done=True, own=f"R{i}_1")
                self._fetch(A, i, "A", "zb + 1", own=f"R{i}_2")
                self._fetch(A, i, "B", "zb + 2", own=f"R{i}_3")
                continue
            if getattr(self, "_role", "all") == "face":
                continue
            A("    for (int k = -HZ; k <= HZ; ++k) {")
            self._fetch(A, i, "A", "zb + k", "        ")
            self._store(A, i, "A", f"(unsigned)(k + HZ) % {L_['D']}u", "        ")
            A("    }")
            self._fetch(A, i, "A", f"zb + {self._lead(i)}")
            self._fetch(A, i, "B", f"zb + 1 + {self._lead(i)}")
        A("    __syncthreads();")

    def _emit_step(self, A, names, buf: str, sq: int, zname: str = "z", t: int = 0) -> None:
        """One plane of the z-march at ring phase ``t`` (0..3): queue source i reads plane z-1
        from ring slot t, z from t+1 and z+1 from t+2 (mod 4), and the plane fetched after the
        step, z+3, replaces slot t."""
        ind = "        "
        z = zname
        ring = lambda i, d: f"R{i}_{(t + d) % 4}"
        A(f"{ind}{{ const int s = {z} - zb + HZ;")
        A(f"{ind}  const bool ez = ({z} < REDGE) || ({z} + REDGE > N0 - 1);")
        # queue sources: the next plane enters the register queue now, transformed
        for i in self.queue_sources():
            self._xf_own(A, i, buf, ind + "  ", own=ring(i, 2))
        role = getattr(self, "_role", "all")
        qa = "".join(f", {ring(i, 0)}[k], {ring(i, 1)}[k], {ring(i, 2)}[k]" for i in self.queue_sources())
        call = f"(p, A, s, {sq}, {z}, y, x, ly, lx, y0, x0, vox{qa});"
        if role != "face":
            self._emit_interior(A, names, ind, z, call)
        if self.fw and role == "face":
            A(f"{ind}  if (fact) {{")
            A(f"{ind}  #pragma unroll")
            A(f"{ind}  for (int k = 0; k < VPT; ++k) {{")
            A(f"{ind}    const int ly = ly0 + k * RP, y = y0 + ly, x = x0 + lx;")
            A(f"{ind}    if (!ez && y >= YLO && y < YHI) {{")
            A(f"{ind}      const IDX vox = (IDX){z} * (N1 * N2) + (IDX)(unsigned)(y * N2 + x);")
            first = True
            for side in self.xface_sides():
                for d in range(self.redge):
                    xc = d if side == "lo" else self.grid[2] - 1 - d
                    A(f"{ind}      {'if' if first else 'else if'} (x == {xc}) {self._variant_name(('xface', side, d))}{call}")
                    first = False
            A(f"{ind}    }}")
            A(f"{ind}  }}")
            A(f"{ind}  }}")
        if self.use_lds:
            for i, L_ in enumerate(self.src_layout):
                q = L_["mode"] == "queue"
                slot = str(sq ^ 1) if q else f"(unsigned)(s + HZ + 1) % {L_['D']}u"
                self._store(A, i, buf, slot, ind + "  ", own_done=q, own=ring(i, 2) if q else None)
            A(f"{ind}  __syncthreads();")
            # always fetch (clamped to the last plane this block needs, an L2 hit) so that
            # every path issues the same loads and the compiler can wait on counts, not zero
            for i, L_ in enumerate(self.src_layout):
                ld = self._lead(i)
                self._fetch(A, i, buf, f"min({z} + 2 + {ld}, ze - 1 + {ld})", ind + "  ",
                            own=ring(i, 0) if L_["mode"] == "queue" else None)
        A(f"{ind}}}")

    def _emit_interior(self, A, names, ind: str, z: str, call: str) -> None:
        """The compute waves' evaluation of one plane (interior variant, optional staging)."""
        A(f"{ind}  #pragma unroll")
        A(f"{ind}  for (int k = 0; k < VPT; ++k) {{")
        A(f"{ind}    const int ly = ly0 + k * RP, y = y0 + ly, x = x0 + lx;")
        staged = bool(self.stage) and not getattr(self, "_flag", False)
        if staged:
            A(f"{ind}    A.vflag = 0;")
        A(f"{ind}    if (y < N1 && x < N2) {{")
        A(f"{ind}      const IDX vox = (IDX){z} * (N1 * N2) + (IDX)(unsigned)(y * N2 + x);")
        if self.edge_mode == "band" and ("lds", False, False) in names:
            # z-edge planes, the y face band and corners belong to the tail blocks; the
            # x-face band to the face wave
            A(f"{ind}      if (!ez && y >= YLO && y < YHI && x >= XLO && x < XHI) {self._variant_name(('lds', False, False))}{call}")
        elif ("lds", False, False) in names:
            # z-edge planes belong to the boundary kernel
            A(f"{ind}      if (ez) {{ }}")
            vn = lambda ex_, ey_: self._variant_name(("lds", ex_, ey_))
            A(f"{ind}      else if (!ex && !ey) {vn(False, False)}{call}")
            A(f"{ind}      else if (ex && !ey) {vn(True, False)}{call}")
            A(f"{ind}      else if (!ex) {vn(False, True)}{call}")
            A(f"{ind}      else {vn(True, True)}{call}")
        else:
            A(f"{ind}      {self._variant_name(('gather',))}{call}")
        A(f"{ind}    }}")
        if staged:
            # lane l of a wave holds voxel (y_0 + l / W, x_0 + l % W), W = min(TX, 64): one run of
            # 64 consecutive voxels, or 64 / TX runs of TX on consecutive rows.  Each lane
            # deposited its C channels inside the variant; element q * 64 + l of the wave's C * 64
            # outputs is stored by lane l, so a store instruction writes whole lines
            W = min(self.sched.tx, 64)
            # element e of the area (voxel v = e / C) lives at C * (lane 0's voxel) + e, plus one
            # row, (N2 - W) * C further, for every W voxels
            A(f"{ind}    {{ const unsigned long long vm = __ballot(A.vflag);")
            A(f"{ind}      if (vm) {{")
            A(f"{ind}        const int lane = tid & 63;")
            A(f"{ind}        const unsigned vlo = (unsigned)vm, vhi = (unsigned)(vm >> 32);")
            A(f"{ind}        const IDX v0 = ((IDX){z} * N1 + (y - lane / {W})) * N2 + (x - lane % {W});")
            A(f"{ind}        wave_sync();")
            # a wave whose 64 voxels were all evaluated here (the interior of the grid) stores
            # without per-element tests: chunk q of 64 elements spans at most two rows (C * W >= 64),
            # which are known at compile time, so only the chunk that straddles a row end picks
            # the row per lane
            full = all(self.info[self.group.outputs[oi]].ch * W >= 64 for oi in self.stage)
            if full:
                A(f"{ind}        if (vm == ~0ull) {{")
                for oi in self.stage:
                    C = self.info[self.group.outputs[oi]].ch
                    A(f"{ind}          {{ const float* st = &so_stage[tid >> 6][{self.stage_off[oi]}];")
                    A(f"{ind}            float* ob = (float*)p.out{oi} + v0 * {C};")
                    for q in range(C):
                        r_lo, r_hi = (64 * q // C) // W, ((64 * q + 63) // C) // W
                        if r_lo == r_hi:
                            off = f"{64 * q} + lane + {r_lo} * ((N2 - {W}) * {C})" if r_lo else f"{64 * q} + lane"
                        else:
                            thr = C * W * r_hi - 64 * q
                            off = f"{64 * q} + lane + (lane >= {thr} ? {r_hi} : {r_lo}) * ((N2 - {W}) * {C})"
                        A(f"{ind}            st_f32(ob, {off}, st[{64 * q} + lane]);")
                    A(f"{ind}          }}")
                A(f"{ind}        }} else {{")
            for oi in self.stage:
                C = self.info[self.group.outputs[oi]].ch
                A(f"{ind}        {{ const float* st = &so_stage[tid >> 6][{self.stage_off[oi]}];")
                A(f"{ind}          float* ob = (float*)p.out{oi} + v0 * {C};")
                A(f"{ind}          #pragma unroll")
                A(f"{ind}          for (int q = 0; q < {C}; ++q) {{")
                A(f"{ind}            const int e = q * 64 + lane, v = e / {C};")
                # the ballot's halves as 32-bit words: one register for the test, not two
                A(f"{ind}            if (((v < 32 ? vlo : vhi) >> (v & 31)) & 1u) st_f32(ob, e + (v / {W}) * ((N2 - {W}) * {C}), st[e]);")
                A(f"{ind}          }} }}")
            if full:
                A(f"{ind}        }}")
            A(f"{ind}        wave_sync();")
            A(f"{ind}      }}")
            A(f"{ind}    }}")
        A(f"{ind}  }}")


# ------------------------------------------------------------------------------ compiled plans
# FMA contraction across statements, except where a function asks for none (the compensated
# accumulators: a product contracted into their running sum breaks the error terms)
_RTC_OPTIONS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                "-Xclang", "-packed-fp32-ops", "-ffp-contract=fast-honor-pragmas"]
_cache_lock = threading.Lock()
_module_cache: Dict[str, L.Module] = {}


def _cache_dir() -> Optional[Path]:
    d = os.environ.get("STENPY_CACHE_DIR")
    p = Path(d) if d else Path.home() / ".cache" / "stenpy_amd"
    try:
        p.mkdir(parents=True, exist_ok=True)
        return p
    except OSError:
        return None


def compile_source(src: str, kernel: str = "fused_kernel") -> Tuple[bytes, float]:
    """hipRTC-compile ``src`` (disk-cached by content hash); returns (code, seconds)."""
    h = hashlib.sha1((src + "|".join(_RTC_OPTIONS)).encode()).hexdigest()
    cd = _cache_dir()
    if cd is not None:
        f = cd / f"{h}.co"
        if f.exists():
            return f.read_bytes(), 0.0
    t = time.perf_counter()
    code = L.rtc_compile(src, "fused.hip", _RTC_OPTIONS)
    dt = time.perf_counter() - t
    if cd is not None:
        tmp = cd / f"{h}.{os.getpid()}.tmp"
        tmp.write_bytes(code)
        os.replace(tmp, cd / f"{h}.co")
    return code, dt


def kernel_registers(code: bytes) -> Dict[str, int]:
    """Vector registers (VGPRs + AGPRs) of each kernel of a gfx950 code object, read from its
    AMDGPU metadata note (an ELF note of type 32 holding msgpack), without a GPU."""
    import msgpack
    shoff, = struct.unpack_from("<Q", code, 0x28)
    shentsize, shnum = struct.unpack_from("<HH", code, 0x3A)
    out: Dict[str, int] = {}
    for i in range(shnum):
        base = shoff + i * shentsize
        sh_type, = struct.unpack_from("<I", code, base + 4)
        off, size = struct.unpack_from("<QQ", code, base + 0x18)
        if sh_type != 7:                                   # SHT_NOTE
            continue
        q = off
        while q + 12 <= off + size:
            nsz, dsz, typ = struct.unpack_from("<III", code, q)
            d0 = q + 12 + ((nsz + 3) & ~3)
            if typ == 32 and code[q + 12:q + 12 + nsz].rstrip(b"\0") == b"AMDGPU":
                md = msgpack.unpackb(code[d0:d0 + dsz], raw=False)
                for k in md.get("amdhsa.kernels", []):
                    out[k[".name"]] = int(k.get(".vgpr_count", 0)) + int(k.get(".agpr_count", 0))
            q = d0 + ((dsz + 3) & ~3)
    return out


def load_module(src: str, kernel: str = "fused_kernel") -> L.Module:
    h = hashlib.sha1(src.encode()).hexdigest()
    with _cache_lock:
        m = _module_cache.get(h)
        if m is None:
            code, _ = compile_source(src, kernel)
            m = L.Module(code, kernel)
            _module_cache[h] = m
    return m


@dataclass
class CompiledGroup:
    gen: KernelGen
    source: str
    module: Optional[L.Module] = None
    others: Dict[str, L.Module] = dc_field(default_factory=dict)

    def kernel(self, name: str) -> L.Module:
        """Another entry point of the same code object (fused_tail, fused_flag, ...)."""
        if name == "fused_kernel":
            if self.module is None:
                self.module = load_module(self.source)
            return self.module
        if name not in self.others:
            code, _ = compile_source(self.source)
            self.others[name] = L.Module(code, name)
        return self.others[name]

    @property
    def group(self) -> Group:
        return self.gen.group

    def nblocks(self) -> Tuple[int, int, int, int]:
        s = self.gen.sched
        n0, n1, n2 = self.gen.grid
        tx, ty, zcn = -(-n2 // s.tx), -(-n1 // s.ty), -(-n0 // s.zc)
        return tx, ty, zcn, tx * ty * zcn


class CapturedGraph:
    """A hipGraph of fused launches together with what its replays need alive: the compiled
    groups whose code-object modules hold the recorded kernels (and, for a batch, the runner
    that owns the recorded buffers).  Dropping the executor that captured it is safe."""

    def __init__(self, graph: "torch.cuda.CUDAGraph", keep: Sequence[Any]):
        self.graph = graph
        self._keep = list(keep)

    def replay(self) -> None:
        self.graph.replay()


Z_ALL = (0, 2 ** 31 - 1)       # reduction window covering every plane


def pack_params(cg: CompiledGroup, srcs: Sequence[torch.Tensor], outs: Sequence[torch.Tensor],
                reds: Sequence[torch.Tensor], dflags: Sequence[torch.Tensor],
                scal: Sequence[torch.Tensor], detflag: Optional[torch.Tensor] = None,
                zwin: Tuple[int, int] = Z_ALL) -> bytes:
    tx, ty, zcn, nb = cg.nblocks()
    ptrs = [t.data_ptr() for t in srcs] + [t.data_ptr() for t in outs] + \
           [t.data_ptr() for t in reds] + [t.data_ptr() for t in dflags] + [t.data_ptr() for t in scal]
    if dflags:
        ptrs.append(detflag.data_ptr())
    return struct.pack(f"<{len(ptrs)}Q6i", *ptrs, tx, ty, zcn, nb, int(zwin[0]), int(zwin[1]))


# ------------------------------------------------------------------------------ scalar steps
_SCALAR_UNARY = {"neg", "exp", "log", "sqrt", "sin", "cos", "tanh", "clamp"}
_SCALAR_BINARY = {"add", "sub", "mul", "div"}


def scalar_op(node: GraphNode, ins: Sequence[Any], out: torch.Tensor) -> Optional["L.SpScalarOp"]:
    """The ``sp_scalar_program`` op computing ``node`` (a point-wise op whose operands are
    single FP64 device values or Python numbers) into the FP64 device value ``out``; None when
    the node is not such an op.  Parameters follow the registered ops' defaults
    (stenpy_engine.py:1248-1367: log clamps at 1e-15, sqrt at 0, clamp to [0, 1], div
    replaces |b| < 1e-15 by +-1e-15)."""
    op, p = node.op_name, node.params
    if op not in _SCALAR_UNARY and op not in _SCALAR_BINARY:
        return None
    if out.dtype != torch.float64 or out.numel() != 1 or not out.is_cuda:
        return None
    vals = []
    for v in ins:
        if isinstance(v, torch.Tensor) and not v.is_cuda and v.numel() == 1:
            vals.append((None, float(v.reshape(-1)[0])))     # a seeded host scalar
        elif isinstance(v, torch.Tensor):
            if v.dtype != torch.float64 or v.numel() != 1 or v.device != out.device:
                return None
            vals.append((v.data_ptr(), 0.0))
        elif isinstance(v, (int, float)) and not isinstance(v, bool):
            vals.append((None, float(v)))
        else:
            return None
    d = L.SpScalarOp()
    d.out = out.data_ptr()
    if op in _SCALAR_BINARY and not (op == "mul" and len(vals) == 1):
        if len(vals) != 2:
            return None
        d.binary, d.op = 1, L.BINARY[op]
        (d.a, d.av), (d.b, d.bv) = vals
        d.p0 = float(p.get("eps", 1e-15)) if op == "div" else 0.0
        return d
    if len(vals) != 1:
        return None
    d.binary = 0
    (d.a, d.av) = vals[0]
    if op == "mul":                                  # mul(a, scalar=c)
        if p.get("scalar") is None:
            return None
        d.op, d.p0 = L.UNARY["scale"], float(p["scalar"])
    else:
        d.op = L.UNARY[op]
        if op == "log":
            d.p0 = float(p.get("eps", 1e-15))
        elif op == "sqrt":
            d.p0 = float(p.get("eps", 0.0))
        elif op == "clamp":
            d.p0, d.p1 = float(p.get("lo", 0.0)), float(p.get("hi", 1.0))
    return d


# ------------------------------------------------------------------------------ executor
class FusedExecutor:
    """Runs graphs through fused plans; caches plans by (graph structure, shapes, schedule)."""

    def __init__(self, device: torch.device, schedule_fn=None, edge_mode: Optional[str] = None):
        self.device = device
        self.schedule_fn = schedule_fn
        self.edge_mode = edge_mode
        self._side: Optional[torch.cuda.Stream] = None
        self._plans: Dict[str, Tuple[List[Step], Dict[str, Info], Dict[int, CompiledGroup]]] = {}
        self.last_timing: Dict[str, float] = {}

    def side_stream(self) -> "torch.cuda.Stream":
        """Stream of the tail kernels (default priority, created once)."""
        if self._side is None:
            self._side = torch.cuda.Stream(device=self.device)
        return self._side

    def schedule_for(self, group: Group, graph: Graph) -> Schedule:
        if self.schedule_fn is not None:
            s = self.schedule_fn(group, graph)
            if s is not None:
                return s
        return fill_small_grid(heuristic_schedule(group.grid), group.grid)

    @staticmethod
    def _fit_registers(gen: "KernelGen", src: str) -> str:
        """Keep the main kernel at >= 4 waves per SIMD (<= MERGED_TAIL_MAX_VGPRS registers) when
        one of two code-size choices pushes it past that: first the edge band moves to its own
        side-stream kernel, then the LDS staging of wide outputs is dropped (direct stores).
        Register counts come from the code object's metadata, so no GPU is needed."""
        can_split = SPLIT_TAIL == "auto" and gen.use_lds and not gen.batched and gen.boundary_blocks() > 0
        if not (gen.stage or can_split):
            return src
        regs = lambda s_: kernel_registers(compile_source(s_)[0]).get("fused_kernel", 0)
        if can_split and regs(src) > MERGED_TAIL_MAX_VGPRS:
            gen.want_split = True
            src = gen.source()
        if gen.stage and regs(src) > MERGED_TAIL_MAX_VGPRS:
            gen.stage, gen.stage_off, gen.stage_c = [], {}, 0
            src = gen.source()
        return src

    def prepare(self, graph: Graph, values: Dict[str, Any], wanted: Sequence[str],
                out_dtype: Optional[torch.dtype] = None, schedule: Optional[Schedule] = None):
        shapes = tuple(sorted((k, _value_signature(v)) for k, v in values.items()))
        key = graph.canonical_key((shapes, tuple(sorted(wanted)), str(out_dtype),
                                   schedule.as_tuple() if schedule else None,
                                   id(self.schedule_fn), self.edge_mode or EDGE_MODE))
        plan = self._plans.get(key)
        if plan is not None:
            return plan
        steps, info = plan_graph(graph, values, wanted)
        compiled: Dict[int, CompiledGroup] = {}
        for si, st in enumerate(steps):
            if st.kind != "group":
                continue
            g = st.group
            sched = schedule or self.schedule_for(g, graph)
            od = {}
            if out_dtype is not None:
                for o in g.outputs:
                    if o in wanted:
                        od[o] = out_dtype
            for nid in g.outputs:
                if graph.node(nid).op_name == "cast_bf16":
                    od[nid] = torch.bfloat16
            gen = KernelGen(graph, g, info, sched, od, {}, self.edge_mode)
            if gen.lds_bytes() > LDS_LIMIT and schedule is None:
                for alt in FALLBACK_SCHEDULES:
                    gen = KernelGen(graph, g, info, alt, od, {}, self.edge_mode)
                    if gen.lds_bytes() <= LDS_LIMIT:
                        break
            if gen.lds_bytes() > LDS_LIMIT:
                raise RuntimeError(f"fused group needs {gen.lds_bytes()} B of LDS (limit {LDS_LIMIT})")
            src = gen.source()
            src = self._fit_registers(gen, src)
            compiled[si] = CompiledGroup(gen, src)
        plan = (steps, info, compiled)
        self._plans[key] = plan
        return plan

    def run(self, graph: Graph, runtime, seed: Optional[Dict[str, Any]] = None,
            wanted: Optional[Sequence[str]] = None, out_dtype: Optional[torch.dtype] = None,
            schedule: Optional[Schedule] = None,
            red_window: Optional[Tuple[int, int]] = None) -> Dict[str, Any]:
        """Evaluate ``graph``; ``red_window = (z0, z1)`` restricts every reduction to the
        planes [z0, z1) of its input (a padded slab's interior, for the out-of-core path)."""
        from .registry import OP_REGISTRY
        results: Dict[str, Any] = dict(seed or {})
        values: Dict[str, Any] = {}
        for nid in graph._order:
            node = graph.node(nid)
            if nid in results:
                values[nid] = results[nid]
            elif node.op_name == "_constant":
                v = runtime._materialize(node.params.get("value"))
                values[nid] = v
                results[nid] = v
        if wanted is None:
            wanted = [s for s in graph.sinks() if graph.node(s).op_name != "_constant"]
        steps, info, compiled = self.prepare(graph, values, wanted, out_dtype, schedule)
        pending: List[Any] = []           # scalar steps gathered into one device program

        def flush():
            if pending:
                L.scalar_program(pending, device=self.device)
                pending.clear()

        for si, st in enumerate(steps):
            if st.kind == "eager":
                node = graph.node(st.node)
                ins = [results[i] for i in node.input_ids]
                if info[st.node].ch == 0 and (node.op_name in _SCALAR_UNARY or node.op_name in _SCALAR_BINARY):
                    out = torch.empty((), dtype=torch.float64, device=self.device)
                    d = scalar_op(node, ins, out)
                    if d is not None:
                        pending.append(d)
                        results[st.node] = out
                        continue
                flush()
                if red_window is not None and node.op_name in REDUCTION_OPS:
                    ins[0] = ins[0][red_window[0]:red_window[1]]
                results[st.node] = OP_REGISTRY[node.op_name](*ins, **node.params)
                continue
            flush()
            self._run_group(compiled[si], graph, results, red_window)
        flush()
        return results

    def capture(self, graph: Graph, runtime, wanted: Sequence[str],
                out_dtype: Optional[torch.dtype] = None, warmup: int = 1,
                schedule: Optional[Schedule] = None):
        """Capture one evaluation of ``graph`` as a hipGraph.

        Returns (graph, results): replaying the graph re-evaluates into the same ``results``
        tensors.  The plans are prepared and compiled by the warm-up runs; a fused run makes no
        host synchronisation (reduction combines and the determinant flag stay on the device),
        so every launch, including the tail kernels' fork/join on the side stream, is recorded.
        Graphs containing non-fusable ops are captured only if those ops do not synchronise."""
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            for _ in range(warmup):
                self.run(graph, runtime, wanted=wanted, out_dtype=out_dtype, schedule=schedule)
        torch.cuda.current_stream(self.device).wait_stream(side)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            results = self.run(graph, runtime, wanted=wanted, out_dtype=out_dtype, schedule=schedule)
        keep = [cg for plan in self._plans.values() for cg in plan[2].values()]
        return CapturedGraph(g, keep), results

    def prepare_launch(self, cg: CompiledGroup, graph: Graph, results: Dict[str, Any]) -> "Launch":
        """Buffers and packed arguments of one launch of ``cg`` (inputs read from ``results``)."""
        g = cg.group
        if cg.module is None:
            cg.module = load_module(cg.source)
        dev = self.device
        n0, n1, n2 = g.grid
        srcs = []
        for s in g.sources:
            t = results[s]
            if t.dtype != torch.float32:
                t = t.to(torch.float32)
            if not t.is_contiguous():
                t = t.contiguous()
            L.require_gpu(t)
            srcs.append(t)
        outs = []
        for o in g.outputs:
            ch = cg.gen.info[o].ch
            shape = (n0, n1, n2) + {1: (), 3: (3,), 9: (3, 3)}[ch]
            dt = cg.gen.out_dtypes.get(o, torch.float32)
            outs.append(torch.empty(shape, dtype=dt, device=dev))
        tx, ty, zcn, nb = cg.nblocks()
        nrec = nb + cg.gen.boundary_blocks()
        reds = [torch.empty(nrec * 4, dtype=torch.float64, device=dev) for _ in g.reductions]
        ndet = len(cg.gen._det_inputs())
        dflags = [torch.empty(nrec * 4, dtype=torch.float64, device=dev) for _ in range(ndet)]
        detflag = torch.zeros(1, dtype=torch.float64, device=dev) if ndet else None
        scal = []
        for d in g.dscalars:
            t = results[d]
            if not isinstance(t, torch.Tensor):
                t = torch.tensor(float(t))
            scal.append(t.to(device=dev, dtype=torch.float64).reshape(1).contiguous())
        side = self.side_stream() if cg.gen.split_tail else None
        return Launch(cg, srcs, outs, reds, dflags, scal, nb, side, detflag)

    def _run_group(self, cg: CompiledGroup, graph: Graph, results: Dict[str, Any],
                   red_window: Optional[Tuple[int, int]] = None) -> None:
        g = cg.group
        n0, n1, n2 = g.grid
        dev = self.device
        ln = self.prepare_launch(cg, graph, results)
        z0, z1 = (0, n0) if red_window is None else (max(0, red_window[0]), min(n0, red_window[1]))
        ln.zwin = (z0, z1)
        stream = L.stream_ptr(dev)
        nb = ln.nblocks + cg.gen.boundary_blocks()
        if ln.dflags:
            # flag pass: sets the device-side any-small-diagonal flag the main pass reads
            ln.detflag.zero_()
            ln.launch(stream, entry="fused_flag")
        ln.launch(stream)
        for o, t in zip(g.outputs, ln.outs):
            results[o] = t
        # every FP64 final combine of the group in one launch (descriptors passed by value)
        descs = []
        for r, ws in zip(g.reductions, ln.reds):
            node = graph.node(r)
            inp = node.input_ids[0]
            count = (z1 - z0) * n1 * n2 * max(cg.gen.info[inp].ch, 1)
            res = torch.empty(2 if node.op_name == "min_max" else 1, dtype=torch.float64, device=dev)
            descs.append(struct.pack("<QQqii", ws.data_ptr(), res.data_ptr(), count,
                                     RED_KIND[node.op_name], nb))
            results[r] = res if node.op_name == "min_max" else res.reshape(())
        for i in range(0, len(descs), COMBINE_MULTI_MAX):
            chunk = descs[i:i + COMBINE_MULTI_MAX]
            buf = ctypes.create_string_buffer(b"".join(chunk))
            L.check(L.lib().sp_reduce_combine_multi(ctypes.addressof(buf), len(chunk), stream), "combine")


@dataclass
class Launch:
    """One prepared launch of a compiled group (inputs, outputs, workspaces, packed args)."""
    cg: CompiledGroup
    srcs: List[torch.Tensor]
    outs: List[torch.Tensor]
    reds: List[torch.Tensor]
    dflags: List[torch.Tensor]
    scal: List[torch.Tensor]
    nblocks: int
    side: Optional["torch.cuda.Stream"] = None
    detflag: Optional[torch.Tensor] = None
    zwin: Tuple[int, int] = Z_ALL

    def args(self) -> bytes:
        return pack_params(self.cg, self.srcs, self.outs, self.reds, self.dflags, self.scal,
                           self.detflag, self.zwin)

    def launch(self, stream: Optional[int] = None, entry: str = "fused_kernel") -> None:
        """Main kernel (or the flag pass, ``entry="fused_flag"``) on ``stream`` (default:
        torch's current stream); the tail kernel, when split, on the side stream between a
        fork and a join with ``stream``."""
        gen = self.cg.gen
        dev = self.srcs[0].device if self.srcs else self.outs[0].device
        st = stream if stream is not None else L.stream_ptr(dev)
        a = self.args()
        nbt = gen.boundary_blocks()
        tail = "fused_flag_tail" if entry == "fused_flag" else "fused_tail"
        if self.side is None or not nbt:
            self.cg.kernel(entry).launch((self.nblocks + nbt, 1, 1), (gen.block_threads(), 1, 1), a, 0, st)
            return
        main = torch.cuda.current_stream(dev) if stream is None else torch.cuda.ExternalStream(st, device=dev)
        self.side.wait_stream(main)
        self.cg.kernel(entry).launch((self.nblocks, 1, 1), (gen.block_threads(), 1, 1), a, 0, st)
        self.cg.kernel(tail).launch((nbt, 1, 1), (gen.block_threads(), 1, 1), a, 0, self.side.cuda_stream)
        main.wait_stream(self.side)

    def run_once(self, stream: Optional[int] = None) -> None:
        """One evaluation of the group's kernels: the determinant flag pass when the group has
        one (its flag cleared by a fill kernel first), then the main launch."""
        if self.dflags:
            self.detflag.zero_()
            self.launch(stream, entry="fused_flag")
        self.launch(stream)

    def time_ms(self, reps: int = 20, warmup: int = 3, rounds: int = 3) -> float:
        """Min over ``rounds`` of the mean time of ``run_once``, HIP events on the launch
        stream."""
        s = L.stream_ptr()
        for _ in range(warmup):
            self.run_once(s)
        best = float("inf")
        for _ in range(rounds):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                self.run_once(s)
            e1.record()
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1) / reps)
        return best

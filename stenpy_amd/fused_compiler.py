"""The reference's fused-path boundary, served by the hipRTC fused executor.

The reference's ``Runtime.run`` reaches fusion through two names of its ``fused_compiler``
module (stenpy_engine.py:46-51, 956-981): ``FUSIBLE_OPS`` (fused_compiler.py:194), consulted by
``_should_fuse`` (any node whose op is in the set makes the graph a fusion candidate), and
``compile_and_execute(graph, runtime, seed) -> Dict[str, Tensor]`` (fused_compiler.py:632),
called by ``_run_fused``.  Importing this module under the name ``fused_compiler`` before the reference's
engine sends those calls to this backend:

* the reference's ``Graph`` (or this package's) is read through the public node fields
  (``_nodes``: id -> node with ``op_name``, ``input_ids``, ``params``) and copied, ids kept;
* the graph runs through ``FusedExecutor`` on ``runtime.device``; ops the planner does not fuse
  run on their plain HIP kernels in the same pass;
* the result holds the seed, every source constant and every sink, as the reference's
  executor returns (stenpy_engine.py:983-1146).  Fields are computed in FP32; when the graph's
  source fields are FP64 (the reference's dtype), the returned field-valued sinks are widened
  to FP64 so that reference code downstream, which asserts FP64, accepts them;
* a non-HIP device raises ``ValueError`` like fused_compiler.py:634-636, so the reference's
  ``Runtime.run`` warns and falls back to its own eager path.
"""
from __future__ import annotations

from typing import Any, Dict, Optional

import torch

from . import fused as _F
from .graph import Graph

# Every op the fused planner can place in a kernel (the reference's set is a subset).
FUSIBLE_OPS = frozenset(_F.UNARY_OPS - {"cast_bf16"}) | _F.BINARY_OPS | _F.STENCIL_OPS | \
    _F.TENSOR_OPS | _F.REDUCTION_OPS | _F.CHANNEL_OPS

_executors: Dict[Any, Any] = {}


def _as_graph(graph) -> Graph:
    if isinstance(graph, Graph):
        return graph
    nodes = getattr(graph, "_nodes", None)
    if nodes is None:
        raise TypeError("compile_and_execute: expected a Stenpy Graph (with _nodes)")
    g = Graph()
    for nid, node in nodes.items():
        g.add(node.op_name, tuple(node.input_ids), dict(node.params), node_id=nid)
    return g


def _runtime_for(device: torch.device):
    """One of this package's runtimes per device (plans and code objects are cached in it)."""
    from .runtime import Runtime
    key = (device.type, device.index)
    rt = _executors.get(key)
    if rt is None:
        rt = Runtime(device=device, fusion_enabled=True)
        _executors[key] = rt
    return rt


def compile_and_execute(graph, runtime, seed: Optional[Dict[str, Any]] = None) -> Dict[str, Any]:
    """Evaluate ``graph`` on the GPU with fused HIP kernels (fused_compiler.py:632)."""
    device = torch.device(getattr(runtime, "device", "cuda"))
    if device.type != "cuda":
        raise ValueError("Fused execution requires a CUDA device. "
                         "Use Runtime(device='cuda') or let the Runtime fall back to original.")
    if device.index is None:
        device = torch.device("cuda", torch.cuda.current_device())
    g = _as_graph(graph)
    rt = _runtime_for(device)
    wanted = [s for s in g.sinks() if g.node(s).op_name != "_constant"]
    res = rt.fused.run(g, rt, seed, wanted=wanted)
    wide = any(isinstance(n.params.get("value"), torch.Tensor) and n.params["value"].ndim >= 3
               and n.params["value"].dtype == torch.float64
               for n in g.nodes.values() if n.op_name == "_constant")
    out: Dict[str, Any] = dict(seed or {})
    for nid, node in g.nodes.items():
        if node.op_name == "_constant" and nid not in out:
            out[nid] = node.params.get("value")
    for nid in wanted:
        v = res[nid]
        if wide and isinstance(v, torch.Tensor) and v.dtype in (torch.float32, torch.bfloat16):
            v = v.to(torch.float64)
        out[nid] = v
    return out

"""Evaluate a Stenpy ``Graph`` on the CPU with the restated operators (oracle/ops.py).

TEST INFRASTRUCTURE ONLY (see oracle/ops.py).  Follows the eager executor's semantics
(stenpy_engine.py:983-1146): nodes in topological order, ``_constant`` nodes yield their
value, every other node calls the restated op with ``(*inputs, **params)``; 0-d constants
keep their FP64 value and broadcast like torch scalars.  ``dtype`` selects the FP64 oracle or
the FP32 twin; field inputs are upcast from the FP32 values the GPU consumed.
"""
from __future__ import annotations

from typing import Any, Dict, Iterable, Optional

import torch

from . import ops as O

_NO_KW = {"alloc", "mm"}


def to_cpu(v: Any, dtype: torch.dtype) -> Any:
    if isinstance(v, torch.Tensor):
        t = v.detach().to("cpu")
        if t.ndim == 0 or t.numel() == 1 and t.ndim <= 1:
            return t.to(torch.float64) if dtype == torch.float64 else t.to(torch.float64)
        return t.to(dtype)
    return v


def evaluate(graph, dtype: torch.dtype = torch.float64, seed: Optional[Dict[str, Any]] = None,
             keep: Optional[Iterable[str]] = None) -> Dict[str, Any]:
    """Return every node's value (CPU tensors of ``dtype``)."""
    res: Dict[str, Any] = {}
    for k, v in (seed or {}).items():
        res[k] = to_cpu(v, dtype)
    for node in graph.topological_sort():
        if node.id in res:
            continue
        if node.op_name == "_constant":
            res[node.id] = to_cpu(node.params["value"], dtype)
            continue
        fn = O.OPS[node.op_name]
        ins = []
        for i in node.input_ids:
            v = res[i]
            # 0-d constants are FP64 in the reference and broadcast without promotion
            if isinstance(v, torch.Tensor) and v.ndim == 0 and dtype != torch.float64:
                v = v.to(dtype)
            ins.append(v)
        kw = {k: v for k, v in node.params.items() if k not in _NO_KW}
        out = fn(*ins, **kw)
        res[node.id] = out
    return res

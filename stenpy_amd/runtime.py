"""Executors: the eager per-operator path and the fused path, behind ``Runtime.run``.

``Runtime.run(graph, seed)`` keeps the reference contract (stenpy_engine.py:945-1146): it
returns a dict holding the seed, every source constant and every sink; inputs are released
once their last consumer has run.  The eager path (``_run_original``) calls
``OP_REGISTRY[op](*inputs, **params)`` in topological order -- one HIP kernel (or a few) per
node, the "plain per-operator" path.  The fused path compiles the graph into hipRTC kernels
(``stenpy_amd.fused``).  ``OPS_FUSION_ENABLED`` selects fusion (default on, as
stenpy_engine.py:46); a fused failure falls back to the eager path with a warning, as
stenpy_engine.py:949-954 does, unless ``STENPY_FUSED_STRICT=1``.
"""
from __future__ import annotations

import inspect
import os
import warnings
from typing import Any, Dict, List, Optional, Sequence, Union

import numpy as np
import torch

from .graph import Graph
from .registry import OP_METADATA, OP_REGISTRY


class LazyFieldProtocol:
    """Marker base class of lazily read fields (stenpy_engine.py:56-80)."""
    __slots__ = ()


def is_lazy_field(x: Any) -> bool:
    return isinstance(x, LazyFieldProtocol)


_LAZY_FIELD_FACTORY = None


def set_lazy_field_factory(factory) -> None:
    global _LAZY_FIELD_FACTORY
    _LAZY_FIELD_FACTORY = factory


def _env_flag(name: str, default: str) -> bool:
    return os.environ.get(name, default).lower() in ("1", "true", "yes")


class MemoryManager:
    """Allocation accounting shim.

    The reference's MemoryManager (memory_manager.py:219-568) pools and zero-fills every
    buffer and spills through zict tiers.  Here the torch caching allocator owns device
    memory and kernels write every output element, so no zero fill is needed; this object
    keeps the ``stats()`` surface and peak-usage numbers.
    """

    def __init__(self, device: Union[str, torch.device] = "cuda", **_: Any) -> None:
        self.device = torch.device(device)

    def stats(self) -> Dict[str, Any]:
        if self.device.type == "cuda" and torch.cuda.is_available():
            return {"allocated_bytes": torch.cuda.memory_allocated(self.device),
                    "peak_bytes": torch.cuda.max_memory_allocated(self.device)}
        return {"allocated_bytes": 0, "peak_bytes": 0}

    def clear_pool(self, full: bool = False) -> None:
        if full and torch.cuda.is_available():
            torch.cuda.empty_cache()


class Runtime:
    """Graph executor (stenpy_engine.py:505-604, 945-1146)."""

    def __init__(self, mm: Optional[MemoryManager] = None,
                 device: Union[str, torch.device] = "cuda", skip_pool: bool = False,
                 fusion_enabled: Optional[bool] = None, schedule_fn=None) -> None:
        dev = torch.device(device)
        if dev.type == "cuda" and dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device() if torch.cuda.is_available() else 0)
        if dev.type != "cuda":
            raise RuntimeError("stenpy_amd executes graphs on a HIP device; device must be 'cuda'")
        self.device = dev
        self.mm = mm or MemoryManager(dev)
        self.skip_pool = skip_pool
        self.fusion_enabled = (_env_flag("OPS_FUSION_ENABLED", "1")
                               if fusion_enabled is None else bool(fusion_enabled))
        self._fused = None
        self._schedule_fn = schedule_fn
        self.last_path = None

    # -- helpers ---------------------------------------------------------------------------
    def _materialize(self, v: Any) -> Any:
        """Move a source value to the device as FP32 (scalars stay host numbers)."""
        if isinstance(v, LazyFieldProtocol):
            v = torch.from_numpy(np.ascontiguousarray(v[:], dtype=np.float32))
        if isinstance(v, np.ndarray):
            v = torch.from_numpy(np.ascontiguousarray(v))
        if isinstance(v, torch.Tensor):
            if v.numel() == 1 and v.ndim <= 1 and not v.is_cuda:
                return v
            if v.dtype in (torch.float64, torch.float16, torch.bfloat16) or not v.is_cuda:
                v = v.to(device=self.device, dtype=torch.float32)
            return v
        return v

    @property
    def fused(self):
        if self._fused is None:
            from .fused import FusedExecutor
            self._fused = FusedExecutor(self.device, self._schedule_fn)
        return self._fused

    # -- entry points -----------------------------------------------------------------------
    def run(self, graph: Graph, seed: Optional[Dict[str, Any]] = None,
            outputs: Optional[Sequence[str]] = None,
            out_dtype: Optional[torch.dtype] = None) -> Dict[str, Any]:
        """Evaluate ``graph``; ``outputs`` (extension) names the nodes to keep besides sinks."""
        with torch.no_grad():
            if self.fusion_enabled:
                try:
                    res = self.fused.run(graph, self, seed, wanted=outputs, out_dtype=out_dtype)
                    self.last_path = "fused"
                    return res
                except Exception as e:
                    if _env_flag("STENPY_FUSED_STRICT", "0"):
                        raise
                    warnings.warn(f"Fusion failed, falling back to original: {e}")
            self.last_path = "eager"
            return self._run_original(graph, seed, outputs, out_dtype)

    def run_eager(self, graph: Graph, seed=None, outputs=None, out_dtype=None) -> Dict[str, Any]:
        with torch.no_grad():
            self.last_path = "eager"
            return self._run_original(graph, seed, outputs, out_dtype)

    def _run_original(self, graph: Graph, seed: Optional[Dict[str, Any]] = None,
                      outputs: Optional[Sequence[str]] = None,
                      out_dtype: Optional[torch.dtype] = None) -> Dict[str, Any]:
        results: Dict[str, Any] = dict(seed or {})
        keep = set(outputs or ())
        consumers = graph.build_consumer_map()
        refs = {nid: len(c) for nid, c in consumers.items()}
        for node in graph.topological_sort():
            if node.id in results:
                continue
            if node.op_name == "_constant":
                results[node.id] = self._materialize(node.params.get("value"))
                continue
            ins = []
            for i in node.input_ids:
                if i not in results:
                    raise RuntimeError(f"Input '{i}' for node '{node.id}' not found in results")
                ins.append(self._materialize(results[i]))
            kwargs = {k: (self._materialize(v) if isinstance(v, LazyFieldProtocol) else v)
                      for k, v in node.params.items()}
            fn = OP_REGISTRY[node.op_name]
            try:
                params = inspect.signature(fn).parameters
            except (TypeError, ValueError):
                params = {}
            if "alloc" in params:
                kwargs["alloc"] = None
            if "mm" in params:
                kwargs["mm"] = self.mm
            out = fn(*ins, **kwargs)
            if (out_dtype is not None and isinstance(out, torch.Tensor) and out.ndim >= 3
                    and not consumers[node.id]):
                from . import ops
                out = ops._unary("copy", out, out_dtype=out_dtype)
            results[node.id] = out
            for dep in node.input_ids:
                refs[dep] -= 1
                if refs[dep] <= 0 and dep in results and dep not in keep:
                    if graph.node(dep).op_name != "_constant":
                        results.pop(dep)
        return results

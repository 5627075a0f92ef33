"""Operator registry: the plugin boundary every executor dispatches through.

``register_operator`` keeps the reference's signature and metadata keys
(stenpy_engine.py:1184-1212): ``stencil_radius``, ``halo_left``/``halo_right``, ``cost``,
``exchange_dims``, ``flops_per_element``, ``bytes_read_per_element``,
``bytes_written_per_element`` and ``supports_fp64_fusion``.  The eager executor calls
``OP_REGISTRY[name](*inputs, **params)`` and adds ``alloc=``/``mm=`` when the callable
declares them (stenpy_engine.py:1074-1081).

One key is added to the metadata, ``kind``, which tells the fused planner how an op can be
evaluated per voxel: ``"source"``, ``"pointwise"``, ``"stencil"``, ``"reduction"``,
``"tensor"`` (per-voxel 3x3 algebra) or ``"opaque"`` (never fused).
"""
from __future__ import annotations

from typing import Callable, Dict, List, Optional

OP_REGISTRY: Dict[str, Callable] = {}
OP_METADATA: Dict[str, dict] = {}


def register_operator(name: str, func: Callable, radius: int = 0, halo_l: int = 0,
                      halo_r: int = 0, cost: str = "low",
                      exchange_dims: Optional[List[int]] = None, flops: int = 0,
                      bytes_in: int = 0, bytes_out: int = 0, fp64_fusion: bool = True,
                      kind: str = "opaque") -> None:
    """Register ``func`` under ``name`` (stenpy_engine.py:1188-1212)."""
    OP_REGISTRY[name] = func
    OP_METADATA[name] = {
        "stencil_radius": radius,
        "halo_left": halo_l if halo_l else radius,
        "halo_right": halo_r if halo_r else radius,
        "cost": cost,
        "exchange_dims": exchange_dims,
        "flops_per_element": flops,
        "bytes_read_per_element": bytes_in,
        "bytes_written_per_element": bytes_out,
        "supports_fp64_fusion": fp64_fusion,
        "kind": kind,
    }


def op_kind(name: str) -> str:
    return OP_METADATA.get(name, {}).get("kind", "opaque")

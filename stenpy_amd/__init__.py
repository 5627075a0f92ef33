"""stenpy_amd: an MI355X (gfx950) execution backend for Stenpy.

Public names follow the reference package (stenpy.py:986-997 / stenpy_engine.py:3780-3790):
``Graph``, ``GraphNode``, ``Runtime``, ``register_operator``, ``OP_REGISTRY``,
``OP_METADATA``, ``compile_expression``, ``parse_expression``, ``Pipeline``,
``LazyFieldProtocol`` ...  ``stenpy_amd.<op>(expr)`` returns a Pipeline exactly like
``stenpy.<op>(expr)`` (stenpy.py:967-984).
"""
from __future__ import annotations

from .graph import Graph, GraphNode, _graph_sinks
from .registry import OP_METADATA, OP_REGISTRY, register_operator
from . import ops as _ops  # registers every operator
from .compiler import UserVar, compile_expression, parse_expression
from .runtime import (LazyFieldProtocol, MemoryManager, Runtime, is_lazy_field,
                      set_lazy_field_factory)
from .pipeline import Pipeline, PerfStats, log_event, set_logging, stream_execute

__all__ = [
    "Graph", "GraphNode", "Runtime", "MemoryManager", "register_operator", "OP_REGISTRY",
    "OP_METADATA", "compile_expression", "parse_expression", "UserVar", "LazyFieldProtocol",
    "is_lazy_field", "set_lazy_field_factory", "Pipeline", "PerfStats", "log_event",
    "set_logging", "stream_execute",
]


class _OperatorDispatch:
    """``stenpy_amd.<op>(expr, **settings)`` -> Pipeline (stenpy.py:967-976)."""

    def __init__(self, op_name: str) -> None:
        self._op_name = op_name

    def __call__(self, expr, *args, **kwargs) -> Pipeline:
        p = Pipeline(self._op_name, expr if isinstance(expr, str) else str(expr))
        if kwargs:
            p.set(**kwargs)
        return p


def __getattr__(name: str):
    if name in OP_REGISTRY:
        return _OperatorDispatch(name)
    raise AttributeError(f"module 'stenpy_amd' has no attribute {name!r}")

"""Graph IR: the DAG of operator nodes that every executor consumes.

Behaviour follows the reference IR (stenpy_engine.py:351-451):
  * node ids are ``"<op>_<n>"`` from one process-wide counter (stenpy_engine.py:88-91);
  * ``topological_sort`` is Kahn's algorithm and raises ``ValueError`` on an unknown input id
    or a cycle (stenpy_engine.py:384-407);
  * ``clone_with_replacement`` swaps one source node for a ``_constant`` holding a chunk
    tensor and keeps every id (stenpy_engine.py:410-419);
  * shape/dtype side tables (stenpy_engine.py:428-437) and the consumer map (439-444).

Additions for the GPU executors: ``sinks()``, ``dead_code_eliminate()`` (the reference executes
dead nodes, e.g. the gradient left behind by the divergence(gradient) rewrite) and
``canonical_key()``, a structural hash used to cache fused plans.  The reference's fused
cache hashes only the set of op names (fused_compiler.py:537-543), which collides; the key
here covers the whole DAG.
"""
from __future__ import annotations

import collections
import hashlib
import itertools
from dataclasses import dataclass, field as dc_field
from typing import Any, Dict, Iterable, List, Optional, Tuple

import torch

_call_counter = itertools.count()


def _uid(prefix: str = "t") -> str:
    """Process-unique node id ``<prefix>_<n>`` (stenpy_engine.py:88-91)."""
    return f"{prefix}_{next(_call_counter)}"


@dataclass(frozen=True)
class GraphNode:
    """One operator application (stenpy_engine.py:351-356)."""
    id: str
    op_name: str
    input_ids: Tuple[str, ...] = dc_field(default_factory=tuple)
    params: Dict[str, Any] = dc_field(default_factory=dict)


class Graph:
    """Insertion-ordered DAG of :class:`GraphNode` (stenpy_engine.py:359-444)."""

    def __init__(self) -> None:
        self._nodes: Dict[str, GraphNode] = {}
        self._order: List[str] = []
        self._dtypes: Dict[str, torch.dtype] = {}
        self._shapes: Dict[str, Tuple[int, ...]] = {}

    # -- construction ---------------------------------------------------------------------
    def add(self, op_name: str, input_ids: Iterable[str] = (),
            params: Optional[Dict[str, Any]] = None, node_id: Optional[str] = None) -> str:
        nid = node_id or _uid(op_name)
        self._nodes[nid] = GraphNode(id=nid, op_name=op_name, input_ids=tuple(input_ids),
                                     params=dict(params or {}))
        if nid not in self._order:
            self._order.append(nid)
        return nid

    # -- traversal ------------------------------------------------------------------------
    def topological_sort(self) -> List[GraphNode]:
        in_deg = {nid: 0 for nid in self._nodes}
        children: Dict[str, List[str]] = {nid: [] for nid in self._nodes}
        for nid, node in self._nodes.items():
            for dep in node.input_ids:
                if dep not in self._nodes:
                    raise ValueError(f"Node '{nid}' references unknown input '{dep}'")
                in_deg[nid] += 1
                children[dep].append(nid)
        queue = collections.deque(nid for nid, d in in_deg.items() if d == 0)
        order: List[GraphNode] = []
        while queue:
            nid = queue.popleft()
            order.append(self._nodes[nid])
            for child in children[nid]:
                in_deg[child] -= 1
                if in_deg[child] == 0:
                    queue.append(child)
        if len(order) != len(self._nodes):
            raise ValueError("Graph contains a cycle")
        return order

    def clone_with_replacement(self, source_id: str, tensor: Any) -> "Graph":
        g = Graph()
        for nid in self._order:
            node = self._nodes[nid]
            if nid == source_id:
                g.add("_constant", (), {"value": tensor}, node_id=nid)
            else:
                g.add(node.op_name, node.input_ids, node.params, node_id=nid)
        return g

    def build_consumer_map(self) -> Dict[str, List[str]]:
        consumers: Dict[str, List[str]] = {nid: [] for nid in self._nodes}
        for nid, node in self._nodes.items():
            for dep in node.input_ids:
                consumers[dep].append(nid)
        return consumers

    def sinks(self) -> List[str]:
        """Nodes nobody consumes, in insertion order (stenpy_engine.py:447-451)."""
        consumed = set()
        for node in self._nodes.values():
            consumed.update(node.input_ids)
        return [nid for nid in self._order if nid not in consumed]

    # -- optimisation ---------------------------------------------------------------------
    def dead_code_eliminate(self, keep: Iterable[str]) -> "Graph":
        """Return a copy holding only the ancestors of ``keep`` (ids preserved)."""
        live = set()
        stack = list(keep)
        while stack:
            nid = stack.pop()
            if nid in live:
                continue
            if nid not in self._nodes:
                raise ValueError(f"unknown node '{nid}'")
            live.add(nid)
            stack.extend(self._nodes[nid].input_ids)
        g = Graph()
        for nid in self._order:
            if nid in live:
                n = self._nodes[nid]
                g.add(n.op_name, n.input_ids, n.params, node_id=nid)
                if nid in self._shapes:
                    g._shapes[nid] = self._shapes[nid]
                    g._dtypes[nid] = self._dtypes[nid]
        return g

    def canonical_key(self, extra: Any = None) -> str:
        """Structural hash over ops, edges, scalar params and source-tensor shapes/dtypes."""
        h = hashlib.sha1()
        idx = {}
        for i, node in enumerate(self.topological_sort()):
            idx[node.id] = i
            parts = [node.op_name, ",".join(str(idx[d]) for d in node.input_ids)]
            for k in sorted(node.params):
                v = node.params[k]
                if isinstance(v, torch.Tensor):
                    if v.ndim == 0 or v.numel() == 1:
                        parts.append(f"{k}=S{float(v.reshape(-1)[0]):.17g}")
                    else:
                        parts.append(f"{k}=T{tuple(v.shape)}{v.dtype}")
                else:
                    parts.append(f"{k}={v!r}")
            h.update(("|".join(parts) + ";").encode())
        if extra is not None:
            h.update(repr(extra).encode())
        return h.hexdigest()

    # -- side tables (stenpy_engine.py:428-437) -------------------------------------------
    def set_node_shape(self, node_id: str, shape: Tuple[int, ...],
                       dtype: torch.dtype = torch.float32) -> None:
        self._shapes[node_id] = tuple(shape)
        self._dtypes[node_id] = dtype

    def get_node_shape(self, node_id: str) -> Optional[Tuple[int, ...]]:
        return self._shapes.get(node_id)

    def get_node_dtype(self, node_id: str) -> torch.dtype:
        return self._dtypes.get(node_id, torch.float32)

    def node(self, nid: str) -> GraphNode:
        return self._nodes[nid]

    @property
    def nodes(self) -> Dict[str, GraphNode]:
        return self._nodes

    def __len__(self) -> int:
        return len(self._nodes)

    def __contains__(self, nid: str) -> bool:
        return nid in self._nodes

    def __repr__(self) -> str:
        return f"Graph({len(self._nodes)} nodes)"


def _graph_sinks(graph: Graph) -> List[str]:
    return graph.sinks()

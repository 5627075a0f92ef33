"""Graph IR: the DAG of operator nodes that every executor consumes.

Public surface kept from the reference IR (stenpy_engine.py:351-451), because drop-in callers
use it: ``GraphNode(id, op_name, input_ids, params)``; ``Graph.add(op_name, input_ids=(),
params=None, node_id=None) -> id`` with generated ids of the form ``"<op>_<n>"``
(stenpy_engine.py:88-91); ``topological_sort()`` raising ``ValueError`` on an unknown input id or
a cycle; ``clone_with_replacement(source_id, tensor)``; the shape/dtype side table;
``build_consumer_map()``; ``_graph_sinks``.  Reference-side code also reads ``graph._nodes``,
``graph._order`` and ``graph._shapes`` directly (fused_compiler.py:237-726,
stenpy_engine.py:449-457, 961-968), so those names exist here as read-only views.

Design.  Nodes live in one insertion-ordered table, and the reverse edges (who consumes a node)
are kept up to date as nodes are added, so the consumer map, the sink list and dead-code
elimination never rescan the graph.  The topological order is a depth-first post-order over the
insertion order: a graph whose nodes were added after their inputs (everything the compiler
builds) comes back in insertion order, and any other valid DAG still comes back with every node
after its inputs.

Additions for the GPU executors: ``sinks()``, ``dead_code_eliminate()`` (the reference executes
dead nodes, e.g. the gradient left behind by the divergence(gradient) rewrite) and
``canonical_key()``, a structural hash used to cache fused plans.  The reference's fused cache
hashes only the set of op names (fused_compiler.py:537-543), which collides; the key here covers
the whole DAG.
"""
from __future__ import annotations

import hashlib
import itertools
import threading
from dataclasses import dataclass, field as dc_field
from typing import Any, Dict, Iterable, List, Optional, Tuple

import torch


class _IdSource:
    """Process-wide source of node ids ``<op>_<n>``; ``n`` never repeats within a process."""

    def __init__(self) -> None:
        self._next = itertools.count()
        self._lock = threading.Lock()

    def __call__(self, op_name: str) -> str:
        with self._lock:
            return f"{op_name}_{next(self._next)}"


_fresh_id = _IdSource()


def _uid(prefix: str = "t") -> str:
    """A new process-unique id ``<prefix>_<n>``."""
    return _fresh_id(prefix)


@dataclass(frozen=True)
class GraphNode:
    """One operator application: ``op_name(*input_ids, **params)``."""
    id: str
    op_name: str
    input_ids: Tuple[str, ...] = dc_field(default_factory=tuple)
    params: Dict[str, Any] = dc_field(default_factory=dict)


class Graph:
    """Insertion-ordered DAG of :class:`GraphNode`, with reverse edges kept incrementally."""

    def __init__(self) -> None:
        self._table: Dict[str, GraphNode] = {}
        self._users: Dict[str, List[str]] = {}
        self._meta: Dict[str, Tuple[Tuple[int, ...], torch.dtype]] = {}

    # -- construction ---------------------------------------------------------------------
    def add(self, op_name: str, input_ids: Iterable[str] = (),
            params: Optional[Dict[str, Any]] = None, node_id: Optional[str] = None) -> str:
        """Insert a node (or replace the node of the same id in place) and return its id."""
        node = GraphNode(node_id or _fresh_id(op_name), op_name, tuple(input_ids),
                         dict(params or {}))
        old = self._table.get(node.id)
        if old is not None:
            for src in old.input_ids:
                self._users[src].remove(old.id)
        self._table[node.id] = node
        for src in node.input_ids:
            self._users.setdefault(src, []).append(node.id)
        return node.id

    # -- read-only views kept for reference-side callers ------------------------------------
    @property
    def _nodes(self) -> Dict[str, GraphNode]:
        return self._table

    @property
    def _order(self) -> List[str]:
        return list(self._table)

    @property
    def _shapes(self) -> Dict[str, Tuple[int, ...]]:
        return {nid: m[0] for nid, m in self._meta.items()}

    @property
    def _dtypes(self) -> Dict[str, torch.dtype]:
        return {nid: m[1] for nid, m in self._meta.items()}

    @property
    def nodes(self) -> Dict[str, GraphNode]:
        return self._table

    def node(self, nid: str) -> GraphNode:
        return self._table[nid]

    # -- traversal ------------------------------------------------------------------------
    def topological_sort(self) -> List[GraphNode]:
        """Every node after all of its inputs; ``ValueError`` on an unknown input or a cycle."""
        OPEN, DONE = 1, 2
        mark: Dict[str, int] = {}
        out: List[GraphNode] = []
        for root in self._table:
            if root in mark:
                continue
            mark[root] = OPEN
            path = [(root, iter(self._table[root].input_ids))]
            while path:
                nid, pending = path[-1]
                src = next(pending, None)
                if src is None:
                    path.pop()
                    mark[nid] = DONE
                    out.append(self._table[nid])
                    continue
                if src not in self._table:
                    raise ValueError(f"node '{nid}' has an unknown input '{src}'")
                state = mark.get(src)
                if state == OPEN:
                    raise ValueError(f"dependency cycle through node '{src}'")
                if state is None:
                    mark[src] = OPEN
                    path.append((src, iter(self._table[src].input_ids)))
        return out

    def clone_with_replacement(self, source_id: str, tensor: Any) -> "Graph":
        """A copy in which node ``source_id`` becomes a ``_constant`` holding ``tensor``.

        Every id is kept, so results of the copy are keyed like the original's (the chunk
        streamer swaps a field for one slab this way)."""
        g = Graph()
        for nid, node in self._table.items():
            if nid == source_id:
                g.add("_constant", (), {"value": tensor}, node_id=nid)
            else:
                g.add(node.op_name, node.input_ids, node.params, node_id=nid)
        return g

    def build_consumer_map(self) -> Dict[str, List[str]]:
        """node id -> ids of the nodes that read it, in insertion order."""
        return {nid: list(self._users.get(nid, ())) for nid in self._table}

    def sinks(self) -> List[str]:
        """Nodes nobody consumes, in insertion order (stenpy_engine.py:447-451)."""
        return [nid for nid in self._table if not self._users.get(nid)]

    # -- optimisation ---------------------------------------------------------------------
    def dead_code_eliminate(self, keep: Iterable[str]) -> "Graph":
        """Return a copy holding only the ancestors of ``keep`` (ids preserved)."""
        live = set()
        stack = list(keep)
        while stack:
            nid = stack.pop()
            if nid in live:
                continue
            if nid not in self._table:
                raise ValueError(f"unknown node '{nid}'")
            live.add(nid)
            stack.extend(self._table[nid].input_ids)
        g = Graph()
        for nid, n in self._table.items():
            if nid in live:
                g.add(n.op_name, n.input_ids, n.params, node_id=nid)
                if nid in self._meta:
                    g._meta[nid] = self._meta[nid]
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

    # -- shape/dtype side table -----------------------------------------------------------
    def set_node_shape(self, node_id: str, shape: Tuple[int, ...],
                       dtype: torch.dtype = torch.float32) -> None:
        self._meta[node_id] = (tuple(shape), dtype)

    def get_node_shape(self, node_id: str) -> Optional[Tuple[int, ...]]:
        m = self._meta.get(node_id)
        return None if m is None else m[0]

    def get_node_dtype(self, node_id: str) -> torch.dtype:
        """FP32 when unset: this backend stores fields in FP32 (the reference assumes FP64)."""
        m = self._meta.get(node_id)
        return torch.float32 if m is None else m[1]

    def __len__(self) -> int:
        return len(self._table)

    def __contains__(self, nid: str) -> bool:
        return nid in self._table

    def __repr__(self) -> str:
        return f"Graph({len(self._table)} nodes)"


def _graph_sinks(graph: Graph) -> List[str]:
    return graph.sinks()

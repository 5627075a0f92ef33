"""Chainable ``Pipeline`` API, logging and perf report (stenpy.py:54-95, 351-400, 617-964).

``Pipeline(op_name, expr)`` collects settings by attribute assignment: known settings
(``stream_i stream_o device log perf_report output boundary dx chunk_mb``), string values as
field paths, anything else as constants / in-memory fields (stenpy.py:664-676).  Setting
``output`` triggers ``run()``.  Defaults match the reference: boundary "periodic", dx 1.0
(stenpy.py:651-656).  The expression string, not ``op_name``, is what gets compiled
(stenpy.py:715-717).

Differences: fields are evaluated on the GPU in FP32; fields may be given as tensors or
NumPy arrays as well as paths; paths ending in ``.npy`` are read with NumPy, others with
h5py when it is importable (datasets ``data``/``field``, ``spacing``/``origin`` attrs).
"""
from __future__ import annotations

import datetime
import logging
import math
import os
import time
import warnings
from typing import Any, Dict, Optional, Tuple, Union

import numpy as np
import torch

from .compiler import compile_expression

_KNOWN_SETTINGS = {"stream_i", "stream_o", "device", "log", "perf_report", "output", "boundary",
                   "dx", "chunk_mb"}
_LOG_ENABLED = os.environ.get("STENPY_LOG", "1").lower() in ("1", "true", "yes")
_LOG_DIR = os.environ.get("STENPY_LOG_DIR", ".")
_logger = logging.getLogger("stenpy")


def set_logging(enabled: bool = True, log_dir: Optional[str] = None) -> None:
    """Toggle the daily event log (stenpy.py:92-95)."""
    global _LOG_ENABLED, _LOG_DIR
    _LOG_ENABLED = bool(enabled)
    if log_dir is not None:
        _LOG_DIR = log_dir


def log_event(msg: str) -> None:
    """Append ``msg`` to ``stenpy_<YYYY-MM-DD>.log`` (stenpy.py:85-90)."""
    if not _LOG_ENABLED:
        return
    try:
        path = os.path.join(_LOG_DIR, f"stenpy_{datetime.date.today().isoformat()}.log")
        with open(path, "a") as fh:
            fh.write(f"{datetime.datetime.now().isoformat()} {msg}\n")
    except OSError:
        pass


class PerfStats:
    """Wall-clock stage timing plus byte counts (stenpy.py:355-382)."""

    def __init__(self) -> None:
        self.t0 = self.t1 = 0.0
        self.bytes_read = 0
        self.bytes_written = 0
        self.voxels = 0
        self.kernel_ms = 0.0

    def start(self) -> None:
        self.t0 = time.perf_counter()

    def stop(self) -> None:
        self.t1 = time.perf_counter()

    @property
    def elapsed_s(self) -> float:
        return self.t1 - self.t0


def _print_perf_report(stats: PerfStats, op: str, device: str) -> None:
    el = max(stats.elapsed_s, 1e-12)
    print(f"stenpy perf: op={op} device={device} elapsed={el:.4f}s "
          f"voxels={stats.voxels} ({stats.voxels / el / 1e9:.3f} Gvoxel/s) "
          f"read={stats.bytes_read / 1e9:.3f} GB written={stats.bytes_written / 1e9:.3f} GB")


def _resolve_device(device: Union[str, torch.device]) -> torch.device:
    dev = torch.device(str(device))
    if dev.type == "cuda":
        if not torch.cuda.is_available():
            raise RuntimeError("device='cuda' was requested but torch.cuda.is_available() is False")
        return torch.device("cuda", dev.index if dev.index is not None else 0)
    raise RuntimeError("stenpy_amd runs graphs on the GPU; use device='cuda'")


def load_tensor(path: str, device: Union[str, torch.device] = "cuda"
                ) -> Tuple[torch.Tensor, Tuple[float, ...], Tuple[float, ...]]:
    """Read a field file into an FP32 device tensor; returns (tensor, spacing, origin)."""
    if path.endswith(".npy"):
        arr = np.load(path, mmap_mode="r")
        spacing, origin = (1.0,) * arr.ndim, (0.0,) * arr.ndim
    else:
        try:
            import h5py
        except ImportError as exc:
            raise RuntimeError(f"reading {path} needs h5py, which is not installed") from exc
        with h5py.File(path, "r") as f:
            name = next((c for c in ("data", "field") if c in f), None) or next(iter(f))
            arr = f[name][...]
            spacing = tuple(float(v) for v in f.attrs.get("spacing", [1.0] * arr.ndim))
            origin = tuple(float(v) for v in f.attrs.get("origin", [0.0] * len(spacing)))
    t = torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float32)).to(device)
    return t, spacing, origin


def save_tensor(t: torch.Tensor, spacing, origin, path: str) -> None:
    arr = t.detach().float().cpu().numpy()
    if path.endswith(".npy"):
        np.save(path, arr)
        return
    try:
        import h5py
    except ImportError as exc:
        raise RuntimeError(f"writing {path} needs h5py, which is not installed") from exc
    with h5py.File(path, "w") as f:
        f.create_dataset("data", data=arr)
        f.attrs["spacing"] = list(spacing)
        f.attrs["origin"] = list(origin)


class Pipeline:
    """stenpy.py:645-964."""

    def __init__(self, op_name: str, expr_str: str) -> None:
        for k, v in (("_op_name", op_name), ("_expr_str", expr_str), ("_stream_i", False),
                     ("_stream_o", False), ("_device", "cuda"), ("_chunk_mb", None),
                     ("_log", True), ("_perf_report", False), ("_boundary", "periodic"),
                     ("_dx", 1.0), ("_field_paths", {}), ("_constants", {}), ("_output", None),
                     ("_result", None), ("_ran", False)):
            object.__setattr__(self, k, v)

    def __setattr__(self, name: str, value: Any) -> None:
        if name in _KNOWN_SETTINGS:
            object.__setattr__(self, f"_{name}", value)
            if name == "output":
                self.run()
            return
        if name.startswith("_"):
            object.__setattr__(self, name, value)
            return
        if isinstance(value, str):
            self._field_paths[name] = value
        else:
            self._constants[name] = value

    def set(self, **kwargs) -> "Pipeline":
        for k, v in kwargs.items():
            setattr(self, k, v)
        return self

    def run(self) -> Any:
        if self._ran:
            return self._result
        object.__setattr__(self, "_ran", True)
        from .runtime import Runtime
        dev = _resolve_device(self._device)
        stats = PerfStats()
        stats.start()
        if self._log:
            log_event(f"pipeline start op={self._op_name} expr={self._expr_str!r} device={self._device}")
        field_map: Dict[str, Any] = {}
        spacing = origin = None
        lazy = {}
        for var, path in self._field_paths.items():
            if self._stream_i:
                from .stream import NpyLazyField
                lazy[var] = NpyLazyField(path)
                field_map[var] = torch.zeros((1,) + tuple(lazy[var].shape[1:]), dtype=torch.float32)
            else:
                t, spacing, origin = load_tensor(path, dev)
                field_map[var] = t
        for var, val in self._constants.items():
            if isinstance(val, (torch.Tensor, np.ndarray)) and np.ndim(val) >= 1:
                field_map[var] = torch.as_tensor(val)
            else:
                field_map.setdefault(var, torch.tensor(float(val), dtype=torch.float64))
        graph, sink, warns, resolved, _ = compile_expression(
            self._expr_str, dx=self._dx, boundary=self._boundary, field_map=field_map)
        for w in warns:
            warnings.warn(w)
        rt = Runtime(device=dev)
        if self._stream_i or self._stream_o:
            from .stream import stream_graph
            if len(lazy) != 1:
                raise NotImplementedError("stream_i/stream_o support a single streamed field")
            (var, lf), = lazy.items()
            if not self._output:
                raise ValueError("stream_o requires .output to be set to a path.")
            meta = stream_graph(graph, sink, resolved[var], lf, self._output, rt,
                                boundary=self._boundary, chunk_mb=self._chunk_mb)
            object.__setattr__(self, "_result", meta)
        else:
            out = rt.run(graph)[sink]
            if self._output:
                save_tensor(out, spacing or (1.0,), origin or (0.0,), self._output)
            object.__setattr__(self, "_result", out)
            if isinstance(out, torch.Tensor):
                stats.voxels = out.numel()
        torch.cuda.synchronize(dev)
        stats.stop()
        if self._log:
            log_event(f"pipeline done op={self._op_name} elapsed={stats.elapsed_s:.3f}s")
        if self._perf_report:
            _print_perf_report(stats, self._op_name, str(dev))
        return self._result

    @property
    def value(self) -> Any:
        return self.run()

    def __getitem__(self, idx):
        return self.run()[idx]

    def __repr__(self) -> str:
        if self._result is not None:
            return f"Pipeline({self._op_name}, result={self._result!r})"
        return f"Pipeline({self._op_name}({self._expr_str}), unconfigured)"


def stream_execute(*args, **kwargs):
    """Single-op streamed evaluation (stenpy.py:474-615); see ``stenpy_amd.stream``."""
    from .stream import stream_execute as _se
    return _se(*args, **kwargs)

"""ctypes binding of libstenpy_hip.so (declared in include/stenpy_hip.h).

Torch is plumbing here: it owns device memory and the current HIP stream; every compute
call goes through the C ABI.  If the library is missing or cannot be loaded the first call
raises ``RuntimeError`` -- there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path
from typing import List, Optional, Sequence

import numpy as np
import torch

LIB_PATH = Path(__file__).resolve().parent / "libstenpy_hip.so"
ABI_VERSION = 5

# enums mirrored from include/stenpy_hip.h
SP_F32, SP_BF16, SP_F64, SP_C64 = 0, 1, 2, 3
BOUNDARY = {"periodic": 0, "neumann": 1, "dirichlet": 2, "reflect": 3}
SP_ZEROPAD = 4
UNARY = {"neg": 0, "exp": 1, "log": 2, "sqrt": 3, "sin": 4, "cos": 5, "tanh": 6, "clamp": 7,
         "scale": 8, "copy": 9, "square": 10}
BINARY = {"add": 0, "sub": 1, "mul": 2, "div": 3}
D1_PERIODIC, D1_EDGE2, D1_EDGE1, D2_PAD = 0, 1, 2, 3
REDUCE = {"sum": 0, "mean": 1, "norm_l2": 2, "variance": 3, "variance_biased": 4,
          "min_max": 5, "integrate": 6, "any_small_diag": 7}
TENSOR3 = {"trace": 0, "determinant": 1, "eigenvalues": 2, "deviatoric": 3, "inverse": 4}

_DTYPE_CODE = {torch.float32: SP_F32, torch.bfloat16: SP_BF16, torch.float64: SP_F64,
               torch.complex64: SP_C64}


def boundary_code(name: str) -> int:
    return BOUNDARY.get(name, SP_ZEROPAD)


class SpTensor(C.Structure):
    _fields_ = [("ptr", C.c_void_p), ("dtype", C.c_int32), ("ndim", C.c_int32),
                ("shape", C.c_int64 * 6), ("stride", C.c_int64 * 6)]


def desc(t: torch.Tensor) -> SpTensor:
    if t.ndim > 6:
        raise ValueError("tensors of rank > 6 are not supported")
    if t.dtype not in _DTYPE_CODE:
        raise TypeError(f"unsupported dtype {t.dtype}")
    d = SpTensor()
    d.ptr = t.data_ptr() if t.numel() else None
    d.dtype = _DTYPE_CODE[t.dtype]
    d.ndim = t.ndim
    for i, (s, st) in enumerate(zip(t.shape, t.stride())):
        d.shape[i] = s
        d.stride[i] = st
    return d


_lib: Optional[C.CDLL] = None
_lock = threading.Lock()


def _declare(lib: C.CDLL) -> None:
    P, I, D, V, Z = C.POINTER(SpTensor), C.c_int, C.c_double, C.c_void_p, C.c_size_t
    sig = {
        "sp_abi_version": ([], I),
        "sp_last_error": ([], C.c_char_p),
        "sp_device_info": ([C.c_char_p, I, C.POINTER(I), C.POINTER(I)], I),
        "sp_op_unary": ([I, P, P, D, D, V], I),
        "sp_op_binary": ([I, P, P, P, D, V], I),
        "sp_op_dot": ([P, P, P, V], I),
        "sp_op_sqnorm": ([P, P, V], I),
        "sp_op_norm_last": ([P, P, D, V], I),
        "sp_op_normalize_last": ([P, P, D, V], I),
        "sp_axis_fir": ([P, I, P, I, C.POINTER(D), I, V], I),
        "sp_fft": ([P, P, I, I, V], I),
        "sp_spectral": ([P, P, C.POINTER(D), I, I, V], I),
        "sp_distance_transform": ([P, P, C.POINTER(D), V], I),
        "sp_cumulative_trapezoid": ([P, P, I, D, V], I),
        "sp_surface_integral": ([P, P, P, V, V], I),
        "sp_pair_stat": ([I, P, P, V, V], I),
        "sp_entropy": ([P, I, D, V, V, V], I),
        "sp_axis_pass": ([P, I, P, I, I, D, D, I, V], I),
        "sp_laplacian3d": ([P, P, C.POINTER(D), I, V], I),
        "sp_gradient3d": ([P, P, D, I, V], I),
        "sp_divergence3d": ([P, P, C.POINTER(D), I, V], I),
        "sp_curl3d": ([P, P, C.POINTER(D), V], I),
        "sp_gradient_nd3d": ([P, P, C.POINTER(D), V], I),
        "sp_hessian3d": ([P, P, C.POINTER(D), V], I),
        "sp_jacobian3d": ([P, P, C.POINTER(D), V], I),
        "sp_mean_curvature3d": ([P, P, C.POINTER(D), I, D, V], I),
        "sp_material_derivative3d": ([P, P, P, C.POINTER(D), V], I),
        "sp_reduce_workspace_bytes": ([C.c_int64], Z),
        "sp_reduce": ([I, P, V, V, C.POINTER(D), V], I),
        "sp_reduce_combine": ([I, V, I, C.c_int64, V, V], I),
        "sp_reduce_combine_batch": ([V, I, V], I),
        "sp_reduce_combine_multi": ([V, I, V], I),
        "sp_scalar_program": ([V, I, V], I),
        "sp_tensor3": ([I, P, P, V, V], I),
        "sp_gen_fourier": ([P, C.POINTER(C.c_int64), C.c_int64, C.POINTER(C.c_int32),
                            C.POINTER(D), C.POINTER(D), C.c_uint64, V], I),
        "sp_rtc_compile": ([C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(V), C.POINTER(Z),
                            C.c_char_p, Z], I),
        "sp_free_code": ([V], None),
        "sp_module_load": ([V, Z, C.c_char_p, C.POINTER(V)], I),
        "sp_module_launch": ([V, C.POINTER(C.c_uint), C.POINTER(C.c_uint), C.c_uint, V, V, Z], I),
        "sp_module_attributes": ([V, C.POINTER(I), C.POINTER(I), C.POINTER(I), C.POINTER(I)], I),
        "sp_module_unload": ([V], None),
        "sp_module_occupancy": ([V, I, C.c_size_t, C.POINTER(I)], I),
        "sp_host_gen_fourier": ([V, C.POINTER(C.c_int64), C.c_int64, C.c_int64, C.POINTER(C.c_int32),
                                 C.POINTER(D), C.POINTER(D), C.c_uint64, I], I),
        "sp_gemm_bf16": ([I, I, I, I, I, V, I, V, I, V, I, C.c_float, I, V], I),
        "sp_mlp_bias_act": ([V, V, I, I, I, V, V, V], I),
        "sp_mlp_relu_grad": ([V, V, C.c_int64, V, V, V], I),
        "sp_mlp_colsum": ([V, I, I, V, V], I),
        "sp_mlp_mse": ([V, V, I, V, V, V, V], I),
        "sp_mlp_adam": ([V, V, V, V, C.c_int64, C.c_float, C.c_float, C.c_float, C.c_float, I,
                         C.c_float, V, V], I),
        "sp_cast_f32_bf16": ([V, C.c_int64, V, V], I),
        "sp_probe_stream": ([I, I, V, V, V, V, C.c_int64, I, V], I),
        "sp_reduce_noalloc": ([I, P, V, C.POINTER(D), V], I),
        "sp_reduce_noalloc_scratch_bytes": ([], Z),
        "sp_probe_blocks": ([I, C.c_int64, I], C.c_int64),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res


def lib() -> C.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not LIB_PATH.exists():
                    raise RuntimeError(
                        f"{LIB_PATH} is missing: build it with `python -m stenpy_amd.build`")
                handle = C.CDLL(str(LIB_PATH))
                _declare(handle)
                v = handle.sp_abi_version()
                if v != ABI_VERSION:
                    raise RuntimeError(f"libstenpy_hip ABI {v}, expected {ABI_VERSION}; rebuild")
                _lib = handle
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().sp_last_error().decode(errors="replace")
        if rc == 1:
            raise ValueError(f"{what}: {msg}")
        if rc == 3:
            raise NotImplementedError(f"{what}: {msg}")
        raise RuntimeError(f"{what}: {msg} (status {rc})")


def stream_ptr(device: Optional[torch.device] = None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def require_gpu(*ts: torch.Tensor) -> None:
    for t in ts:
        if not t.is_cuda:
            raise RuntimeError(
                "stenpy_amd executes on the GPU only: tensors must live on a HIP device "
                f"(got {t.device})")


def _dx3(dx) -> C.Array:
    if isinstance(dx, (int, float)):
        vals = (float(dx),) * 3
    else:
        vals = tuple(float(v) for v in dx)
        if len(vals) != 3:
            raise ValueError(f"dx must have 3 entries, got {len(vals)}")
    return (C.c_double * 3)(*vals)


# ------------------------------------------------------------------------- wrappers
def unary(op: str, a: torch.Tensor, out: torch.Tensor, p0: float = 0.0, p1: float = 0.0) -> torch.Tensor:
    require_gpu(a, out)
    check(lib().sp_op_unary(UNARY[op], C.byref(desc(a)), C.byref(desc(out)), p0, p1,
                            stream_ptr(a.device)), f"unary {op}")
    return out


def binary(op: str, a: torch.Tensor, b: torch.Tensor, out: torch.Tensor, eps: float = 0.0) -> torch.Tensor:
    require_gpu(a, b, out)
    check(lib().sp_op_binary(BINARY[op], C.byref(desc(a)), C.byref(desc(b)), C.byref(desc(out)),
                             eps, stream_ptr(a.device)), f"binary {op}")
    return out


class SpScalarOp(C.Structure):
    """One op of ``sp_scalar_program`` (include/stenpy_hip.h)."""
    _fields_ = [("a", C.c_void_p), ("b", C.c_void_p), ("out", C.c_void_p),
                ("av", C.c_double), ("bv", C.c_double), ("p0", C.c_double), ("p1", C.c_double),
                ("binary", C.c_int32), ("op", C.c_int32)]


SCALAR_PROGRAM_MAX = 48


def scalar_program(ops: List[SpScalarOp], stream: Optional[int] = None, device=None) -> None:
    """Evaluate the scalar ops in order, SCALAR_PROGRAM_MAX per launch (later chunks see the
    earlier chunks' results, written in stream order)."""
    st = stream if stream is not None else stream_ptr(device)
    for i in range(0, len(ops), SCALAR_PROGRAM_MAX):
        chunk = ops[i:i + SCALAR_PROGRAM_MAX]
        arr = (SpScalarOp * len(chunk))(*chunk)
        check(lib().sp_scalar_program(C.addressof(arr), len(chunk), st), "scalar program")


def channel(kind: str, a: torch.Tensor, out: torch.Tensor, b: torch.Tensor = None,
            eps: float = 1e-15) -> torch.Tensor:
    require_gpu(a, out)
    L, s = lib(), stream_ptr(a.device)
    if kind == "dot":
        check(L.sp_op_dot(C.byref(desc(a)), C.byref(desc(b)), C.byref(desc(out)), s), "dot")
    elif kind == "sqnorm":
        check(L.sp_op_sqnorm(C.byref(desc(a)), C.byref(desc(out)), s), "sqnorm")
    elif kind == "norm_last":
        check(L.sp_op_norm_last(C.byref(desc(a)), C.byref(desc(out)), eps, s), "norm_last")
    else:
        check(L.sp_op_normalize_last(C.byref(desc(a)), C.byref(desc(out)), eps, s), "normalize")
    return out


def axis_pass(x: torch.Tensor, axis: int, out: torch.Tensor, mode: int, pad: int, h: float,
              scale: float = 1.0, accumulate: bool = False) -> torch.Tensor:
    require_gpu(x, out)
    check(lib().sp_axis_pass(C.byref(desc(x)), axis, C.byref(desc(out)), mode, pad, float(h),
                             float(scale), int(accumulate), stream_ptr(x.device)), "axis pass")
    return out


def axis_fir(x: torch.Tensor, axis: int, out: torch.Tensor, weights, pad: int) -> torch.Tensor:
    require_gpu(x, out)
    w = (C.c_double * len(weights))(*[float(v) for v in weights])
    check(lib().sp_axis_fir(C.byref(desc(x)), axis, C.byref(desc(out)), len(weights), w, pad,
                            stream_ptr(x.device)), "axis fir")
    return out


NORM = {"backward": 0, "ortho": 1, "forward": 2}
SPECTRAL_GRADIENT, SPECTRAL_LAPLACIAN = 0, 1


def fft(x: torch.Tensor, out: torch.Tensor, inverse: bool, norm: str) -> torch.Tensor:
    require_gpu(x, out)
    if norm not in NORM:
        raise ValueError(f"unknown FFT norm {norm!r} (backward, ortho or forward)")
    check(lib().sp_fft(C.byref(desc(x)), C.byref(desc(out)), int(bool(inverse)), NORM[norm],
                       stream_ptr(x.device)), "fft")
    return out


def cumulative_trapezoid(x: torch.Tensor, out: torch.Tensor, dim: int, dx: float) -> torch.Tensor:
    require_gpu(x, out)
    check(lib().sp_cumulative_trapezoid(C.byref(desc(x)), C.byref(desc(out)), int(dim), float(dx),
                                        stream_ptr(x.device)), "cumulative integral")
    return out


def surface_integral(t: torch.Tensor, normals: torch.Tensor, weights: torch.Tensor) -> torch.Tensor:
    require_gpu(t, normals, weights)
    res = torch.empty((), dtype=torch.float64, device=t.device)
    check(lib().sp_surface_integral(C.byref(desc(t)), C.byref(desc(normals)), C.byref(desc(weights)),
                                    C.c_void_p(res.data_ptr()), stream_ptr(t.device)), "surface integral")
    return res


PAIR_STAT = {"covariance": 0, "correlation": 1}


def pair_stat(kind: str, a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    require_gpu(a, b)
    res = torch.empty((), dtype=torch.float64, device=a.device)
    check(lib().sp_pair_stat(PAIR_STAT[kind], C.byref(desc(a)), C.byref(desc(b)), C.c_void_p(res.data_ptr()),
                             stream_ptr(a.device)), kind)
    return res


def entropy(x: torch.Tensor, dim: Optional[int], eps: float) -> torch.Tensor:
    require_gpu(x)
    if dim is None:
        res = torch.empty((), dtype=torch.float64, device=x.device)
        check(lib().sp_entropy(C.byref(desc(x)), -1, float(eps), C.c_void_p(res.data_ptr()), None,
                               stream_ptr(x.device)), "entropy")
        return res
    out = torch.empty(x.shape[:dim] + x.shape[dim + 1:], dtype=torch.float32, device=x.device)
    check(lib().sp_entropy(C.byref(desc(x)), int(dim), float(eps), None, C.byref(desc(out)),
                           stream_ptr(x.device)), "entropy")
    return out


def distance_transform(x: torch.Tensor, out: torch.Tensor, dx) -> torch.Tensor:
    require_gpu(x, out)
    d = (C.c_double * x.ndim)(*[float(v) for v in dx])
    check(lib().sp_distance_transform(C.byref(desc(x)), C.byref(desc(out)), d, stream_ptr(x.device)),
          "distance transform")
    return out


def spectral(x: torch.Tensor, out: torch.Tensor, dx, kind: int, dim: int = 0) -> torch.Tensor:
    require_gpu(x, out)
    d = (C.c_double * x.ndim)(*[float(v) for v in dx])
    check(lib().sp_spectral(C.byref(desc(x)), C.byref(desc(out)), d, int(kind), int(dim),
                            stream_ptr(x.device)), "spectral")
    return out


def stencil3d(name: str, x: torch.Tensor, out: torch.Tensor, dx=1.0, boundary: int = 0,
              v: torch.Tensor = None, eps: float = 1e-15) -> torch.Tensor:
    require_gpu(x, out)
    L, s = lib(), stream_ptr(x.device)
    dxi, dxo = C.byref(desc(x)), C.byref(desc(out))
    if name == "laplacian":
        rc = L.sp_laplacian3d(dxi, dxo, _dx3(dx), boundary, s)
    elif name == "gradient":
        rc = L.sp_gradient3d(dxi, dxo, float(dx), boundary, s)
    elif name == "divergence":
        rc = L.sp_divergence3d(dxi, dxo, _dx3(dx), boundary, s)
    elif name == "curl":
        rc = L.sp_curl3d(dxi, dxo, _dx3(dx), s)
    elif name == "gradient_nd":
        rc = L.sp_gradient_nd3d(dxi, dxo, _dx3(dx), s)
    elif name == "hessian":
        rc = L.sp_hessian3d(dxi, dxo, _dx3(dx), s)
    elif name == "jacobian":
        rc = L.sp_jacobian3d(dxi, dxo, _dx3(dx), s)
    elif name == "mean_curvature":
        rc = L.sp_mean_curvature3d(dxi, dxo, _dx3(dx), boundary, eps, s)
    elif name == "material_derivative":
        rc = L.sp_material_derivative3d(dxi, C.byref(desc(v)), dxo, _dx3(dx), s)
    else:
        raise KeyError(name)
    check(rc, name)
    return out


_ws_cache = {}


def _workspace(device: torch.device) -> torch.Tensor:
    key = (device.type, device.index)
    ws = _ws_cache.get(key)
    if ws is None:
        nbytes = lib().sp_reduce_workspace_bytes(0)
        ws = torch.empty(nbytes // 8, dtype=torch.float64, device=device)
        _ws_cache[key] = ws
    return ws


def reduce(kind: str, x: torch.Tensor, wparams: Sequence[float] = None) -> torch.Tensor:
    """Device FP64 result (0-d, or shape (2,) for min_max)."""
    require_gpu(x)
    if not x.is_contiguous():
        x = x.contiguous()
    out = torch.empty(2 if kind == "min_max" else 1, dtype=torch.float64, device=x.device)
    wp = (C.c_double * 4)(*(wparams or (1.0, 1.0, 1.0, 0.0)))
    ws = _workspace(x.device)
    check(lib().sp_reduce(REDUCE[kind], C.byref(desc(x)), out.data_ptr(), ws.data_ptr(), wp,
                          stream_ptr(x.device)), f"reduce {kind}")
    return out if kind == "min_max" else out.reshape(())


def tensor3(kind: str, h: torch.Tensor, out: torch.Tensor, flag: torch.Tensor = None) -> torch.Tensor:
    require_gpu(h, out)
    check(lib().sp_tensor3(TENSOR3[kind], C.byref(desc(h)), C.byref(desc(out)),
                           flag.data_ptr() if flag is not None else None,
                           stream_ptr(h.device)), kind)
    return out


def gen_fourier(out: torch.Tensor, shape, z0: int, ks: np.ndarray, amps: np.ndarray,
                phis: np.ndarray, seed: int) -> torch.Tensor:
    require_gpu(out)
    n = (C.c_int64 * 3)(*shape)
    k = np.ascontiguousarray(ks, dtype=np.int32).reshape(-1)
    a = np.ascontiguousarray(amps, dtype=np.float64)
    p = np.ascontiguousarray(phis, dtype=np.float64)
    check(lib().sp_gen_fourier(C.byref(desc(out)), n, int(z0),
                               k.ctypes.data_as(C.POINTER(C.c_int32)),
                               a.ctypes.data_as(C.POINTER(C.c_double)),
                               p.ctypes.data_as(C.POINTER(C.c_double)),
                               C.c_uint64(seed & ((1 << 64) - 1)), stream_ptr(out.device)),
          "gen_fourier")
    return out


def device_info():
    name = C.create_string_buffer(64)
    cu, xcd = C.c_int(0), C.c_int(0)
    check(lib().sp_device_info(name, 64, C.byref(cu), C.byref(xcd)), "device_info")
    return name.value.decode(), cu.value, xcd.value


# ---- hipRTC --------------------------------------------------------------------------------
def rtc_compile(source: str, name: str, options: Sequence[str]) -> bytes:
    code, size = C.c_void_p(), C.c_size_t()
    log = C.create_string_buffer(1 << 16)
    rc = lib().sp_rtc_compile(source.encode(), name.encode(), "\n".join(options).encode(),
                              C.byref(code), C.byref(size), log, len(log))
    if rc != 0:
        raise RuntimeError(f"hipRTC failed for {name}: {lib().sp_last_error().decode()}")
    try:
        return C.string_at(code, size.value)
    finally:
        lib().sp_free_code(code)


class Module:
    """A loaded code object holding one kernel."""

    def __init__(self, code: bytes, kernel: str):
        self._code = code
        h = C.c_void_p()
        check(lib().sp_module_load(code, len(code), kernel.encode(), C.byref(h)), "module_load")
        self.handle = h

    def launch(self, grid, block, args: bytes, shared: int = 0, stream: int = None) -> None:
        g = (C.c_uint * 3)(*grid)
        b = (C.c_uint * 3)(*block)
        s = stream if stream is not None else stream_ptr()
        check(lib().sp_module_launch(self.handle, g, b, shared, s, args, len(args)), "launch")

    def occupancy(self, block_threads: int, dyn_shared: int = 0) -> int:
        """Blocks per CU that fit at once (needs a device)."""
        n = C.c_int()
        check(lib().sp_module_occupancy(self.handle, int(block_threads), int(dyn_shared), C.byref(n)),
              "occupancy")
        return n.value

    def attributes(self):
        v, l, s, m = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().sp_module_attributes(self.handle, C.byref(v), C.byref(l), C.byref(s),
                                         C.byref(m)), "attributes")
        return {"vgprs": v.value, "lds": l.value, "spill": s.value, "max_threads": m.value}

    def __del__(self):
        try:
            if self.handle:
                lib().sp_module_unload(self.handle)
        except Exception:
            pass


def host_gen_fourier(out_ptr: int, shape, z0: int, nz: int, ks: np.ndarray, amps: np.ndarray,
                     phis: np.ndarray, seed: int, threads: int = 0) -> None:
    """Fill host memory at ``out_ptr`` with planes [z0, z0+nz) of the seeded field."""
    n = (C.c_int64 * 3)(*shape)
    k = np.ascontiguousarray(ks, dtype=np.int32).reshape(-1)
    a = np.ascontiguousarray(amps, dtype=np.float64)
    p = np.ascontiguousarray(phis, dtype=np.float64)
    rc = lib().sp_host_gen_fourier(out_ptr, n, int(z0), int(nz), k.ctypes.data_as(C.POINTER(C.c_int32)),
                                   a.ctypes.data_as(C.POINTER(C.c_double)),
                                   p.ctypes.data_as(C.POINTER(C.c_double)),
                                   C.c_uint64(seed & ((1 << 64) - 1)), int(threads))
    if rc != 0:
        raise ValueError("host_gen_fourier: bad arguments")


PROBE = {"copy": 0, "4r1w": 1, "read": 2, "write": 3, "4r1w_aos": 4}


def probe_stream(mix: str, persistent: bool, a: torch.Tensor, u: Optional[torch.Tensor] = None,
                 out: Optional[torch.Tensor] = None, partials: Optional[torch.Tensor] = None,
                 n_cu: int = 256) -> None:
    """One launch of an HBM streaming probe (``sp_probe_stream``) on torch's current stream."""
    require_gpu(a)
    check(lib().sp_probe_stream(PROBE[mix], int(persistent), a.data_ptr(),
                                u.data_ptr() if u is not None else None,
                                out.data_ptr() if out is not None else None,
                                partials.data_ptr() if partials is not None else None,
                                a.numel(), n_cu, stream_ptr(a.device)), "probe " + mix)


def probe_blocks(persistent: bool, n: int, n_cu: int = 256) -> int:
    return int(lib().sp_probe_blocks(int(persistent), n, n_cu))


def reduce_noalloc(kind: str, x: torch.Tensor, out: torch.Tensor,
                   wparams: Sequence[float] = None) -> torch.Tensor:
    """``sp_reduce_noalloc``: the reduction of ``reduce`` into the caller's FP64 ``out`` (1
    element, 2 for min_max), with the block records in the library's static device array; the
    call allocates nothing."""
    require_gpu(x, out)
    if not x.is_contiguous():
        raise ValueError("reduce_noalloc: the input must be contiguous (no copy is made)")
    if out.dtype != torch.float64 or out.numel() < (2 if kind == "min_max" else 1):
        raise ValueError("reduce_noalloc: out must hold FP64 results")
    wp = (C.c_double * 4)(*(wparams or (1.0, 1.0, 1.0, 0.0)))
    check(lib().sp_reduce_noalloc(REDUCE[kind], C.byref(desc(x)), out.data_ptr(), wp,
                                  stream_ptr(x.device)), f"reduce_noalloc {kind}")
    return out

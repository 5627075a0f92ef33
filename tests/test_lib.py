"""The C-ABI library builds, loads and exports every entry point include/*.h declares (CPU)."""
import ctypes
import re
from pathlib import Path

import pytest

from stenpy_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(sp_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "sp_abi_version" in names and "sp_reduce" in names and len(names) >= 25


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"


def test_abi_version_and_error_channel():
    L = _lib.lib()
    assert L.sp_abi_version() == _lib.ABI_VERSION
    # a bad argument fails with a status and a message, without touching the GPU
    rc = L.sp_op_unary(99, None, None, 0.0, 0.0, None)
    assert rc != 0
    assert L.sp_last_error()


def test_python_binding_covers_every_export():
    src = (ROOT / "stenpy_amd" / "_lib.py").read_text()
    missing = [n for n in declared_functions() if f'"{n}"' not in src]
    assert not missing, f"not bound in _lib.py: {missing}"


def test_gpu_only_contract():
    import torch
    from stenpy_amd import ops
    with pytest.raises(RuntimeError, match="GPU"):
        ops.laplacian(torch.zeros(4, 4, 4))

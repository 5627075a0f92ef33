"""Build libstenpy_hip.so in-tree with hipcc for gfx950.

Usage: ``python -m stenpy_amd.build`` (also called by ``__graft_entry__.build()``).
Every .hip file under ``stenpy_amd/csrc`` is compiled in parallel to an object file under
``stenpy_amd/_build`` and linked with hipRTC into ``stenpy_amd/libstenpy_hip.so``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libstenpy_hip.so"
ARCH = os.environ.get("STENPY_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are disabled for every device compile (hipcc emits them from
# ordinary float code otherwise); the host pass warns that it does not know the feature.
COMMON_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # FMA contraction except where a function's pragma forbids it (compensated sums)
    "-ffp-contract=fast-honor-pragmas",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found; the HIP backend cannot be built")
    return exe


def _compile(src: Path) -> Path:
    obj = OBJ / (src.stem + ".o")
    cmd = [hipcc(), *COMMON_FLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stderr[-6000:]}")
    return obj


def _compile_host(src: Path) -> Path:
    """Host-only C++ (the slab generator): g++ with OpenMP and per-ISA clones."""
    obj = OBJ / (src.stem + ".o")
    cmd = ["g++", "-O3", "-fPIC", "-std=c++17", "-fopenmp", f"-I{ROOT / 'include'}", "-c",
           str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for {src.name}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    host_sources = sorted(CSRC.glob("*.cpp"))
    jobs = min(len(sources) + len(host_sources), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)), 16)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        futs = [ex.submit(_compile, s) for s in sources] + [ex.submit(_compile_host, s) for s in host_sources]
        objs = [f.result() for f in futs]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-fopenmp", "-o", str(tmp),
           *map(str, objs), "-L/opt/rocm/lib", "-lhiprtc", "-lrocfft", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(sources)} sources")
    return LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as exc:
        print(exc, file=sys.stderr)
        sys.exit(1)

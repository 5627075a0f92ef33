"""The per-voxel 3x3 routines of stenpy_amd/csrc/sp_sym3.h built as plain C++ for the host and
checked on the CPU: ``div3f`` (x / 3 by multiplication and one FMA correction) against the FP64
quotient rounded to FP32, over whole binades; ``atan2_upper`` against numpy's FP64 arctan2;
``sym3_eigvalsh`` against numpy's FP64 eigvalsh on random, nearly degenerate and widely scaled
symmetric matrices."""
import ctypes
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
HDR = ROOT / "stenpy_amd" / "csrc" / "sp_sym3.h"

_SHIM = r"""
#include <cmath>
#define __device__
#define __forceinline__ inline
#include "sp_sym3.h"
extern "C" void div3(const float* x, float* y, long n) { for (long i = 0; i < n; ++i) y[i] = div3f(x[i]); }
extern "C" void eig(const float* m, float* e, long n) { for (long i = 0; i < n; ++i) sym3_eigvalsh(m + 9 * i, e + 3 * i); }
extern "C" void dev(const float* m, float* d, long n) { for (long i = 0; i < n; ++i) sym3_deviatoric(m + 9 * i, d + 9 * i); }
extern "C" void atan2u(const float* y, const float* x, float* r, long n) { for (long i = 0; i < n; ++i) r[i] = atan2_upper(y[i], x[i]); }
extern "C" void rnorm(const float* s, float eps, float* r, long n) { for (long i = 0; i < n; ++i) r[i] = unit_rnorm(s[i], eps); }
"""


@pytest.fixture(scope="module")
def sym3(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    d = tmp_path_factory.mktemp("sym3")
    (d / "shim.cpp").write_text(_SHIM)
    so = d / "libsym3.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", f"-I{HDR.parent}",
                    "-o", str(so), str(d / "shim.cpp")], check=True)
    lib = ctypes.CDLL(str(so))
    for fn in (lib.div3, lib.eig, lib.dev):
        fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
    lib.atan2u.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
    lib.rnorm.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p, ctypes.c_long]
    return lib


def _div3(lib, x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    lib.div3(x.ctypes.data, y.ctypes.data, x.size)
    return y


def test_div3f_is_correctly_rounded(sym3):
    """Normal inputs with normal quotients: the scheme commutes with scaling by powers of two,
    so the binade [1, 2) covers them; the binades where x / 3 is subnormal or x is the largest
    are checked whole, with both signs.  The reference rounding is FP64 x / 3 rounded to FP32,
    which is the correct rounding (x / 3 is never within 2^-51 of an FP32 midpoint)."""
    def binade(e):
        bits = (np.arange(1 << 23, dtype=np.uint32) | np.uint32((e + 127) << 23)) if e > -127 else \
            np.arange(1 << 23, dtype=np.uint32)
        return bits.view(np.float32)
    for e in (0, -126, -127, -125, 127, 126):
        x = binade(e)
        x = np.concatenate([x, -x])
        got = _div3(sym3, x)
        want = (x.astype(np.float64) / 3.0).astype(np.float32)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), e
    sp = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 3.0, -3.0, 1.5, 3.4028235e38], dtype=np.float32)
    got = _div3(sym3, sp)
    want = (sp.astype(np.float64) / 3.0).astype(np.float32)
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    assert np.array_equal(got[ok].view(np.uint32), want[ok].view(np.uint32))


def test_eigvalsh_against_numpy_fp64(sym3):
    """Eigenvalues within a few FP32 ulp of the spectral radius (the bound of the closed form's
    angle), on well separated, nearly double and triple, and widely scaled spectra."""
    rng = np.random.default_rng(3)
    mats = []
    for scale in (1e-30, 1e-3, 1.0, 1e3, 1e25):
        a = rng.standard_normal((2000, 3, 3))
        mats.append(scale * (a + a.transpose(0, 2, 1)))
        q, _ = np.linalg.qr(rng.standard_normal((2000, 3, 3)))
        lam = np.stack([np.ones(2000), 1 + 1e-6 * rng.standard_normal(2000), 1 + 1e-3 * rng.standard_normal(2000)], -1)
        mats.append(scale * (q * lam[:, None, :]) @ q.transpose(0, 2, 1))
    m = np.concatenate(mats).astype(np.float32)
    m = 0.5 * (m + m.transpose(0, 2, 1))
    e = np.empty((m.shape[0], 3), dtype=np.float32)
    sym3.eig(np.ascontiguousarray(m).ctypes.data, e.ctypes.data, m.shape[0])
    ref = np.linalg.eigvalsh(m.astype(np.float64))
    rad = np.abs(ref).max(axis=1, keepdims=True)
    err = np.abs(e.astype(np.float64) - ref) / (rad * 2.0 ** -24)
    print("eig max err (ulp of radius)", err.max())
    assert err.max() <= 16.0, err.max()


def test_atan2_upper_against_numpy(sym3):
    """The closed form's angle: atan2(y, x) for y >= 0 within 2 ulp of pi over a grid of both
    signs of x, magnitudes 2^-60 .. 2^60, and the axes and zeros."""
    rng = np.random.default_rng(5)
    y = np.abs(rng.standard_normal(200000) * 2.0 ** rng.integers(-60, 60, 200000)).astype(np.float32)
    x = (rng.standard_normal(200000) * 2.0 ** rng.integers(-60, 60, 200000)).astype(np.float32)
    special_y = np.array([0, 0, 0, 1, 1, 1, 1e-30, 3e38, 1, 0], dtype=np.float32)
    special_x = np.array([0, -0.0, 1, 0, -0.0, -1, -1e30, 1, 1, -5], dtype=np.float32)
    y, x = np.concatenate([y, special_y, y[:1000]]), np.concatenate([x, special_x, y[:1000]])
    got = np.empty_like(y)
    sym3.atan2u(y.ctypes.data, x.ctypes.data, got.ctypes.data, y.size)
    ref = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    err = np.abs(got - ref) / (np.pi * 2.0 ** -24)
    assert err.max() <= 2.0, (err.max(), y[err.argmax()], x[err.argmax()])
    nan = np.array([np.nan, 1.0], dtype=np.float32), np.array([1.0, np.nan], dtype=np.float32)
    out = np.empty(2, dtype=np.float32)
    sym3.atan2u(nan[0].ctypes.data, nan[1].ctypes.data, out.ctypes.data, 2)
    assert np.isnan(out).all()


def test_unit_rnorm_values(sym3):
    """1 / max(sqrt(s), eps) after one FP32 Newton step: within 1 ulp of the FP64 value over 40
    binades of s (from a correctly rounded start here; the device starts from the 1-ulp hardware
    reciprocal square root, which the step absorbs), and the clamp / NaN / inf cases."""
    rng = np.random.default_rng(5)
    s = np.exp(rng.uniform(np.log(1e-20), np.log(1e20), 200000)).astype(np.float32)
    eps = np.float32(1e-15)
    r = np.empty_like(s)
    sym3.rnorm(s.ctypes.data, float(eps), r.ctypes.data, s.size)
    exact = 1.0 / np.sqrt(s.astype(np.float64))
    ulp = np.spacing(exact.astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(r - exact) / ulp) <= 1.0
    special = np.array([0.0, 1e-31, np.inf, np.nan, 1e-30], dtype=np.float32)
    out = np.empty_like(special)
    sym3.rnorm(special.ctypes.data, float(eps), out.ctypes.data, special.size)
    # below eps^2: 1 / eps to about an ulp; an infinite s: 0; NaN: NaN
    for v in (out[0], out[1], out[4]):
        assert abs(v - 1e15) <= 1e15 * 2.5e-7
    assert out[2] == 0.0 and np.isnan(out[3])

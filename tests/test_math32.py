"""The FP32 exp / tanh / log / sin / cos of the generated kernels (stenpy_amd/csrc/sp_math32.h), built as
plain C++ for the host and checked on the CPU against FP64 numpy over a strided sweep of every
FP32 bit pattern plus the special and threshold values.  The bound is the one the header
states: 0.5 ulp of the FP32 result plus 2^-16 ulp.  The reduction constants are checked against
60-digit decimal values."""
import ctypes
import os
import re
import shutil
import subprocess
from decimal import Decimal, getcontext
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
HDR = ROOT / "stenpy_amd" / "csrc" / "sp_math32.h"

_SHIM = r"""
#include "sp_math32.h"
extern "C" void run(int fn, const float* x, float* y, long n) {
    for (long i = 0; i < n; ++i)
        y[i] = fn == 0 ? spm_expf(x[i]) : fn == 1 ? spm_tanhf(x[i]) : fn == 2 ? spm_logf(x[i])
             : fn == 3 ? spm_sinf(x[i]) : spm_cosf(x[i]);
}
"""


@pytest.fixture(scope="module")
def math32(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    d = tmp_path_factory.mktemp("math32")
    (d / "shim.cpp").write_text(_SHIM)
    so = d / "libmath32.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", f"-I{HDR.parent}",
                    "-o", str(so), str(d / "shim.cpp")], check=True)
    lib = ctypes.CDLL(str(so))
    lib.run.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]

    def call(fn, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty_like(x)
        lib.run(fn, x.ctypes.data, y.ctypes.data, x.size)
        return y
    return call


def _inputs(lo_bits=0, stride=1021):
    bits = np.arange(lo_bits, 1 << 32, stride, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -1.0, 1e-45, -1e-45, 1.1754942e-38,
                        3.4028235e38, -3.4028235e38, 88.72283, 88.72284, -87.33654, -103.97208, -103.97209,
                        9.01, 9.02, -9.02, 1e-8, -1e-8, 0.5, 0.75, 1.5, 0.99999994, 1.0000001],
                       dtype=np.float32)
    return np.concatenate([x, special])


def _ulp_errors(got, ref64):
    """|got - ref| in units of the last place of the FP32 rounding of ref (finite refs only)."""
    with np.errstate(over="ignore", invalid="ignore"):
        r32 = ref64.astype(np.float32)
        fin = np.isfinite(r32) & np.isfinite(ref64)
        _, e = np.frexp(ref64[fin])
        ulp = np.ldexp(1.0, np.maximum(e - 24, -149))
        err = np.abs(got[fin].astype(np.float64) - ref64[fin]) / ulp
    return fin, err


@pytest.mark.parametrize("fn,name", [(0, "exp"), (1, "tanh"), (2, "log"), (3, "sin"), (4, "cos")])
def test_accuracy_over_fp32_inputs(math32, fn, name):
    x = _inputs(lo_bits=fn * 7)
    got = math32(fn, x)
    xd = x.astype(np.float64)
    with np.errstate(all="ignore"):
        ref = {0: np.exp, 1: np.tanh, 2: np.log, 3: np.sin, 4: np.cos}[fn](xd)
    fin, err = _ulp_errors(got, ref)
    assert err.max() <= 0.5 + 2.0 ** -16, f"{name}: max error {err.max():.9f} ulp"
    # non-finite FP32 results: the same inf / NaN as the FP32 rounding of the FP64 value
    with np.errstate(over="ignore", invalid="ignore"):
        r32 = ref.astype(np.float32)
    nf = ~fin
    assert np.array_equal(got[nf], r32[nf], equal_nan=True), name
    # mostly correctly rounded
    assert np.mean(got[fin] != r32[fin]) < 1e-5, name


def test_edge_values(math32):
    assert math32(0, [100.0, -200.0, np.inf, -np.inf]).tolist() == [np.inf, 0.0, np.inf, 0.0]
    assert np.isnan(math32(0, [np.nan])[0]) and np.isnan(math32(1, [np.nan])[0])
    assert math32(1, [20.0, -20.0, np.inf, -np.inf, 1e-30]).tolist() == [1.0, -1.0, 1.0, -1.0, np.float32(1e-30)]
    assert np.signbit(math32(1, [-0.0])[0])
    lg = math32(2, [0.0, -0.0, -1.0, np.inf, 1.0, np.nan])
    assert lg[0] == -np.inf and lg[1] == -np.inf and np.isnan(lg[2]) and lg[3] == np.inf and lg[4] == 0.0
    assert np.isnan(lg[5])
    tr = math32(3, [0.0, -0.0, np.inf, np.nan, 1e-30, 3e38, 524287.97, 524288.0])
    assert tr[0] == 0.0 and np.signbit(tr[1]) and np.isnan(tr[2]) and np.isnan(tr[3]) and tr[4] == np.float32(1e-30)
    assert tr[5] == np.float32(np.sin(np.float64(np.float32(3e38))))
    assert math32(4, [0.0, np.inf, 1e-30]).tolist()[0::2] == [1.0, 1.0] and np.isnan(math32(4, [np.inf])[0])


def _nearest_double(d):
    f = float(d)
    cands = [np.nextafter(f, -np.inf), f, np.nextafter(f, np.inf)]
    return float(min(cands, key=lambda v: abs(Decimal(float(v)) - d)))


def test_reduction_constants_are_correctly_rounded():
    getcontext().prec = 60
    ln2 = Decimal(2).ln()
    text = HDR.read_text()
    pi = Decimal("3.14159265358979323846264338327950288419716939937510582097494459")
    p1 = _nearest_double(pi / 2)
    for value, exact in (("0x1.62e42fefa39efp-1", ln2), ("0x1.71547652b82fep+0", 1 / ln2),
                         ("0x1.921fb54442d18p+0", pi / 2), ("0x1.1a62633145c07p-54", pi / 2 - Decimal(p1)),
                         ("0x1.45f306dc9c883p-1", 2 / pi)):
        assert value in text and float.fromhex(value) == _nearest_double(exact)
    # the integer-rounding shifter and the subnormal scale are exact powers (1.5 2^52, 2^23)
    assert "0x1.8p52" in text and "0x1p23f" in text


_SWEEP = r"""
#include <math.h>
#include <string.h>
#include "sp_math32.h"
// every FP32 bit pattern: max error in ulps of the FP32 rounding of the x87 long double value,
// and the count of results that differ from that rounding
extern "C" void sweep(int fn, double* max_ulp, long long* not_rounded) {
    double mx = 0.0;
    long long miss = 0;
    #pragma omp parallel for reduction(max:mx) reduction(+:miss) schedule(static, 65536)
    for (long long i = 0; i < (1LL << 32); ++i) {
        const unsigned u = (unsigned)i;
        float x;
        memcpy(&x, &u, 4);
        float got;
        long double ref;
        switch (fn) {
            case 0: got = spm_expf(x); ref = expl(x); break;
            case 1: got = spm_tanhf(x); ref = tanhl(x); break;
            case 2: got = spm_logf(x); ref = logl(x); break;
            case 3: got = spm_sinf(x); ref = sinl(x); break;
            default: got = spm_cosf(x); ref = cosl(x); break;
        }
        const float r32 = (float)ref;
        if (isnan(r32) || isinf(r32)) {
            if (!(got == r32 || (isnan(got) && isnan(r32)))) { ++miss; mx = 1e9; }
            continue;
        }
        int e;
        frexpl(ref, &e);
        const long double ulp = ldexpl(1.0L, e - 24 > -149 ? e - 24 : -149);
        const double err = (double)(fabsl((long double)got - ref) / ulp);
        if (err > mx) mx = err;
        if (got != r32) ++miss;
    }
    *max_ulp = mx;
    *not_rounded = miss;
}
"""


@pytest.mark.skipif(os.environ.get("STENPY_MATH32_EXHAUSTIVE") != "1",
                    reason="exhaustive sweep (minutes); set STENPY_MATH32_EXHAUSTIVE=1")
@pytest.mark.parametrize("fn,name", [(0, "exp"), (1, "tanh"), (2, "log"), (3, "sin"), (4, "cos")])
def test_exhaustive_over_all_fp32_inputs(tmp_path, fn, name):
    """All 2^32 inputs against x86 80-bit long double (64-bit significand): within 0.5 + 2^-16 ulp,
    and correctly rounded for all but a handful."""
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    (tmp_path / "sweep.cpp").write_text(_SWEEP)
    so = tmp_path / "libsweep.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared",
                    f"-I{HDR.parent}", "-o", str(so), str(tmp_path / "sweep.cpp")], check=True)
    lib = ctypes.CDLL(str(so))
    mx, miss = ctypes.c_double(), ctypes.c_longlong()
    lib.sweep(fn, ctypes.byref(mx), ctypes.byref(miss))
    assert mx.value <= 0.5 + 2.0 ** -16, f"{name}: max error {mx.value:.9f} ulp"
    assert miss.value <= 64, f"{name}: {miss.value} results not correctly rounded"

"""Golden fixtures (tests/golden/fixtures.npz, written by tests/golden/make_golden.py).

CPU: the generator reproduces its committed output bit for bit and matches a pure-Python
evaluation of the formula; the oracle reproduces the committed outputs; committed outputs
satisfy closed-form invariants.  GPU: the HIP paths reproduce the committed oracle outputs
within the 2x error rule.
"""
import math
from pathlib import Path

import numpy as np
import pytest
import torch

import stenpy_amd as sp
from stenpy_amd import gen
from oracle import graph_eval

FIX = np.load(Path(__file__).resolve().parent / "golden" / "fixtures.npz")
STENCIL_GRAPH = "divergence(gradient(f)) + dot(u, gradient(f)) + sqnorm(curl(u))"


def test_generator_fixture_bit_exact():
    assert np.array_equal(gen.generate_host((16, 16, 16), 2), FIX["gen16_seed2"])
    assert np.array_equal(gen.generate_host((12, 14, 16), 2), FIX["f"])


def test_generator_formula_pure_python():
    shape, seed = (3, 4, 5), 7
    ks, amps, phis = gen.fourier_params(seed)
    assert ks.shape == (8, 3) and ks.min() >= 1 and ks.max() <= 8
    host = gen.generate_host(shape, seed)
    M = (1 << 64) - 1
    for i in range(shape[0]):
        for j in range(shape[1]):
            for k in range(shape[2]):
                v = sum(float(amps[m]) * math.cos(2 * math.pi * (int(ks[m, 0]) * i / shape[0] + int(ks[m, 1]) * j / shape[1]
                                                                   + int(ks[m, 2]) * k / shape[2]) + float(phis[m]))
                        for m in range(8))
                lin = (i * shape[1] + j) * shape[2] + k
                z = (seed + (lin + 1) * 0x9E3779B97F4A7C15) & M
                z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
                z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
                z ^= z >> 31
                v += (2.0 * ((z >> 11) / 2.0 ** 53) - 1.0) * 1e-3
                assert host[i, j, k] == np.float32(v)


def test_generator_slabs_compose():
    full = gen.generate_host((10, 6, 7), 3)
    parts = np.concatenate([gen.generate_host((10, 6, 7), 3, z0=a, nz=b - a) for a, b in ((0, 3), (3, 8), (8, 10))])
    assert np.array_equal(full, parts)


def _graph(expr, f, u):
    g, sink, *_ = sp.compile_expression(expr, boundary="periodic", field_map={"f": f, "u": u})
    return g, sink


@pytest.mark.parametrize("name,expr", [("pointwise", "f**2 + tanh(f)*exp(f)"), ("laplacian", "laplacian(f)"),
                                       ("stencil_graph", STENCIL_GRAPH), ("hessian_eig", "eigenvalues(hessian(f))")])
def test_oracle_reproduces_fixtures(name, expr):
    f, u = torch.from_numpy(FIX["f"]), torch.from_numpy(FIX["u"])
    g, sink = _graph(expr, f, u)
    o64 = graph_eval.evaluate(g, torch.float64)[sink]
    np.testing.assert_array_equal(o64.numpy(), FIX[f"{name}__out64"])


def test_fixture_invariants():
    # periodic Laplacian sums to zero over the grid; gradient of a periodic field too
    assert abs(FIX["laplacian__out64"].sum()) < 1e-11
    assert np.abs(FIX["gradient__out64"].sum(axis=(0, 1, 2))).max() < 1e-11
    # eigenvalues ascending; sum of eigenvalues = trace of the hessian
    e = FIX["hessian_eig__out64"]
    assert (np.diff(e, axis=-1) >= -1e-12).all()


@pytest.mark.gpu
@pytest.mark.parametrize("name,expr", [("pointwise", "f**2 + tanh(f)*exp(f)"), ("laplacian", "laplacian(f)"),
                                       ("gradient", "gradient(f)"), ("stencil_graph", STENCIL_GRAPH),
                                       ("hessian_eig", "eigenvalues(hessian(f))"), ("hessian_det", "determinant(hessian(f))"),
                                       ("mean_curvature", "mean_curvature(f)")])
@pytest.mark.parametrize("path", ["fused", "eager"])
def test_hip_paths_match_fixtures(name, expr, path):
    from tests._util import assert_2x
    f = torch.from_numpy(FIX["f"]).cuda()
    u = torch.from_numpy(FIX["u"]).cuda()
    g, sink = _graph(expr, f, u)
    rt = sp.Runtime(device="cuda")
    res = rt.fused.run(g, rt, wanted=[sink]) if path == "fused" else rt.run_eager(g)
    assert_2x(res[sink], torch.from_numpy(FIX[f"{name}__out64"]), torch.from_numpy(FIX[f"{name}__out32"]), name)


@pytest.mark.gpu
def test_device_generator_matches_host():
    d = gen.generate((16, 16, 16), 2, "cuda").cpu().numpy()
    h = FIX["gen16_seed2"]
    ulp = np.spacing(np.abs(h).astype(np.float32))
    assert (np.abs(d.astype(np.float64) - h) <= ulp).all()

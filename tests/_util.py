"""Shared helpers for the parity tests: seeded inputs and the 2x error rule."""
from __future__ import annotations

import torch

from stenpy_amd import gen


def field(shape, seed, dev="cuda"):
    return gen.generate(shape, seed, dev)


def vfield(shape, seeds=(3, 4, 5), dev="cuda"):
    return torch.stack([gen.generate(shape, s, dev) for s in seeds], dim=-1).contiguous()


def errs(build, o64, o32):
    """(e_build, e_ref32): max |x - oracle64| of the build and of the FP32 twin."""
    b = build.detach().double().cpu()
    r = o64.double()
    assert b.shape == r.shape, f"shape {tuple(b.shape)} vs oracle {tuple(r.shape)}"
    nan_b, nan_r = torch.isnan(b), torch.isnan(r)
    assert torch.equal(nan_b, nan_r), "NaN pattern differs from the oracle"
    b, r, t = b[~nan_b], r[~nan_r], o32.double()[~nan_r]
    if b.numel() == 0:
        return 0.0, 0.0
    return (b - r).abs().max().item(), (t - r).abs().max().item()


def assert_2x(build, o64, o32, what="", slack=0.0):
    """Error rule: max|build - o64| <= 2 max|ref32 - o64| (+ ``slack``, stated by the caller)."""
    e_b, e_32 = errs(build, o64, o32)
    assert e_b <= 2.0 * e_32 + slack, f"{what}: build error {e_b:.3e} > 2 x ref32 error {e_32:.3e} + {slack:.1e}"
    return e_b, e_32

"""Seeded synthetic fields: a sum of 8 Fourier modes plus counter-based noise.

Specification (fixed; every benchmark and test input comes from it):

    rng = numpy.random.Generator(PCG64(seed))
    for m in 0..7:  k_m = rng.integers(1, 9, 3); a_m = rng.standard_normal() / |k_m|;
                    phi_m = rng.uniform(0, 2 pi)
    f[i,j,k] = sum_m a_m cos(2 pi (k_m0 i/n0 + k_m1 j/n1 + k_m2 k/n2) + phi_m)     (FP64)
             + (2 u - 1) * 1e-3,   u = top 53 bits of the (lin+1)-th output of a SplitMix64
               stream started at ``seed``, lin = C-order linear index in the full grid
    stored as FP32 (rounded once).

The noise is counter-based, so any z-slab of any field can be produced independently (the
out-of-core streamer relies on that).  ``generate`` runs the FP64 formula on the GPU through
the C-ABI; ``generate_host`` is the NumPy twin used for host-side slabs.  They agree to
within one FP32 ulp (FP64 ``cos`` of device and host libraries may differ in the last bit).
"""
from __future__ import annotations

import math
from typing import Sequence, Tuple

import numpy as np

N_MODES = 8
NOISE_AMPLITUDE = 1e-3
_GAMMA = 0x9E3779B97F4A7C15
_M1 = 0xBF58476D1CE4E5B9
_M2 = 0x94D049BB133111EB
_MASK = (1 << 64) - 1


def fourier_params(seed: int) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Mode wavenumbers (8x3 int), amplitudes (8) and phases (8) for ``seed``."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ks, amps, phis = [], [], []
    for _ in range(N_MODES):
        k = rng.integers(1, 9, 3)
        a = rng.standard_normal() / float(np.linalg.norm(k))
        phi = rng.uniform(0.0, 2.0 * math.pi)
        ks.append(k)
        amps.append(a)
        phis.append(phi)
    return (np.asarray(ks, dtype=np.int32), np.asarray(amps, dtype=np.float64),
            np.asarray(phis, dtype=np.float64))


def splitmix64_u53(seed: int, lin: np.ndarray) -> np.ndarray:
    """Uniform [0,1) doubles from the (lin+1)-th SplitMix64 output of a stream at ``seed``."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed & _MASK) + (lin.astype(np.uint64) + np.uint64(1)) * np.uint64(_GAMMA))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(_M1)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(_M2)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def generate_host(shape: Sequence[int], seed: int, z0: int = 0, nz: int = None,
                  modes=None) -> np.ndarray:
    """FP32 planes ``[z0, z0+nz)`` of the field of full ``shape`` (n0, n1, n2), on the host,
    by the direct formula (one cosine per voxel and mode).  ``modes`` = (ks, amps, phis)
    replaces the seed's modes (the noise still follows ``seed``)."""
    n0, n1, n2 = (int(s) for s in shape)
    nz = n0 - z0 if nz is None else int(nz)
    ks, amps, phis = fourier_params(seed) if modes is None else modes
    i = np.arange(z0, z0 + nz, dtype=np.float64)[:, None, None]
    j = np.arange(n1, dtype=np.float64)[None, :, None]
    k = np.arange(n2, dtype=np.float64)[None, None, :]
    acc = np.zeros((nz, n1, n2), dtype=np.float64)
    two_pi = 2.0 * math.pi
    for m in range(N_MODES):
        ph = two_pi * (ks[m, 0] * i / n0 + ks[m, 1] * j / n1 + ks[m, 2] * k / n2) + phis[m]
        acc += amps[m] * np.cos(ph)
    lin = (np.arange(z0, z0 + nz, dtype=np.int64)[:, None, None] * n1
           + np.arange(n1, dtype=np.int64)[None, :, None]) * n2 \
        + np.arange(n2, dtype=np.int64)[None, None, :]
    acc += (2.0 * splitmix64_u53(seed, lin) - 1.0) * NOISE_AMPLITUDE
    return acc.astype(np.float32)


def generate(shape: Sequence[int], seed: int, device="cuda", out=None, z0: int = 0,
             nz: int = None):
    """FP32 field (or planes ``[z0, z0+nz)``) generated on the GPU by the C-ABI kernel."""
    import torch
    from . import _lib
    n0, n1, n2 = (int(s) for s in shape)
    nz = n0 - z0 if nz is None else int(nz)
    if out is None:
        out = torch.empty((nz, n1, n2), dtype=torch.float32, device=device)
    ks, amps, phis = fourier_params(seed)
    _lib.gen_fourier(out, (n0, n1, n2), z0, ks, amps, phis, seed)
    return out


def generate_vector(shape: Sequence[int], seeds: Sequence[int], device="cuda"):
    """Channel-last (n0, n1, n2, C) field whose component c uses ``seeds[c]``."""
    import torch
    comps = [generate(shape, s, device) for s in seeds]
    return torch.stack(comps, dim=-1).contiguous()


def quadratic_host(shape: Sequence[int], z0: int = 0, nz: int = None) -> np.ndarray:
    """i^2 + j^2 + k^2 in FP32 (exact for coordinates below 2^12)."""
    n0, n1, n2 = (int(s) for s in shape)
    nz = n0 - z0 if nz is None else int(nz)
    i = np.arange(z0, z0 + nz, dtype=np.float64)[:, None, None]
    j = np.arange(n1, dtype=np.float64)[None, :, None]
    k = np.arange(n2, dtype=np.float64)[None, None, :]
    return (i * i + j * j + k * k).astype(np.float32)

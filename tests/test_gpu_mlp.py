"""Schedule-MLP kernels: MFMA GEMM, forward/backward gradients, Adam, training (GPU)."""
import math

import numpy as np
import pytest
import torch

from stenpy_amd import _lib as L
from stenpy_amd.mlp import ScheduleMLP

pytestmark = pytest.mark.gpu


def bf16_round(x: torch.Tensor) -> torch.Tensor:
    return x.to(torch.bfloat16).to(torch.float64)


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(64, 64, 16), (100, 37, 70), (1, 256, 256), (256, 1, 513)])
def test_gemm_bf16_layouts(ta, tb, M, N, K):
    g = torch.Generator().manual_seed(M * 7 + N)
    A = torch.randn((K, M) if ta else (M, K), generator=g).to(torch.bfloat16)
    B = torch.randn((N, K) if tb else (K, N), generator=g).to(torch.bfloat16)
    Ad, Bd = A.cuda(), B.cuda()
    C = torch.full((M, N), 3.0, device="cuda")
    L.check(L.lib().sp_gemm_bf16(ta, tb, M, N, K, Ad.data_ptr(), A.shape[1], Bd.data_ptr(), B.shape[1],
                                 C.data_ptr(), N, 0.5, 1, L.stream_ptr()), "gemm")
    opA = A.double().t() if ta else A.double()
    opB = B.double().t() if tb else B.double()
    ref = 3.0 + 0.5 * (opA @ opB)
    err = (C.cpu().double() - ref).abs().max().item()
    assert err <= 1e-5 * math.sqrt(K) * max(1.0, ref.abs().max().item()), err


def _reference_step(model: ScheduleMLP, X: torch.Tensor, t: torch.Tensor):
    """FP64 restatement with the kernels' bf16 rounding points; returns (loss, gW, gb)."""
    W = [bf16_round(w.cpu().double()) for w in model.W]
    b = [v.cpu().double() for v in model.b]
    a = [bf16_round(X.double())]
    zs = []
    for l in range(len(W)):
        z = a[-1] @ W[l] + b[l]
        zs.append(z)
        if l < len(W) - 1:
            a.append(bf16_round(torch.relu(z)))
    y = zs[-1].reshape(-1)
    B = y.shape[0]
    loss = ((y - t.double()) ** 2).mean()
    dz = (2 * (y - t.double()) / B).reshape(B, 1)
    gW, gb = [None] * len(W), [None] * len(W)
    for l in range(len(W) - 1, -1, -1):
        dzb = bf16_round(dz)
        gW[l] = a[l].t() @ dzb
        gb[l] = dz.sum(0)
        if l:
            dz = (dzb @ W[l].t()) * (zs[l - 1] > 0)
    return loss, gW, gb


def test_gradients_match_cpu_oracle():
    torch.manual_seed(0)
    model = ScheduleMLP(n_in=28, width=256, depth=4, device="cuda", seed=1)
    X = torch.randn(96, 28)
    t = torch.randn(96)
    model.forward(model.to_bf16(X))
    loss = model.backward(t.cuda())
    rl, rgW, rgb = _reference_step(model, X, t)
    assert abs(loss.item() - rl.item()) <= 1e-5 * max(1.0, abs(rl.item()))
    for l in range(4):
        gw, rw = model.gW[l].cpu().double(), rgW[l]
        scale = max(rw.abs().max().item(), 1e-12)
        mismatch = ((gw - rw).abs() > 1e-3 * scale).double().mean().item()
        assert mismatch < 1e-3, f"layer {l}: {mismatch:.2e} of dW entries off"
        assert (model.gb[l].cpu().double() - rgb[l]).abs().max().item() <= 1e-3 * max(rgb[l].abs().max().item(), 1e-12)


def test_gradients_match_autograd_fp64():
    """Against plain FP64 autograd (no bf16 anywhere): agreement at bf16 resolution.

    Inputs, weights, activations and back-propagated errors are each rounded to bf16 (8-bit
    mantissa) by the kernels, so layer-0 gradients drift by a few percent after three layers;
    the exact-rounding check above pins the arithmetic, this one the overall direction."""
    model = ScheduleMLP(n_in=28, width=256, depth=4, device="cuda", seed=2)
    X = torch.randn(64, 28)
    t = torch.randn(64)
    model.forward(model.to_bf16(X))
    model.backward(t.cuda())
    W = [w.cpu().double().requires_grad_(True) for w in model.W]
    b = [v.cpu().double().requires_grad_(True) for v in model.b]
    a = X.double()
    for l in range(4):
        a = a @ W[l] + b[l]
        if l < 3:
            a = torch.relu(a)
    loss = ((a.reshape(-1) - t.double()) ** 2).mean()
    loss.backward()
    for l in range(4):
        ref = W[l].grad
        rel = (model.gW[l].cpu().double() - ref).norm() / max(ref.norm().item(), 1e-12)
        assert rel < 0.15, f"layer {l} relative gradient error {rel:.3e}"


def test_training_fits_a_smooth_function():
    rng = np.random.default_rng(0)
    X = rng.uniform(-1, 1, (2048, 6)).astype(np.float32)
    y = (np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2] + 0.5 * X[:, 3] ** 2).astype(np.float32)
    model = ScheduleMLP(n_in=6, width=256, depth=4, device="cuda", seed=3)
    losses = model.fit(X[:1536], y[:1536], epochs=60, batch=256, lr=2e-3)
    assert losses[-1] < 0.1 * losses[0]
    pred = model.predict(X[1536:])
    assert np.mean((pred - y[1536:]) ** 2) < 0.1 * np.var(y[1536:])

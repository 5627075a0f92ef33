"""Learned schedule model: a 4-layer width-256 MLP trained on autotuner timings.

The network maps (group features, schedule features) to the log of the kernel time per
voxel; the autotuner's exhaustive search stays the fallback and the measure of its regret.
Weights are FP32 masters with BF16 copies; every matmul is ``sp_gemm_bf16`` (MFMA, FP32
accumulate) and every other step (bias + ReLU, ReLU backward, bias gradients, MSE, Adam) is
a HIP kernel of libstenpy_hip -- torch only allocates the buffers.

Layer l: z_l = a_l W_l + b_l (fp32), a_{l+1} = bf16(relu(z_l)) (no ReLU after the last).
Backward with dz_L = 2 (y - t) / B:
    dW_l = a_l^T bf16(dz_l),  db_l = colsum(dz_l),  da_l = bf16(dz_l) W_l^T,
    dz_{l-1} = da_l * (z_{l-1} > 0).
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _lib as L


def _p(t: Optional[torch.Tensor]):
    return None if t is None else t.data_ptr()


class ScheduleMLP:
    def __init__(self, n_in: int, width: int = 256, depth: int = 4, device="cuda", seed: int = 0):
        self.device = torch.device(device)
        self.dims = [n_in] + [width] * (depth - 1) + [1]
        g = torch.Generator().manual_seed(seed)
        self.W, self.b, self.Wb = [], [], []
        for k, n in zip(self.dims[:-1], self.dims[1:]):
            w = torch.randn(k, n, generator=g) * math.sqrt(2.0 / k)
            self.W.append(w.to(self.device).contiguous())
            self.b.append(torch.zeros(n, device=self.device))
            self.Wb.append(torch.empty(k, n, dtype=torch.bfloat16, device=self.device))
        self.mW = [torch.zeros_like(w) for w in self.W]
        self.vW = [torch.zeros_like(w) for w in self.W]
        self.mb = [torch.zeros_like(b) for b in self.b]
        self.vb = [torch.zeros_like(b) for b in self.b]
        self.step_count = 0
        self.x_mean = torch.zeros(n_in)
        self.x_std = torch.ones(n_in)
        self.y_mean, self.y_std = 0.0, 1.0
        self._refresh_bf16()

    # -- helpers ------------------------------------------------------------------------
    def _s(self) -> int:
        return L.stream_ptr(self.device)

    def _refresh_bf16(self) -> None:
        for w, wb in zip(self.W, self.Wb):
            L.check(L.lib().sp_cast_f32_bf16(w.data_ptr(), w.numel(), wb.data_ptr(), self._s()), "cast")

    def _gemm(self, ta, tb, M, N, K, A, lda, B, ldb, Cm, ldc, alpha=1.0, acc=0):
        L.check(L.lib().sp_gemm_bf16(ta, tb, M, N, K, A.data_ptr(), lda, B.data_ptr(), ldb,
                                     Cm.data_ptr(), ldc, alpha, acc, self._s()), "gemm")

    def to_bf16(self, x: torch.Tensor) -> torch.Tensor:
        x = x.to(self.device, torch.float32).contiguous()
        out = torch.empty(x.shape, dtype=torch.bfloat16, device=self.device)
        L.check(L.lib().sp_cast_f32_bf16(x.data_ptr(), x.numel(), out.data_ptr(), self._s()), "cast")
        return out

    # -- forward / backward --------------------------------------------------------------
    def forward(self, xb: torch.Tensor) -> torch.Tensor:
        """xb: (B, n_in) bf16 on the device -> (B,) fp32 predictions (normalised target)."""
        B = xb.shape[0]
        self._acts = [xb]
        self._z = []
        a = xb
        nl = len(self.W)
        for l in range(nl):
            K, N = self.dims[l], self.dims[l + 1]
            zpre = torch.empty(B, N, device=self.device)
            self._gemm(0, 0, B, N, K, a, K, self.Wb[l], N, zpre, N)
            z = torch.empty(B, N, device=self.device)
            last = l == nl - 1
            an = None if last else torch.empty(B, N, dtype=torch.bfloat16, device=self.device)
            L.check(L.lib().sp_mlp_bias_act(zpre.data_ptr(), self.b[l].data_ptr(), B, N,
                                            0 if last else 1, z.data_ptr(), _p(an), self._s()),
                    "bias_act")
            self._z.append(z)
            if not last:
                self._acts.append(an)
                a = an
        return self._z[-1].reshape(B)

    def backward(self, target: torch.Tensor) -> torch.Tensor:
        """MSE against ``target`` (B,) fp32; fills self.gW / self.gb; returns the loss (device)."""
        y = self._z[-1].reshape(-1)
        B = y.shape[0]
        t = target.to(self.device, torch.float32).contiguous()
        dz = torch.empty(B, 1, device=self.device)
        dzb = torch.empty(B, 1, dtype=torch.bfloat16, device=self.device)
        loss = torch.empty(1, device=self.device)
        L.check(L.lib().sp_mlp_mse(y.data_ptr(), t.data_ptr(), B, dz.data_ptr(), dzb.data_ptr(),
                                   loss.data_ptr(), self._s()), "mse")
        nl = len(self.W)
        self.gW = [None] * nl
        self.gb = [None] * nl
        for l in range(nl - 1, -1, -1):
            K, N = self.dims[l], self.dims[l + 1]
            gW = torch.empty(K, N, device=self.device)
            self._gemm(1, 0, K, N, B, self._acts[l], K, dzb, N, gW, N)
            gb = torch.empty(N, device=self.device)
            L.check(L.lib().sp_mlp_colsum(dz.data_ptr(), B, N, gb.data_ptr(), self._s()), "colsum")
            self.gW[l], self.gb[l] = gW, gb
            if l == 0:
                break
            da = torch.empty(B, K, device=self.device)
            self._gemm(0, 1, B, K, N, dzb, N, self.Wb[l], N, da, K)
            dz = torch.empty(B, K, device=self.device)
            dzb = torch.empty(B, K, dtype=torch.bfloat16, device=self.device)
            L.check(L.lib().sp_mlp_relu_grad(da.data_ptr(), self._z[l - 1].data_ptr(), B * K,
                                             dz.data_ptr(), dzb.data_ptr(), self._s()), "relu_grad")
        return loss

    def adam_step(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999,
                  eps: float = 1e-8, weight_decay: float = 0.0) -> None:
        self.step_count += 1
        s = self._s()
        for l in range(len(self.W)):
            L.check(L.lib().sp_mlp_adam(self.W[l].data_ptr(), self.gW[l].data_ptr(),
                                        self.mW[l].data_ptr(), self.vW[l].data_ptr(),
                                        self.W[l].numel(), lr, beta1, beta2, eps, self.step_count,
                                        weight_decay, self.Wb[l].data_ptr(), s), "adam")
            L.check(L.lib().sp_mlp_adam(self.b[l].data_ptr(), self.gb[l].data_ptr(),
                                        self.mb[l].data_ptr(), self.vb[l].data_ptr(),
                                        self.b[l].numel(), lr, beta1, beta2, eps, self.step_count,
                                        0.0, None, s), "adam")

    # -- training / inference on raw features ------------------------------------------------
    def fit(self, X: np.ndarray, y: np.ndarray, epochs: int = 200, batch: int = 512,
            lr: float = 2e-3, seed: int = 0, verbose: bool = False,
            weight_decay: float = 0.0) -> List[float]:
        X = torch.as_tensor(X, dtype=torch.float32)
        y = torch.as_tensor(y, dtype=torch.float32)
        self.x_mean, self.x_std = X.mean(0), X.std(0).clamp(min=1e-6)
        self.y_mean, self.y_std = float(y.mean()), float(max(y.std().item(), 1e-6))
        Xn = ((X - self.x_mean) / self.x_std)
        yn = (y - self.y_mean) / self.y_std
        Xd = self.to_bf16(Xn)
        yd = yn.to(self.device)
        n = Xd.shape[0]
        g = torch.Generator().manual_seed(seed)
        losses = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).to(self.device)
            tot = torch.zeros(1, device=self.device)
            nb = 0
            for i in range(0, n, batch):
                idx = perm[i:i + batch]
                xb = Xd.index_select(0, idx).contiguous()
                self.forward(xb)
                loss = self.backward(yd.index_select(0, idx))
                self.adam_step(lr * 0.5 * (1 + math.cos(math.pi * ep / epochs)),
                               weight_decay=weight_decay)
                tot += loss
                nb += 1
            losses.append(float(tot.item()) / nb)
            if verbose and ep % 5 == 0:
                print(f"epoch {ep} loss {losses[-1]:.4f}", flush=True)
        return losses

    def predict(self, X: np.ndarray) -> np.ndarray:
        X = torch.as_tensor(X, dtype=torch.float32)
        Xd = self.to_bf16((X - self.x_mean) / self.x_std)
        with torch.no_grad():
            out = self.forward(Xd)
        return (out.cpu().numpy() * self.y_std + self.y_mean)

    # -- persistence ---------------------------------------------------------------------------
    def state_dict(self) -> Dict[str, object]:
        return {"dims": self.dims, "W": [w.cpu() for w in self.W], "b": [b.cpu() for b in self.b],
                "x_mean": self.x_mean, "x_std": self.x_std, "y_mean": self.y_mean,
                "y_std": self.y_std}

    def load_state_dict(self, st: Dict[str, object]) -> None:
        for l, (w, b) in enumerate(zip(st["W"], st["b"])):
            self.W[l].copy_(w)
            self.b[l].copy_(b)
        self.x_mean, self.x_std = st["x_mean"], st["x_std"]
        self.y_mean, self.y_std = float(st["y_mean"]), float(st["y_std"])
        self._refresh_bf16()

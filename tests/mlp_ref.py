"""Plain PyTorch FP32 reference of the surrogate network's kernels (test code).

`forward` / `backward` restate the network of include/sfmlp.h with the same
BF16 rounding points as the kernels when `rounded` is set (inputs, weights,
hidden activations and pre-activation gradients rounded to BF16, products
accumulated in FP32), and without any rounding otherwise.  The unrounded
manual backward is pinned against torch autograd on the CPU.
"""
from __future__ import annotations

import math

import torch

DIMS = (32, 256, 256, 256, 256, 32)
OUT = 6


def bf16(x: torch.Tensor) -> torch.Tensor:
    return x.to(torch.bfloat16).to(torch.float32)


def gelu(z: torch.Tensor) -> torch.Tensor:
    return 0.5 * z * (1.0 + torch.erf(z / math.sqrt(2.0)))


def gelu_grad(z: torch.Tensor) -> torch.Tensor:
    return 0.5 * (1.0 + torch.erf(z / math.sqrt(2.0))) + z * torch.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(Ws, bs, x, rounded: bool = True):
    """Ws[l]: [out, in] padded FP32 master weights; x: [B, 32] FP32 (BF16 values).
    Returns the padded output [B, 32] and the cache for backward."""
    r = bf16 if rounded else (lambda t: t)
    H, Z = [x], []
    h = x
    for l in range(4):
        z = h @ r(Ws[l]).T + bs[l]
        Z.append(z)
        h = r(gelu(z))
        H.append(h)
    y = h @ r(Ws[4]).T + bs[4]
    return y, (H, Z)


def loss_and_grad_out(y, target):
    """Mean squared error over the 6 real outputs and dLoss/dy (padded)."""
    B = y.shape[0]
    d = y[:, :OUT] - target
    loss = (d * d).sum() / (B * OUT)
    dy = torch.zeros_like(y)
    dy[:, :OUT] = 2.0 * d / (B * OUT)
    return loss, dy


def backward(Ws, cache, dy, rounded: bool = True):
    """Gradients (dW[l], db[l]) for l = 0..4."""
    r = bf16 if rounded else (lambda t: t)
    H, Z = cache
    dz = r(dy)
    dWs, dbs = [None] * 5, [None] * 5
    for l in range(4, -1, -1):
        dWs[l] = dz.T @ H[l]
        dbs[l] = dz.sum(0)
        if l > 0:
            dh = dz @ r(Ws[l])
            dz = r(dh * gelu_grad(Z[l - 1]))
    return dWs, dbs


def init_weights(seed: int, scale: float = 1.0):
    """Random padded weights with zero padding (real dims 13 in, 6 out)."""
    g = torch.Generator().manual_seed(seed)
    real = (13, 256, 256, 256, 256, 6)
    Ws, bs = [], []
    for l in range(5):
        w = torch.zeros(DIMS[l + 1], DIMS[l])
        b = torch.zeros(DIMS[l + 1])
        bound = scale / math.sqrt(real[l])
        w[:real[l + 1], :real[l]] = (torch.rand(real[l + 1], real[l], generator=g) * 2 - 1) * bound
        b[:real[l + 1]] = (torch.rand(real[l + 1], generator=g) * 2 - 1) * 0.1
        Ws.append(w)
        bs.append(b)
    return Ws, bs


def random_inputs(seed: int, B: int):
    g = torch.Generator().manual_seed(seed)
    x = torch.zeros(B, 32)
    x[:, :13] = torch.randn(B, 13, generator=g)
    t = torch.randn(B, OUT, generator=g)
    return bf16(x), t

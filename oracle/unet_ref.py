"""CPU reference of the learned preconditioner's network: the 3-level U-Net in torch FP64 autograd.

TEST INFRASTRUCTURE ONLY: tests/ check the GPU network (libvfxsim3d.so vfx_net_*) against it; the
product never imports it.  The reference has no network (its solver is MGPCG only), so this is a
restatement of the build's own network definition with torch's conv3d / avg_pool3d as the
arithmetic, pinned by torch's autograd (a library routine, independent of the GPU kernels).

Rounding: the GPU stores activations and the gradients with respect to stored activations in
BF16 and accumulates in FP32.  `rbf` rounds to BF16 in the forward pass AND in the backward pass,
and is placed exactly where the GPU stores a value, so the FP64 result differs from the GPU's
only by FP32 accumulation order (and the occasional BF16 tie it moves).
"""
from __future__ import annotations

import torch
import torch.nn.functional as F

LAYERS = ((2, 16, True), (16, 16, True), (16, 32, True), (32, 32, True), (32, 64, True), (64, 64, True),
          (96, 32, True), (32, 32, True), (48, 16, True), (16, 16, True), (16, 1, False))


class _RoundBF16(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        return x.to(torch.bfloat16).to(torch.float64)

    @staticmethod
    def backward(ctx, g):
        return g.to(torch.bfloat16).to(torch.float64)


class _RoundBF16Forward(torch.autograd.Function):
    """BF16 copy of the FP32 master weights: rounded forward, the FP32 gradient passes through."""

    @staticmethod
    def forward(ctx, x):
        return x.to(torch.float32).to(torch.bfloat16).to(torch.float64)

    @staticmethod
    def backward(ctx, g):
        return g


def rbf(x):
    return _RoundBF16.apply(x)


def rbf_weight(w):
    return _RoundBF16Forward.apply(w)


def _up(x):
    return x.repeat_interleave(2, 2).repeat_interleave(2, 3).repeat_interleave(2, 4)


def forward(params, inp):
    """params: [(w (cout, cin, 3, 3, 3), b (cout,)) float64 tensors requiring grad]; inp (2, N, N, N)
    float64 -> out (N, N, N) float64 (the GPU's FP32 output)."""
    def conv(x, i):
        w, b = params[i]
        ci = LAYERS[i][0]
        y = F.conv3d(x[:, :ci], rbf_weight(w), b, padding=1)   # weights are used as BF16
        if LAYERS[i][2]:
            return rbf(torch.relu(y))
        return y

    N = inp.shape[-1]
    x = torch.zeros(1, 16, N, N, N, dtype=torch.float64)
    x[0, :2] = inp
    x = rbf(x)
    e0a = conv(x, 0)
    e0b = conv(e0a, 1)
    p0 = rbf(F.avg_pool3d(e0b, 2))
    e1a = conv(p0, 2)
    e1b = conv(e1a, 3)
    p1 = rbf(F.avg_pool3d(e1b, 2))
    e2a = conv(p1, 4)
    e2b = conv(e2a, 5)
    c1 = rbf(torch.cat([_up(e2b), e1b], 1))
    d1a = conv(c1, 6)
    d1b = conv(d1a, 7)
    c0 = rbf(torch.cat([_up(d1b), e0b], 1))
    d0a = conv(c0, 8)
    d0b = conv(d0a, 9)
    return conv(d0b, 10)[0, 0]


def grads(flat_params, inp, grad_out):
    """(out, flat parameter gradient) of the network for d loss / d out = grad_out (rounded to
    BF16 first, as the GPU stores it)."""
    import numpy as np
    params, k = [], 0
    for ci, co, _ in LAYERS:
        w = torch.tensor(np.asarray(flat_params[k:k + co * ci * 27], np.float64).reshape(co, ci, 3, 3, 3),
                         requires_grad=True)
        k += co * ci * 27
        b = torch.tensor(np.asarray(flat_params[k:k + co], np.float64), requires_grad=True)
        k += co
        params.append((w, b))
    out = forward(params, torch.as_tensor(np.asarray(inp), dtype=torch.float64))
    g = torch.as_tensor(np.asarray(grad_out), dtype=torch.float64).to(torch.bfloat16).to(torch.float64)
    out.backward(g)
    flat = torch.cat([t for w, b in params for t in (w.grad.reshape(-1), b.grad)])
    return out.detach().numpy(), flat.numpy()

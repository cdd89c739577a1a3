"""GPU tests of the learned preconditioner's network (BF16 MFMA U-Net, vfx_net_*): forward and
backward against the CPU reference (oracle/unet_ref.py, torch FP64 autograd with the GPU's BF16
rounding points), Adam against its closed form."""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

import unet_ref
from vfx.sim3d.learned import LAYERS, UNet, init_params, split_params

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _inputs(N, seed):
    g = np.random.default_rng(seed)
    inp = np.zeros((2, N, N, N), np.float32)
    inp[0] = g.standard_normal((N, N, N))
    inp[1] = (g.random((N, N, N)) > 0.2).astype(np.float32)
    return inp


@pytest.mark.parametrize("N", [8, 16])
def test_net_forward_backward_match_reference(cuda, hiplib, N):
    """Output and every layer's weight / bias gradient equal the FP64 reference (same BF16
    rounding points) to FP32-accumulation level."""
    _say(f"unet fwd/bwd N={N}")
    params = init_params(seed=3)
    # non-zero biases so that their gradients are exercised too
    rng = np.random.default_rng(4)
    k = 0
    for ci, co, _ in LAYERS:
        k += co * ci * 27
        params[k:k + co] = rng.uniform(-0.05, 0.05, co).astype(np.float32)
        k += co
    net = UNet(N, params=params)
    inp = _inputs(N, 5)
    gout = np.random.default_rng(6).standard_normal((N, N, N)).astype(np.float32)
    out = net.forward(torch.from_numpy(inp).to(DEV))
    grad = net.backward(torch.from_numpy(gout).to(DEV))
    torch.cuda.synchronize()
    ref_out, ref_grad = unet_ref.grads(params, inp, gout)
    e = _rel_l2(out.cpu().numpy(), ref_out)
    _say(f"  output rel L2 {e:.2e}")
    assert e <= 2e-3
    g = grad.cpu().numpy()
    k = 0
    for i, (ci, co, _) in enumerate(LAYERS):
        nw = co * ci * 27
        ew = _rel_l2(g[k:k + nw], ref_grad[k:k + nw])
        eb = _rel_l2(g[k + nw:k + nw + co], ref_grad[k + nw:k + nw + co])
        _say(f"  layer {i} ({ci}->{co}): weight grad {ew:.2e}, bias grad {eb:.2e}")
        assert ew <= 1e-2 and eb <= 1e-2, i
        k += nw + co


def test_net_adam_closed_form(cuda, hiplib):
    """Two Adam steps equal the closed form (Kingma & Ba 2015, bias-corrected)."""
    N = 8
    p0 = init_params(seed=1)
    net = UNet(N, params=p0)
    rng = np.random.default_rng(2)
    lr, b1, b2, eps = 1e-3, 0.9, 0.999, 1e-8
    m = np.zeros_like(p0, dtype=np.float64)
    v = np.zeros_like(p0, dtype=np.float64)
    p = p0.astype(np.float64)
    for t in (1, 2):
        g = rng.standard_normal(p0.shape).astype(np.float32)
        net.adam(torch.from_numpy(g).to(DEV), lr, b1, b2, eps)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g.astype(np.float64) ** 2
        p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    torch.cuda.synchronize()
    got = net.params()
    e = float(np.abs(got - p).max() / np.abs(p).max())
    _say(f"  adam max rel err {e:.2e}")
    assert e <= 1e-5

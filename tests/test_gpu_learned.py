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


@pytest.mark.parametrize("N", [8, 16, 32, 64])
def test_net_forward_backward_match_reference(cuda, hiplib, N):
    """Output and every layer's weight / bias gradient equal the FP64 reference (same BF16
    rounding points) to FP32-accumulation level.  The sizes cover every launch shape of the
    convolutions: K split over 4 and 2 waves (grids of fewer than 4096 tiles of 16 voxels), one
    wave per tile and 4 tiles per wave (16384 tiles and more), the LDS-tiled kernel (grids from 48).  The difference is BF16 rounding
    flips (a sum that lands within its accumulation error of a BF16 rounding boundary rounds the
    other way) carried through the 11 layers; it grows with the grid (4e-5, 8e-4, 5e-3 at 8, 16,
    32: the earlier single-shape kernel gave the same), so the bound widens with N while staying
    far below what a lost K slice (about 1/14 of a layer's sum) would give.  The gradients flip
    the same way through the backward's BF16 stores (5e-3 at 16, 5e-2 at 32); at 64 only the
    output is compared."""
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
    tol_out, tol_grad = {8: (2e-3, 1e-2), 16: (2e-3, 1e-2), 32: (2e-2, 1.5e-1), 64: (5e-2, None)}[N]
    assert e <= tol_out
    if tol_grad is None:
        return
    g = grad.cpu().numpy()
    k = 0
    for i, (ci, co, _) in enumerate(LAYERS):
        nw = co * ci * 27
        ew = _rel_l2(g[k:k + nw], ref_grad[k:k + nw])
        eb = _rel_l2(g[k + nw:k + nw + co], ref_grad[k + nw:k + nw + co])
        _say(f"  layer {i} ({ci}->{co}): weight grad {ew:.2e}, bias grad {eb:.2e}")
        assert ew <= tol_grad and eb <= tol_grad, i
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


# ---------------------------------------------------------------------------- flexible CG
import dataclasses  # noqa: E402

from vfx.sim3d import Solver3D, _lib as _vl  # noqa: E402
from vfx.sim3d.learned import generate_pairs, noise_pairs, random_scene, scene_config, train  # noqa: E402
from vfx.sim3d.solver import pack_params  # noqa: E402


def _pre_projection_state(N, seed, steps=6):
    """State u* of a random scene right before a projection (advection + forces of step `steps`)."""
    forces, obst, beta = random_scene(seed)
    cfg = scene_config(beta, seed)
    s = Solver3D(N, cfg)
    st = s.reset(obst)
    for i in range(steps):
        s.step(st, forces, 1.0 / 60.0)
    h = pack_params(cfg, N, forces, 1.0 / 60.0, st.meta.step)
    for stage in (_vl.STAGE_ADVECT, _vl.STAGE_FORCES):
        _vl.check(s._lib.vfx_step_stage(s._ctx, stage, _ptr(st.data), _ptr(st.solid),
                                        h.ctypes.data_as(__import__("ctypes").c_void_p) if stage == 0 else None,
                                        _stream()))
    torch.cuda.synchronize()
    return st, cfg


def _ptr(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    import ctypes
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _project(N, st, cfg, net=None, mode=0, maxit=None, sweeps=0, sync="device", reps=1):
    """Cold projection of a copy of st (relative residual 1e-5), `reps` times (the last is
    returned); returns (stats, projected data).  sweeps: damped-Jacobi corrections after each
    application of a full-resolution network (0 = the bare network)."""
    cfg = dataclasses.replace(cfg, pressure_warm_start=False, pressure_rel_tol=1e-5, pressure_max_iters=400,
                              pressure_restarts=2, pressure_sync=sync)
    s = Solver3D(N, cfg)
    s.set_option(_vl.OPT_TRUE_RESIDUAL, 1)
    if mode:
        s.set_option(_vl.OPT_FCG_MG, mode)
    if maxit is not None:
        s.set_option(_vl.OPT_LEARNED_MAXIT, maxit)
    s.set_option(_vl.OPT_LEARNED_SWEEPS, sweeps)
    if net is not None:
        s.set_preconditioner(net)
    from vfx.core.fields3d import FieldState3D
    for _ in range(reps):
        data = st.data.clone()
        work = FieldState3D(data=data, solid=st.solid.clone(), meta=st.meta)
        s.project(work)
        ps = s.pressure_stats()
    ps["history"] = [float(f"{h / max(ps['b_norm'], 1e-300):.2e}") for h in s.fcg_history()]
    return ps, data


def test_fcg_with_mg_matches_pcg(cuda, hiplib):
    """Flexible CG with the (fixed, symmetric) MG V-cycle as preconditioner takes PCG's iterations
    and reaches the same projected field."""
    N = 48
    st, cfg = _pre_projection_state(N, 11)
    pcg, a = _project(N, st, cfg)
    fcg, b = _project(N, st, cfg, mode=1)
    _say(f"  PCG {pcg['iters']} iterations (true rel-res {pcg['true_rel_res']:.2e}), "
         f"FCG+MG {fcg['learned_iters']} (true rel-res {fcg['true_rel_res']:.2e}, fallback {fcg['fallback']})")
    assert fcg["fallback"] == 0
    assert abs(fcg["learned_iters"] - pcg["iters"]) <= 1
    assert pcg["true_rel_res"] <= 1e-5 and fcg["true_rel_res"] <= 1e-5
    v = a[1:4].cpu().numpy()
    e = float(np.abs(b[1:4].cpu().numpy() - v).max() / np.abs(v).max())
    _say(f"  velocity max rel difference {e:.2e}")
    assert e <= 1e-3


@pytest.mark.parametrize("bad", ["zero", "nan", "budget"])
def test_learned_fallback(cuda, hiplib, bad):
    """A useless network (zero output, NaN weights) or an exhausted budget falls back to MGPCG,
    which converges: the true residual decides."""
    N = 32
    st, cfg = _pre_projection_state(N, 12, steps=4)
    p = init_params(0)
    if bad == "zero":
        p[:] = 0.0
    elif bad == "nan":
        p[:100] = np.nan
    net = UNet(N, params=p)
    ps, _ = _project(N, st, cfg, net=net, maxit=1 if bad == "budget" else None)
    _say(f"  {bad}: fallback {ps['fallback']}, learned iterations {ps['learned_iters']}, total {ps['iters']}, "
         f"true rel-res {ps['true_rel_res']:.2e}")
    assert ps["fallback"] == 1
    assert ps["true_rel_res"] <= 1e-5


def test_trained_network_gains(cuda, hiplib):
    """Trained on solver-generated (b, phi) pairs (plus consistent multi-scale noise pairs), the
    network preconditions flexible CG on a held-out scene: after the same number of iterations its
    residual is far below that of CG without it and below the untrained network's, and the
    projection converges (the true residual judges; MGPCG takes over if the network stalls)."""
    N = 32
    pairs = generate_pairs(N, seeds=range(20, 26), steps=14, every=1)
    pairs += noise_pairs(N, [p[0] for p in pairs[::11]], count=len(pairs), seed=5)
    _say(f"  {len(pairs)} training pairs at {N}^3")
    untrained = UNet(N, seed=1)
    net = UNet(N, params=untrained.params())
    losses = train(net, pairs, iters=800, lr=2e-3, seed=0, log=_say)
    first, last = float(np.mean(losses[:50])), float(np.mean(losses[-50:]))
    _say(f"  loss {first:.4f} -> {last:.4f}")
    assert last < 0.5 * first
    st, cfg = _pre_projection_state(N, 99)
    plain, _ = _project(N, st, cfg, mode=2)
    before, _ = _project(N, st, cfg, net=untrained)
    after, _ = _project(N, st, cfg, net=net)
    hybrid, _ = _project(N, st, cfg, net=net, sweeps=2)
    mg, _ = _project(N, st, cfg)
    _say(f"  trained net + 2 Jacobi corrections: {hybrid['learned_iters']} its (fallback {hybrid['fallback']}), "
         f"history {hybrid['history']}")
    _say(f"  held-out scene: plain CG {plain['learned_iters']} its (fallback {plain['fallback']}), untrained net "
         f"{before['learned_iters']} (fallback {before['fallback']}), trained net {after['learned_iters']} (fallback "
         f"{after['fallback']}, true rel-res {after['true_rel_res']:.2e}, {after['iters']} MGPCG), MGPCG {mg['iters']}")
    _say(f"  relative residual histories: plain CG {plain['history']}")
    _say(f"    untrained net {before['history']}")
    _say(f"    trained net {after['history']}")
    assert after["true_rel_res"] <= 1e-5
    k = min(len(after["history"]), len(plain["history"])) - 1
    assert k >= 3
    assert after["history"][k] <= 0.3 * plain["history"][k]
    kb = min(len(after["history"]), len(before["history"])) - 1
    assert after["history"][kb] < before["history"][kb] or before["fallback"] == 1
    # with the Jacobi corrections (the library default) the network alone reaches 1e-5
    assert hybrid["fallback"] == 0 and hybrid["true_rel_res"] <= 1e-5


def _coarse_net(N, iters=300):
    """The V-cycle's coarse correction (a U-Net on the (N/4)^3 grid) trained briefly at N^3 on what
    the V-cycle leaves of solver pairs of seeded scenes 20-23 and noise pairs."""
    from vfx.sim3d.learned import train_coarse, vcycle_pairs
    pairs = generate_pairs(N, seeds=range(20, 24), steps=10, every=1)
    pairs += noise_pairs(N, [p[0] for p in pairs[::11]], count=len(pairs) // 2, seed=5)
    net = UNet(N // 4, seed=1)
    losses = train_coarse(net, vcycle_pairs(N, pairs), iters=iters, lr=2e-3, seed=0)
    _say(f"  coarse correction at {N}^3: loss {np.mean(losses[:30]):.3f} -> {np.mean(losses[-30:]):.3f}")
    return net


@pytest.mark.parametrize("N", [64, 96])
def test_learned_device_matches_polled(cuda, hiplib, N):
    """The coarse correction's flexible CG sequenced on the device (no host synchronisation; the
    full z = z_V + P c is never stored, its dots come from block sums) takes the iterations of the
    host-polled flexible CG, which stores z and reduces its dots directly, reaches the same
    projected velocity, and both meet the true relative residual 1e-5 without fallback."""
    net = _coarse_net(N)
    st, cfg = _pre_projection_state(N, 97)
    polled, a = _project(N, st, cfg, net=net, sync="poll")
    dev, b = _project(N, st, cfg, net=net, sync="device", reps=2)
    mg, _ = _project(N, st, cfg)
    _say(f"  {N}^3: polled {polled['learned_iters']} its (fallback {polled['fallback']}, true rel-res "
         f"{polled['true_rel_res']:.2e}), device {dev['learned_iters']} its (fallback {dev['fallback']}, "
         f"{dev['iters']} MGPCG, true rel-res {dev['true_rel_res']:.2e}), MGPCG {mg['iters']}")
    assert polled["fallback"] == 0 and dev["fallback"] == 0
    assert abs(dev["learned_iters"] - polled["learned_iters"]) <= 1
    assert polled["true_rel_res"] <= 1e-5 and dev["true_rel_res"] <= 1e-5 and dev["converged"]
    v = a[1:4].cpu().numpy()
    e = float(np.abs(b[1:4].cpu().numpy() - v).max() / np.abs(v).max())
    _say(f"  velocity max rel difference {e:.2e}")
    assert e <= 1e-3


@pytest.mark.parametrize("bad", ["nan", "limit"])
def test_learned_device_fallback(cuda, hiplib, bad):
    """On the device, a network that breaks the flexible CG (NaN weights) or a learned run cut at
    one iteration hands the solve to the gated MGPCG run, which converges from the current phi."""
    N = 64
    p = init_params(0)
    if bad == "nan":
        p[:] = np.nan   # (a ReLU layer maps NaN to 0; the last layer's NaN bias reaches the output)
    net = UNet(N // 4, params=p)
    st, cfg = _pre_projection_state(N, 12, steps=4)
    ps, data = _project(N, st, cfg, net=net, maxit=1 if bad == "limit" else None, reps=3)
    _say(f"  {bad}: fallback {ps['fallback']}, learned iterations {ps['learned_iters']}, MGPCG {ps['iters']}, "
         f"true rel-res {ps['true_rel_res']:.2e}, budget stops {ps['budget_stops']}")
    assert ps["fallback"] == 1 and ps["iters"] > 0
    assert 0 <= ps["true_rel_res"] <= 1e-5 and ps["converged"]
    assert torch.isfinite(data).all()


def test_learned_step_graph_replay(cuda, hiplib):
    """Steps with the coarse correction on the device replay as captured step graphs and equal the
    same steps enqueued without graphs bit for bit; step() returns before the GPU has finished."""
    from vfx.sim3d import Solver3D
    N = 64
    net = _coarse_net(N, iters=100)
    forces, obst, beta = random_scene(96)
    cfg = dataclasses.replace(scene_config(beta, 96), pressure_rel_tol=1e-5, pressure_restarts=1)
    runs = []
    for graph in (True, False):
        Solver3D.STEP_GRAPH = graph
        try:
            s = Solver3D(N, cfg)
            s.set_preconditioner(net)
            st = s.reset(obst)
            for _ in range(12):
                s.step(st, forces, 1.0 / 60.0)
            ps = s.pressure_stats()
        finally:
            Solver3D.STEP_GRAPH = True
        _say(f"  graph {graph}: learned its {ps['learned_iters']}, fallback {ps['fallback']}, captures "
             f"{ps['graph_captures']}, replays {ps['graph_replays']}, unconverged {ps['unconverged']}")
        assert ps["converged"]
        if graph:
            assert ps["graph_replays"] > 0
        runs.append(st.data.clone())
    assert torch.equal(runs[0], runs[1])

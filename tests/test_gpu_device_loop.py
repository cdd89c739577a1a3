"""GPU tests of the device-sequenced projection (VFX_OPT_SYNC = device, the default) and of whole
steps replayed as captured hipGraphs: step() does not wait for the device, its results equal the
host-polled loop's bit for bit while the iteration budget suffices, a budget that runs out is
reported and grows, and batched frames grow their budgets without a wait.

The reference runs its PCG loop on the device and replays the whole step as one CUDA graph
(vfx/sim3d/solver.py:146-155, :388-420; pressure.py:940-950)."""
from __future__ import annotations

import dataclasses
import sys
import time

import numpy as np
import pytest
import torch

from vfx.sim3d import Solver3D, _lib
from vfx.sim3d.batch import BatchSolver3D
from vfx.sim3d.learned import random_scene, scene_config

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
DT = 1.0 / 60.0


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _scene(seed=21, **kw):
    forces, obst, beta = random_scene(seed)
    cfg = dataclasses.replace(scene_config(beta, seed), **kw)
    return forces, obst, cfg


def test_step_returns_before_the_device_finishes(cuda, hiplib):
    """With the GPU busy (a spin kernel queued first), ten step() calls return to the host while the
    stream is still running: no host synchronisation inside step(), restarts included."""
    forces, obst, cfg = _scene(pressure_restarts=2)
    N = 48
    _say(f"no host wait: {N}^3, restarts 2")
    s = Solver3D(N, cfg, device=DEV)
    st = s.reset(obst)
    for _ in range(4):
        s.step(st, forces, DT)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    torch.cuda._sleep(300_000_000)   # about 0.1 s or more of device time
    t0 = time.perf_counter()
    for _ in range(10):
        s.step(st, forces, DT)
    host_ms = 1000.0 * (time.perf_counter() - t0)
    busy = not stream.query()
    t1 = time.perf_counter()
    torch.cuda.synchronize()
    wait_ms = 1000.0 * (time.perf_counter() - t1)
    ps = s.pressure_stats()
    _say(f"  10 steps enqueued in {host_ms:.1f} ms; stream busy after them: {busy}; then waited {wait_ms:.1f} ms; "
         f"graphs captured {ps['graph_captures']} replayed {ps['graph_replays']}")
    assert busy, "step() waited for the device"
    assert wait_ms > 20.0, (host_ms, wait_ms)
    assert ps["graph_replays"] >= 10 and ps["converged"]
    assert torch.isfinite(st.data).all()


@pytest.mark.parametrize("graph", [True, False])
def test_device_loop_equals_polled_loop(cuda, hiplib, graph):
    """The device-sequenced projection (graph replays or direct launches) equals the host-polled
    loop bit for bit over 20 steps with a restart allowed, as long as its budgets suffice; both
    meet the true relative residual 1e-5."""
    forces, obst, cfg = _scene(pressure_restarts=1)
    N = 48
    _say(f"device vs polled loop: {N}^3, graph={graph}")
    dev = Solver3D(N, cfg, device=DEV)
    dev.STEP_GRAPH = graph
    pol = Solver3D(N, dataclasses.replace(cfg, pressure_sync="poll"), device=DEV)
    sa, sb = dev.reset(obst), pol.reset(obst)
    for k in range(20):
        dev.step(sa, forces, DT)
        pol.step(sb, forces, DT)
        pa, pb = dev.pressure_stats(), pol.pressure_stats()
        assert pa["iters"] == pb["iters"] and pa["restarts"] == pb["restarts"], (k, pa, pb)
        assert torch.equal(sa.data, sb.data), (k, (sa.data - sb.data).abs().max().item())
        assert pa["converged"] and pb["converged"], (k, pa, pb)
        assert 0 <= pa["true_rel_res"] <= 1e-5, (k, pa)
        if k % 5 == 4:
            _say(f"  step {k}: iters {pa['iters']} budget {pa['budget']} restarts {pa['restarts']} "
                 f"true rel-res {pa['true_rel_res']:.2e}")
    assert pa["budget_stops"] == 0
    assert (pa["graph_replays"] > 0) == graph
    assert pb["budget"] == 0 and pb["graph_replays"] == 0


def test_budget_stop_is_reported_and_grows(cuda, hiplib):
    """A fixed budget of 2 iterations stops the projection there (reported unconverged, counted as a
    budget stop); back on the adaptive budget, it grows from the read-back counts until the
    projections converge again."""
    forces, obst, cfg = _scene(pressure_restarts=0, pressure_warm_start=False)
    N = 48
    _say(f"budget stop: {N}^3")
    s = Solver3D(N, cfg, device=DEV)
    st = s.reset(obst)
    for _ in range(5):
        s.step(st, forces, DT)
    full = s.pressure_stats()
    assert full["converged"] and full["iters"] > 3, full
    s.set_option(_lib.OPT_CG_BUDGET, 2)
    s.step(st, forces, DT)
    ps = s.pressure_stats()
    _say(f"  budget 2: iters {ps['iters']} converged {ps['converged']} budget stops {ps['budget_stops']}")
    assert ps["iters"] == 2 and not ps["converged"] and ps["budget_stops"] == 1
    s.set_option(_lib.OPT_CG_BUDGET, 0)
    budgets = []
    for _ in range(8):
        s.step(st, forces, DT)
        ps = s.pressure_stats()   # waits, so every read-back has arrived before the next step
        budgets.append(ps["budget"])
    _say(f"  adaptive budgets after it: {budgets}, last iters {ps['iters']}")
    assert ps["converged"] and budgets[-1] >= ps["iters"]
    assert torch.isfinite(st.data).all()


@pytest.mark.skip(reason="aborted the process on the GPU (inside vfx_batch_step, no GPU fault reported) when "
                         "the next frame read the last finished frame's state without waiting; the cause is not "
                         "found, so that reading is switched off and budgets grow in vfx_batch_wait only")
def test_batch_budgets_grow_without_wait(cuda, hiplib):
    """Batched frames whose projections need more iterations than their budget: frames enqueued
    one after another with no vfx_batch_wait in between still grow the budgets (the next frame
    reads the last finished frame's state without waiting), and the count of budget stops is
    available without a wait."""
    sizes = (48, 32)
    sc = [_scene(seed, pressure_restarts=0, pressure_warm_start=False, pressure_rel_tol=1e-6) for seed in (31, 32)]
    batch = BatchSolver3D(sizes, [c for _, _, c in sc], streams=2, graph=True, device=DEV)
    states = batch.reset([o for _, o, _ in sc])
    forces = [f for f, _, _ in sc]
    _say(f"batch budgets without wait: sizes {sizes}")
    for _ in range(3):
        batch.step(states, forces, DT)
    torch.cuda.synchronize()
    batch.set_budget(2)
    for fr in range(10):
        batch.step(states, forces, DT)
        torch.cuda.synchronize()   # the frame has finished; the next step() reads its state
    n_trunc = batch.truncated()
    budgets = batch.budgets
    info = batch.wait()
    _say(f"  budget stops {n_trunc}, budgets {budgets.tolist()}, last iters {info['iters'].tolist()} "
         f"converged {info['converged'].tolist()}")
    assert n_trunc >= 1
    assert np.all(budgets > 2) and np.all(budgets >= info["iters"]), (budgets, info)
    assert np.all(info["converged"]), info

"""GPU tests of batched frames (vfx_batch_*): several simulations of different sizes stepped as one
frame on several streams, directly and as a replayed hipGraph, against the same simulations
stepped one by one with Solver3D.step (host-polled PCG).  The per-simulation step is the one the
parity tests pin against the oracle, so equality with it carries that parity over."""
from __future__ import annotations

import dataclasses
import sys

import numpy as np
import pytest
import torch

from vfx.sim3d import Solver3D
from vfx.sim3d.batch import BatchSolver3D
from vfx.sim3d.learned import random_scene, scene_config

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
SIZES = (48, 32, 64)
SEEDS = (11, 12, 13)
DT = 1.0 / 60.0


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _scenes():
    """Seeded scenes; no restarts, so the single solvers run exactly the batch's kernels."""
    out = []
    for sd in SEEDS:
        forces, obst, beta = random_scene(sd)
        out.append((forces, obst, dataclasses.replace(scene_config(beta, sd), pressure_restarts=0)))
    return out


def _pair(streams=2, graph=True):
    """A batch and, for the same scenes, independent single solvers with equal initial states."""
    sc = _scenes()
    batch = BatchSolver3D(SIZES, [c for _, _, c in sc], streams=streams, graph=graph, device=DEV)
    states = batch.reset([o for _, o, _ in sc])
    singles = [Solver3D(n, c, device=DEV) for n, (_, _, c) in zip(SIZES, sc)]
    ref = [s.reset(o) for s, (_, o, _) in zip(singles, sc)]
    forces = [f for f, _, _ in sc]
    return batch, states, singles, ref, forces


def _step_singles(singles, ref, forces):
    for s, st, f in zip(singles, ref, forces):
        s.step(st, f, DT)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_frames_equal_single_steps(cuda, hiplib, graph):
    """Every frame, every simulation's fields equal its own Solver3D.step bit for bit (the budget
    is enough, so the extra enqueued iterations are no-ops); graph frames are replays of one
    capture, checked after each replay with eager work and synchronisation in between."""
    _say(f"batch frames == single steps, graph={graph}, sizes {SIZES}")
    batch, states, singles, ref, forces = _pair(graph=graph)
    for fr in range(5):
        batch.step(states, forces, DT)
        info = batch.wait()
        _step_singles(singles, ref, forces)
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(states, ref)):
            assert info["converged"][i], (fr, i, info)
            assert info["iters"][i] == singles[i].pressure_iters(), (fr, i)
            assert torch.equal(a.data, b.data), (fr, i, (a.data - b.data).abs().max().item())
        lat = info["latency_ms"]
        assert np.all(lat > 0) and np.all(lat <= info["frame_ms"] + 1e-3), info
        _say(f"  frame {fr}: iters {info['iters'].tolist()} latency {np.round(lat, 3).tolist()} "
             f"frame {info['frame_ms']:.3f} ms")
        assert states[0].meta.step == fr + 1


def test_budget_grows_when_used_up(cuda, hiplib):
    """A budget of 2 iterations stops the projections early (reported unconverged, finished like
    the reference's fixed-iteration mode); the budgets then grow and later frames converge."""
    _say("batch budget growth")
    batch, states, singles, ref, forces = _pair(graph=True)
    for _ in range(2):   # flow to project
        batch.step(states, forces, DT)
        batch.wait()
    batch.set_budget(2)
    batch.step(states, forces, DT)
    info = batch.wait()
    _say(f"  budget 2: iters {info['iters'].tolist()} converged {info['converged'].tolist()}")
    assert np.all(info["iters"] <= 2)
    assert not np.all(info["converged"])
    grown = batch.budgets
    assert np.all(grown > 2), grown
    for _ in range(4):
        batch.step(states, forces, DT)
        info = batch.wait()
    _say(f"  after growth: budgets {batch.budgets.tolist()} iters {info['iters'].tolist()}")
    assert np.all(info["converged"]), info
    for st in states:
        assert torch.isfinite(st.data).all()


def test_tune_keeps_states_and_picks_valid_shapes(cuda, hiplib):
    """The autotuner leaves every state as it was, reports settings from its candidate sets, and
    frames after it still match the single solvers (to FP64-reduction-order level: the cell
    kernels' block count changes the order of the FP64 partial sums)."""
    _say("batch autotuner")
    batch, states, singles, ref, forces = _pair(graph=True)
    for _ in range(3):
        batch.step(states, forces, DT)
        batch.wait()
        _step_singles(singles, ref, forces)
    before = [st.data.clone() for st in states]
    rep = batch.tune(states, forces, DT, reps=1)
    torch.cuda.synchronize()
    _say(f"  report {rep}")
    for a, b in zip(states, before):
        assert torch.equal(a.data, b)
    assert set(rep["sizes"]) == set(SIZES)
    for v in rep["sizes"].values():
        assert v["cells_per_thread"] in (1, 2, 4) and v["coarse_m"] in (4, 8, 16) and v["ms"] > 0
    assert rep["streams"] in (1, 2, 4, 8)
    for _ in range(3):
        batch.step(states, forces, DT)
        info = batch.wait()
        _step_singles(singles, ref, forces)
        assert np.all(info["converged"])
    torch.cuda.synchronize()
    for i, (a, b) in enumerate(zip(states, ref)):
        err = ((a.data - b.data).abs().max() / b.data.abs().max()).item()
        _say(f"  sim {i} (N={SIZES[i]}): max rel diff after tuning {err:.2e}")
        assert err < 1e-4


@pytest.mark.parametrize("zc", [2, 4])
def test_cells_per_thread_option(cuda, hiplib, zc):
    """VFX_OPT_CELLS_PER_THREAD changes only the launch shape: steps agree with zc = 1 to
    FP64-reduction-order level."""
    from vfx.sim3d import _lib
    _say(f"cells per thread {zc}")
    forces, obst, cfg = _scenes()[0]
    a, b = Solver3D(48, cfg, device=DEV), Solver3D(48, cfg, device=DEV)
    b.set_option(_lib.OPT_CELLS_PER_THREAD, zc)
    sa, sb = a.reset(obst), b.reset(obst)
    for _ in range(4):
        a.step(sa, forces, DT)
        b.step(sb, forces, DT)
    torch.cuda.synchronize()
    err = ((sa.data - sb.data).abs().max() / sa.data.abs().max()).item()
    _say(f"  max rel diff {err:.2e}")
    assert err < 1e-5

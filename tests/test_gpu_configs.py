"""GPU tests of the benchmark configurations at their stated workloads.

- Mixed precision (256^3 smoke plume past a sphere, 50 steps): the FP32 solver with FP32, FP16
  and BF16 multigrid level storage against the FP64 run of the same solver; every projection of
  every run meets the true relative residual 1e-5, and the pressure, divergence and density
  errors against FP64 are reported and bounded.
- Learned preconditioner (128^3, held-out scenes): the V-cycle's coarse correction trained on
  solver pairs, flexible CG with the MGPCG fallback against plain MGPCG at equal true relative
  residual (no fallback, no more iterations than MGPCG).
- Batched uneven sims (32 sims of sizes drawn by seed from {48, 64, 96, 128}^3): every projection
  of every frame converges, and sims of each size equal their own Solver3D.step bit for bit.

The per-simulation step is the one the parity tests pin against the oracle (reference
vfx/sim3d/solver.py:146-155), so these carry that parity to the full workloads.
"""
from __future__ import annotations

import ctypes
import dataclasses
import sys
import time

import numpy as np
import pytest
import torch

from vfx.core.fields3d import Emitter3D, FieldState3D, Forces3D, Obstacle3D
from vfx.sim3d import Solver3D, SolverConfig3D, _lib

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _plume_config(storage="fp32"):
    """The smoke plume past a sphere: buoyancy beta 1 per K, confinement 0.3, emitter noise
    U[0.8, 1.2], pressure to true relative residual 1e-5 (up to 2 restarts)."""
    return SolverConfig3D(buoyancy_alpha=0.0, buoyancy_beta=1.0, vorticity_eps=0.3, pressure_rel_tol=1e-5,
                          pressure_max_iters=200, pressure_restarts=2, emitter_noise=0.2, noise_seed=0,
                          mg_storage=storage)


PLUME = Forces3D(emitters=[Emitter3D(pos=(0.5, 0.1, 0.5), radius=0.08, density_rate=60.0, temperature=294.0,
                                     fuel_rate=0.0, velocity=(0.0, 0.0, 0.0))])
SPHERE = [Obstacle3D(kind="sphere", pos=(0.5, 0.5, 0.5), size=(0.15, 0.15, 0.15))]


def _run_plume(N, steps, dtype, storage):
    s = Solver3D(N, _plume_config(storage), dtype=dtype)
    s.set_option(_lib.OPT_TRUE_RESIDUAL, 1)
    st = s.reset(SPHERE)
    worst, iters, unconv = 0.0, [], 0
    for _ in range(steps):
        s.step(st, PLUME, DT)
        ps = s.pressure_stats()
        worst = max(worst, ps["true_rel_res"])
        iters.append(ps["iters"])
        unconv += 0 if ps["converged"] else 1
    phi = torch.empty(N, N, N, device=st.data.device, dtype=dtype)
    s._chk(s._api.vfx_get_phi(s._ctx, ctypes.c_void_p(phi.data_ptr()), s._stream()))
    div = s.divergence(st)
    out = {"density": st.data[0].double().cpu().numpy(), "pressure": phi.double().cpu().numpy(),
           "divergence": div.double().cpu().numpy(), "worst": worst, "iters": iters, "unconverged": unconv,
           "max_div": float(div.max())}
    del s, st, phi, div
    torch.cuda.empty_cache()
    return out


def test_mixed_precision_256_against_fp64(cuda, hiplib):
    """256^3 plume past a sphere, 50 steps: FP32 fields with FP32 / FP16 / BF16 levels against the
    FP64 run.  Bounds: every projection at true rel-res <= 1e-5; normalised max |div u| <= 1e-4
    (the reference's gate, tests3d/test_sim3d_layer1.py:135-145); relative L2 errors against FP64
    of pressure <= 2e-2 and density <= 2e-3."""
    N, steps = 256, 50
    _say(f"mixed precision: {N}^3 plume past a sphere, {steps} steps")
    ref = _run_plume(N, steps, torch.float64, "fp32")
    _say(f"  fp64: worst true rel-res {ref['worst']:.2e}, iterations mean {np.mean(ref['iters']):.2f} "
         f"max {max(ref['iters'])}, max |div| {ref['max_div']:.2e}")
    assert ref["worst"] <= 1e-5 and ref["unconverged"] == 0
    for storage in ("fp32", "fp16", "bf16"):
        r = _run_plume(N, steps, torch.float32, storage)
        errs = {}
        for f in ("pressure", "divergence", "density"):
            a, b = r[f], ref[f]
            errs[f] = (float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)),
                       float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)))
        _say(f"  fp32 fields, {storage} levels: worst true rel-res {r['worst']:.2e}, iterations mean "
             f"{np.mean(r['iters']):.2f} max {max(r['iters'])}, max |div| {r['max_div']:.2e}; "
             + ", ".join(f"{f} max {e[0]:.2e} L2 {e[1]:.2e}" for f, e in errs.items()))
        assert r["worst"] <= 1e-5 and r["unconverged"] == 0, storage
        assert r["max_div"] <= 1e-4, storage
        assert errs["pressure"][1] <= 2e-2, storage
        assert errs["density"][1] <= 2e-3, storage


def _pre_projection(N, seed):
    """The seeded scene after 6 steps, advected and forced for the next step: the state a
    projection starts from (cold start, true relative residual 1e-5, up to 2 restarts)."""
    from vfx.sim3d.learned import random_scene, scene_config
    from vfx.sim3d.solver import pack_params
    forces, obst, beta = random_scene(seed)
    cfg = dataclasses.replace(scene_config(beta, seed), pressure_warm_start=False, pressure_rel_tol=1e-5,
                              pressure_max_iters=400, pressure_restarts=2)
    s = Solver3D(N, cfg)
    st = s.reset(obst)
    for _ in range(6):
        s.step(st, forces, DT)
    h = pack_params(cfg, N, forces, DT, st.meta.step)
    for stage in (_lib.STAGE_ADVECT, _lib.STAGE_FORCES):
        s._chk(s._api.vfx_step_stage(s._ctx, stage, ctypes.c_void_p(st.data.data_ptr()),
                                      ctypes.c_void_p(st.solid.data_ptr()),
                                      h.ctypes.data_as(ctypes.c_void_p) if stage == 0 else None, s._stream()))
    torch.cuda.synchronize()
    return s, st


def test_learned_preconditioner_128_held_out(cuda, hiplib):
    """The coarse correction of the V-cycle (a U-Net on the 32^3 grid) trained at 128^3 on what the
    V-cycle leaves of solver pairs (seeded scenes 20-27) and noise pairs, applied in flexible CG on
    the held-out scenes 99 and 98: every projection meets the true relative residual 1e-5 without
    the MGPCG fallback and in no more flexible-CG iterations than plain MGPCG takes on the same
    pre-projection state; times are reported."""
    from vfx.sim3d.learned import UNet, generate_pairs, noise_pairs, train_coarse, vcycle_pairs
    N = 128
    pairs = generate_pairs(N, seeds=range(20, 28), steps=14, every=1)
    pairs += noise_pairs(N, [p[0] for p in pairs[::11]], count=len(pairs) // 2, seed=5)
    data = vcycle_pairs(N, pairs)
    net = UNet(N // 4, seed=1)
    losses = train_coarse(net, data, iters=600, lr=2e-3, seed=0)
    first, last = float(np.mean(losses[:50])), float(np.mean(losses[-50:]))
    _say(f"learned 128^3: coarse correction trained on {len(data)} pairs, loss {first:.3f} -> {last:.3f}")
    assert last < first
    del pairs, data
    for seed in (99, 98):
        s, st = _pre_projection(N, seed)
        s.set_option(_lib.OPT_LEARNED_SWEEPS, 0)
        row = {}
        for name in ("mgpcg", "learned"):
            s.set_option(_lib.OPT_TRUE_RESIDUAL, 1)
            s.set_preconditioner(net if name == "learned" else None)
            times, ps = [], None
            for _ in range(3):
                work = FieldState3D(data=st.data.clone(), solid=st.solid, meta=st.meta)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                s.project(work)
                ps = s.pressure_stats()
                times.append(1000.0 * (time.perf_counter() - t0))
            row[name] = (float(np.median(times)), ps)
            _say(f"  scene {seed} {name}: {row[name][0]:.2f} ms, MGPCG iterations {ps['iters']}, flexible-CG "
                 f"iterations {ps['learned_iters']}, fallback {ps['fallback']}, true rel-res {ps['true_rel_res']:.2e}")
            assert 0 <= ps["true_rel_res"] <= 1e-5 and ps["converged"], (seed, name)
        s.set_preconditioner(None)
        lps, mps = row["learned"][1], row["mgpcg"][1]
        _say(f"  scene {seed}: learned / MGPCG time {row['learned'][0] / row['mgpcg'][0]:.2f}")
        assert lps["fallback"] == 0, seed
        assert lps["learned_iters"] <= mps["iters"], (seed, lps["learned_iters"], mps["iters"])
        del s, st
        torch.cuda.empty_cache()


BATCH_SIZES = (48, 64, 96, 128)


def test_batch_32_uneven_sims(cuda, hiplib):
    """32 sims of sizes drawn by seed 0 from {48, 64, 96, 128}^3 on the seeded scenes, as graph
    frames: every projection of every frame converges to relative residual 1e-5 within its budget,
    and one sim of each size equals its own Solver3D.step bit for bit after every frame."""
    from vfx.sim3d.batch import BatchSolver3D
    from vfx.sim3d.learned import random_scene, scene_config
    rng = np.random.default_rng(0)
    sizes = [int(x) for x in rng.choice(BATCH_SIZES, 32)]
    scenes = [random_scene(i) for i in range(32)]
    cfgs = [dataclasses.replace(scene_config(beta, i), pressure_restarts=0) for i, (_, _, beta) in enumerate(scenes)]
    forces = [f for f, _, _ in scenes]
    _say(f"batch: 32 sims, sizes {sorted(sizes)}")
    batch = BatchSolver3D(sizes, cfgs, streams=4, graph=True)
    states = batch.reset([o for _, o, _ in scenes])
    picks = [sizes.index(n) for n in BATCH_SIZES if n in sizes]
    singles = {i: Solver3D(sizes[i], cfgs[i]) for i in picks}
    ref = {i: singles[i].reset(scenes[i][1]) for i in picks}
    for fr in range(6):
        batch.step(states, forces, DT)
        info = batch.wait()
        for i in picks:
            singles[i].step(ref[i], forces[i], DT)
        torch.cuda.synchronize()
        assert np.all(info["converged"]), (fr, np.nonzero(~info["converged"])[0], info["iters"])
        for i in picks:
            assert torch.equal(states[i].data, ref[i].data), (fr, i, sizes[i])
        _say(f"  frame {fr}: iterations {int(info['iters'].min())}-{int(info['iters'].max())}, "
             f"frame {info['frame_ms']:.2f} ms, latency p50 {np.percentile(info['latency_ms'], 50):.2f} ms")
    for st in states:
        assert torch.isfinite(st.data).all()


def test_batch_32_uneven_sims_learned(cuda, hiplib):
    """Config 4 with the learned preconditioner: the V-cycle's coarse correction trained at 64^3
    runs in every one of the 32 sims of sizes {48, 64, 96, 128}^3 (one network instance per sim on
    its quarter-resolution grid), sequenced on the device inside graph frames with the MGPCG
    fallback; every projection converges, and one sim of each size equals its own Solver3D.step
    with the same network bit for bit after every frame."""
    from vfx.sim3d.batch import BatchSolver3D
    from vfx.sim3d.learned import (UNet, generate_pairs, noise_pairs, random_scene, scene_config, train_coarse,
                                   vcycle_pairs)
    pairs = generate_pairs(64, seeds=range(20, 24), steps=10, every=1)
    pairs += noise_pairs(64, [p[0] for p in pairs[::11]], count=len(pairs) // 2, seed=5)
    net = UNet(16, seed=1)
    train_coarse(net, vcycle_pairs(64, pairs), iters=200, lr=2e-3, seed=0)
    params = net.params()
    del pairs, net
    rng = np.random.default_rng(0)
    sizes = [int(x) for x in rng.choice(BATCH_SIZES, 32)]
    scenes = [random_scene(i) for i in range(32)]
    cfgs = [dataclasses.replace(scene_config(beta, i), pressure_restarts=1) for i, (_, _, beta) in enumerate(scenes)]
    forces = [f for f, _, _ in scenes]
    _say(f"learned batch: 32 sims, sizes {sorted(sizes)}")
    batch = BatchSolver3D(sizes, cfgs, streams=4, graph=True)
    batch.set_preconditioner(params)
    states = batch.reset([o for _, o, _ in scenes])
    picks = [sizes.index(n) for n in BATCH_SIZES if n in sizes]
    singles = {i: Solver3D(sizes[i], cfgs[i]) for i in picks}
    for i in picks:
        singles[i].set_preconditioner(UNet(sizes[i] // 4, params=params))
    ref = {i: singles[i].reset(scenes[i][1]) for i in picks}
    for fr in range(6):
        batch.step(states, forces, DT)
        info = batch.wait()
        for i in picks:
            singles[i].step(ref[i], forces[i], DT)
        torch.cuda.synchronize()
        assert np.all(info["converged"]), (fr, np.nonzero(~info["converged"])[0], info["iters"])
        for i in picks:
            ps = singles[i].pressure_stats()
            assert ps["converged"] and ps["learned_iters"] > 0, (fr, i, ps)
            assert torch.equal(states[i].data, ref[i].data), (fr, i, sizes[i])
        _say(f"  frame {fr}: iterations {int(info['iters'].min())}-{int(info['iters'].max())}, "
             f"frame {info['frame_ms']:.2f} ms, latency p50 {np.percentile(info['latency_ms'], 50):.2f} ms")
    for st in states:
        assert torch.isfinite(st.data).all()

"""Benchmark of the sim3d step on one MI355X: 256^3 fire, plain MGPCG projection to relative
residual 1e-5 (true residual judged), FP32 fields.

    python bench.py [--steps K] [--warmup W] [--res N] [--no-cpu]

Prints ONE JSON line: the metric, the step time, the stage split and iterations of a few extra
profiled steps, the live roofline of the dominant kernel (HIP events on the solver's stream) and
the CPU baseline (the oracle, timed on the host cores on one step of the same workload).
"""
from __future__ import annotations

import argparse
import ctypes
import dataclasses
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "neural-vfx_amd"))

HBM_PEAK_GBS = 8000.0
COPY_PROBE_BYTES = 256 << 20   # device copy of known size that scales the PMC byte counters

# Algorithmic (compulsory) bytes per grid cell of each kernel, per launch: every distinct input
# element read once, every output written once, at its storage type (FP32 fields, uint8 mask,
# multigrid levels of s bytes per value; the level-1 multigrid has N^3 unknowns, 8 per level-1
# cell, so its figures are per fine cell too).  Standard multigrid cells take their coefficients
# from a 32-float table, so the level kernels read a 1-byte flag per level-1 cell (1/8 byte per
# unknown) instead of diag/off.  The CG kernels are the steady-state forms (iteration k >= 2).
# The table is written for s = 4 (FP32 levels); kernel_bytes_per_cell rescales the level terms.
KERNEL_BYTES_PER_CELL = {
    "adv_sl": 24 + 24 + 4,              # data (6 ch, velocity 3 of them) -> hat (6 ch), trace corner
    "adv_mc": 24 + 24 + 4 + 1 + 24,     # data + hat + trace corner + mask -> adv
    "sources": 24 + 1 + 24,             # adv + mask -> data (density, velocity), hat (T, fuel)
    "diffuse": 8 + 1 + 8,               # hat (T, fuel) + mask -> data (T, fuel)
    "upd_matvec": 4 * 3 + 1 + 4 * 3,    # z, p_old, phi + mask -> p, phi, q
    "inv_post": 4 + 4 + 1 + 4,          # level-1 x, r + mask -> z
    "pre_resid_fwd": 4 + 4 + 1 + 4 + 4 + 2 + 2,   # r, q + mask -> r, level-1 r; red half: diag -> x
    "smooth0": 2 + 4 + 0.125 + 2,       # one colour: r (half), x (all) + flags -> x (half)
    "restrict0": 4 + 4 + 0.125 + 0.5 + 0.25 + 0.25,   # level-1 r, x + flags -> level-2 r; red: diag -> x
    "prolong0": 2 + 4 + 0.5 + 0.125 + 2,   # black: r, x (all), level-2 x + flags -> x (half)
}

# level-array bytes (at s = 4) in each kernel's figure above
LEVEL_BYTES_PER_CELL = {"inv_post": 4, "pre_resid_fwd": 4 + 2, "smooth0": 2 + 4 + 2, "restrict0": 4 + 4 + 0.5 + 0.25,
                        "prolong0": 2 + 4 + 0.5 + 2}


def kernel_bytes_per_cell(name: str, s: int = 4) -> float:
    """KERNEL_BYTES_PER_CELL with the multigrid level values stored in s bytes."""
    return KERNEL_BYTES_PER_CELL[name] - LEVEL_BYTES_PER_CELL.get(name, 0.0) * (1.0 - s / 4.0)


# Algorithmic bytes per launch of the finer scalar grid's kernels (S^3 fine cells, N^3 coarse
# cells, FP32, 3 fine channels): every distinct element read once and written once, the coarse
# velocity (12 bytes per coarse cell) read by both advection passes, fine solid (4 bytes) read by
# the first pass which writes the 1-byte fluid flag the later passes read.
def fine_kernel_bytes(name: str, N: int, S: int) -> float:
    f, c = float(S) ** 3, float(N) ** 3
    return {"fadv_sl": f * (12 + 4 + 12 + 4 + 1) + 12 * c,      # hi, solid -> hat, trace corner, flag
            "fadv_mc": f * (12 + 12 + 4 + 1 + 12) + 12 * c,     # hi, hat, corner, flag -> adv
            "fsources": f * (12 + 1 + 12),                      # adv, flag -> hat / hi
            "fdiffuse": f * (8 + 1 + 8),                        # one substep: T, fuel + flag -> T, fuel
            "frestrict": f * 12 + c * (1 + 12)}.get(name, 0.0)  # hi, coarse mask -> 3 coarse channels


# launches per step of each timed kernel for n CG iterations (the per-step share picks the
# dominant kernel)
def launches_per_step(name: str, n: int) -> float:
    return {"adv_sl": 1, "adv_mc": 1, "sources": 1, "diffuse": 1, "upd_matvec": max(n - 1, 0), "inv_post": n,
            "pre_resid_fwd": n, "smooth0": 2 * n, "restrict0": n, "prolong0": n, "coarse": n,
            "fadv_sl": 1, "fadv_mc": 1, "fsources": 1, "fdiffuse": 1, "frestrict": 1}.get(name, 0)


def step_bytes_per_cell(n: int, s: int = 4) -> float:
    """Implementation-independent algorithmic bytes per cell of one whole step with n CG iterations
    and multigrid level storage of s bytes per value: advection 49 (read 24 + mask 1, write 24),
    sources + diffusion 49, the vorticity / force chain 37, the projection's fixed cost 58 (rhs 21,
    apply 29, p init 8), one preconditioner application B_prec = (5 + s) + 15.1 s + (9 + s) =
    14 + 17.1 s, and per iteration 9 (matvec + dot) + 32 (r, phi, p updates) + B_prec:
    193 + B_prec + n (41 + B_prec)."""
    b_prec = 14 + 17.1 * s
    return 193 + b_prec + n * (41 + b_prec)


def pmc_traffic(kernel: str, res: int, scalar_res=None):
    """HBM bytes per launch of `kernel` from the committed PMC summary of the same grid size
    (profiles/pmc_traffic_<res>_fire[_scalars<S>].md: FETCH_SIZE + WRITE_SIZE, scaled by a copy of
    known size), or None when there is none."""
    tag = "" if scalar_res is None else f"_scalars{scalar_res}"
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_{res}_fire{tag}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cols = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cols) == 9 and cols[0] == kernel:
            return int(cols[6])
    return None


def fire_config(seed: int):
    """Fire scene settings: buoyancy beta 1.5 per K of excess temperature, confinement 0.3,
    combustion at 2.0 /s above 0.3 K excess, pressure to ||r||/||b|| <= 1e-5."""
    from vfx.sim3d import SolverConfig3D
    return SolverConfig3D(buoyancy_alpha=0.0, buoyancy_beta=1.5, vorticity_eps=0.3, burn_rate=2.0,
                          ignition_temp=293.3, ignition_ramp=0.1, burn_heat=2.0, smoke_from_fuel=0.35,
                          temperature_cooling=0.5, density_dissipation=0.05,
                          pressure_rel_tol=1e-5, pressure_max_iters=200, pressure_warm_start=True,
                          pressure_restarts=2, emitter_noise=0.2, noise_seed=seed)


def fire_forces():
    """Fuel disc of radius 0.1 m on the floor: a sphere emitter centred on the floor (its
    footprint on the floor is the disc), injecting fuel and a 1 K-excess ignition temperature."""
    from vfx.core.fields3d import Emitter3D, Forces3D
    return Forces3D(emitters=[Emitter3D(pos=(0.5, 0.0, 0.5), radius=0.1, density_rate=1.0, temperature=294.0,
                                        fuel_rate=1.0, velocity=(0.0, 0.0, 0.0))])


def plume_config(mg: str = "fp32"):
    """Smoke plume past a sphere: buoyancy beta 1 per K, confinement 0.3, emitter noise
    U[0.8, 1.2], pressure to ||r||/||b|| <= 1e-5 (true residual judged, up to 2 restarts)."""
    from vfx.sim3d import SolverConfig3D
    return SolverConfig3D(buoyancy_alpha=0.0, buoyancy_beta=1.0, vorticity_eps=0.3, pressure_rel_tol=1e-5,
                          pressure_max_iters=200, pressure_restarts=2, emitter_noise=0.2, noise_seed=0,
                          mg_storage=mg)


def plume_scene():
    """Emitter sphere r 0.08 m at (0.5, 0.1, 0.5) injecting density 1 per step and a 1 K excess
    temperature target (both x U[0.8, 1.2] per cell); solid sphere r 0.15 m at the centre."""
    from vfx.core.fields3d import Emitter3D, Forces3D, Obstacle3D
    forces = Forces3D(emitters=[Emitter3D(pos=(0.5, 0.1, 0.5), radius=0.08, density_rate=60.0, temperature=294.0,
                                          fuel_rate=0.0, velocity=(0.0, 0.0, 0.0))])
    return forces, [Obstacle3D(kind="sphere", pos=(0.5, 0.5, 0.5), size=(0.15, 0.15, 0.15))]


def precision_report(N: int, steps: int):
    """Pressure, divergence and density of the FP32 solver (multigrid levels in FP32, FP16, BF16)
    against the FP64 run of the same solver on the same plume, after `steps` steps."""
    import numpy as np
    import torch
    from vfx.sim3d import Solver3D, _lib
    forces, obst = plume_scene()
    dt = 1.0 / 60.0
    res = {}
    for name, dtype, mg in (("fp64", torch.float64, "fp32"), ("fp32", torch.float32, "fp32"),
                            ("fp16_levels", torch.float32, "fp16"), ("bf16_levels", torch.float32, "bf16")):
        print(f"# {name}: {steps} steps at {N}^3", file=sys.stderr, flush=True)
        s = Solver3D(N, plume_config(mg), dtype=dtype)
        s.set_option(_lib.OPT_TRUE_RESIDUAL, 1)
        st = s.reset(obst)
        iters, worst = [], 0.0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            s.step(st, forces, dt)
            ps = s.pressure_stats()
            iters.append(ps["iters"])
            worst = max(worst, ps["true_rel_res"])
        torch.cuda.synchronize()
        ms = 1000.0 * (time.perf_counter() - t0) / steps
        phi = torch.empty(N, N, N, device=st.data.device, dtype=dtype)
        s._chk(s._api.vfx_get_phi(s._ctx, ctypes.c_void_p(phi.data_ptr()), s._stream()))
        div = s.divergence(st)
        torch.cuda.synchronize()
        res[name] = {"density": st.data[0].double().cpu().numpy(), "pressure": phi.double().cpu().numpy(),
                     "divergence": div.double().cpu().numpy(), "iters_mean": round(float(np.mean(iters)), 2),
                     "iters_max": int(max(iters)), "worst_true_rel_res": float(f"{worst:.3g}"),
                     "max_div": float(f"{float(div.max()):.3g}"), "ms_per_step_with_sync": round(ms, 2)}
        del s, st
        torch.cuda.empty_cache()
    ref = res["fp64"]
    out = {"report": f"{N}^3 smoke plume past a sphere, {steps} steps: FP32 solver vs FP64 run of the same solver",
           "res": N, "steps": steps, "fp64": {k: v for k, v in ref.items() if not isinstance(v, np.ndarray)}}
    for name in ("fp32", "fp16_levels", "bf16_levels"):
        r = res[name]
        row = {k: v for k, v in r.items() if not isinstance(v, np.ndarray)}
        for f in ("pressure", "divergence", "density"):
            a, b = r[f], ref[f]
            scale = max(float(np.abs(b).max()), 1e-300)
            row[f"{f}_max_err"] = float(f"{float(np.abs(a - b).max()) / scale:.3g}")
            row[f"{f}_l2_err"] = float(f"{float(np.linalg.norm(a - b)) / max(float(np.linalg.norm(b)), 1e-300):.3g}")
        out[name] = row
    print(json.dumps(out))


def net_flops(N: int):
    """(algorithmic, executed) FLOPs of one U-Net application at N^3: 2 * 27 * cin * cout per output
    voxel of each convolution, with the nominal channel counts and with the MFMA-padded ones."""
    from vfx.sim3d.learned import LAYERS
    level = (0, 0, 1, 1, 2, 2, 1, 1, 0, 0, 0)
    pad = lambda c: max(16, (c + 15) // 16 * 16)
    alg = exe = 0.0
    for (ci, co, _), lv in zip(LAYERS, level):
        vox = (N >> lv) ** 3
        alg += 2.0 * 27 * ci * co * vox
        exe += 2.0 * 27 * pad(ci) * pad(co) * vox
    return alg, exe


def train_learned(design: str, train_res: int, iters: int, say=None):
    """Train the learned part on solver pairs of seeded scenes 20-27 plus multi-scale noise pairs at
    train_res: design "coarse" (the default) is the coarse correction of the V-cycle (a U-Net on the
    (train_res/4)^3 grid, trained on what the V-cycle leaves, in the fine energy norm), "full" the
    full-resolution U-Net preconditioner.  Returns (params, training summary)."""
    import numpy as np
    import torch
    from vfx.sim3d.learned import UNet, generate_pairs, noise_pairs, train, train_coarse, vcycle_pairs
    t0 = time.perf_counter()
    pairs = generate_pairs(train_res, seeds=range(20, 28), steps=14, every=1)
    pairs += noise_pairs(train_res, [p[0] for p in pairs[::11]], count=len(pairs) // (2 if design == "coarse" else 1),
                         seed=5)
    if design == "coarse":
        data = vcycle_pairs(train_res, pairs)
        net = UNet(train_res // 4, seed=1)
        losses = train_coarse(net, data, iters=iters, lr=2e-3, seed=0, log=say)
    else:
        data = pairs
        net = UNet(train_res, seed=1)
        losses = train(net, data, iters=iters, lr=2e-3, seed=0, log=say)
    torch.cuda.synchronize()
    summary = {"design": design, "res": train_res, "pairs": len(data), "iters": iters,
               "seconds": round(time.perf_counter() - t0, 1),
               "loss_first50": round(float(np.mean(losses[:50])), 4),
               "loss_last50": round(float(np.mean(losses[-50:])), 4)}
    params = net.params()
    del net, pairs, data
    torch.cuda.empty_cache()
    return params, summary


def compare_learned(N: int, params, design: str, reps: int = 3):
    """Learned flexible CG (MGPCG fallback) against plain MGPCG on the same pre-projection state of
    the held-out scene 99 after 6 steps at N^3: cold start, ||r||/||b|| <= 1e-5 with the true
    residual judged; median wall ms of `reps` projections each, iterations, fallback."""
    import numpy as np
    import torch
    from vfx.core.fields3d import FieldState3D
    from vfx.sim3d import Solver3D, _lib
    from vfx.sim3d.learned import UNet, random_scene, scene_config
    from vfx.sim3d.solver import pack_params
    forces, obst, beta = random_scene(99)
    cfg = dataclasses.replace(scene_config(beta, 99), pressure_warm_start=False, pressure_rel_tol=1e-5,
                              pressure_max_iters=400, pressure_restarts=2)
    s = Solver3D(N, cfg)
    st = s.reset(obst)
    for _ in range(6):
        s.step(st, forces, 1.0 / 60.0)
    h = pack_params(cfg, N, forces, 1.0 / 60.0, st.meta.step)
    for stage in (_lib.STAGE_ADVECT, _lib.STAGE_FORCES):
        s._chk(s._api.vfx_step_stage(s._ctx, stage, ctypes.c_void_p(st.data.data_ptr()),
                                      ctypes.c_void_p(st.solid.data_ptr()),
                                      h.ctypes.data_as(ctypes.c_void_p) if stage == 0 else None, s._stream()))
    torch.cuda.synchronize()
    netN = UNet(N // 4 if design == "coarse" else N, params=params)
    s.set_option(_lib.OPT_LEARNED_SWEEPS, 0 if design == "coarse" else 2)
    row = {}
    for name in ("mgpcg", "learned"):
        s.set_option(_lib.OPT_TRUE_RESIDUAL, 1)
        s.set_preconditioner(netN if name == "learned" else None)
        times = []
        for _ in range(reps):
            work = FieldState3D(data=st.data.clone(), solid=st.solid, meta=st.meta)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            s.project(work)
            ps = s.pressure_stats()   # synchronizes
            times.append(1000.0 * (time.perf_counter() - t1))
        row[name] = {"pressure_ms": round(float(np.median(times)), 3), "mgpcg_iters": ps["iters"],
                     "learned_iters": ps["learned_iters"], "fallback": ps["fallback"],
                     "true_rel_res": float(f"{ps['true_rel_res']:.3g}")}
        if name == "learned":
            row[name]["history_first"] = [float(f"{x / max(ps['b_norm'], 1e-300):.2e}") for x in s.fcg_history()[:10]]
    s.set_preconditioner(None)
    # the network's application alone (on its own grid)
    Nn = netN.res
    ss = Solver3D(Nn)
    sol = torch.zeros(Nn, Nn, Nn, device="cuda")
    r = torch.randn(Nn, Nn, Nn, device="cuda")
    for _ in range(2):
        netN.apply(ss, sol, r)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(10):
        netN.apply(ss, sol, r)
    ev[1].record()
    ev[1].synchronize()
    ms = ev[0].elapsed_time(ev[1]) / 10
    alg, exe = net_flops(Nn)
    row["net_grid"] = Nn
    row["net_apply_ms"] = round(ms, 3)
    row["net_mfma_tflops_per_s"] = round(exe / ms / 1e9, 1)
    row["net_mfma_frac_of_bf16_dense_peak"] = round(exe / ms / 1e9 / 2500.0, 4)
    row["speedup_learned_vs_mgpcg"] = round(row["mgpcg"]["pressure_ms"] / row["learned"]["pressure_ms"], 3)
    del s, st, netN, ss, r
    torch.cuda.empty_cache()
    return row


def learned_report(train_res: int, eval_res, iters: int, design: str = "coarse"):
    """Learned flexible CG against plain MGPCG at each eval_res (compare_learned), after training
    (train_learned)."""
    import torch
    say = lambda *a: print("#", *a, file=sys.stderr, flush=True)
    params, summary = train_learned(design, train_res, iters, say)
    say(f"trained: {summary}")
    out = {"report": "learned flexible CG (" + ("MG V-cycle + coarse U-Net correction" if design == "coarse" else
                                                "full-resolution U-Net + 2 Jacobi corrections")
                     + ", MGPCG fallback) vs plain MGPCG, held-out scene 99, cold start, rel-res 1e-5",
           "train": summary}
    for N in eval_res:
        row = compare_learned(N, params, design)
        say(f"{N}^3: {row}")
        out[str(N)] = row
    torch.cuda.empty_cache()
    print(json.dumps(out))


BATCH_SIZES = (48, 64, 96, 128)


def pair_report(steps: int, warmup: int, lr_res: int = 64, hr_res: int = 256):
    """Paired runs (vfx/datagen3d/sim.py run_pair): the free LR runtime run and the HR run guided to
    follow it, on the reference's plume-past-a-sphere scene with sandbox dissipation.  Each phase of
    a paired step is bracketed by device events on the stream: the guidance velocity nudge, the HR
    step, the LR step, the scalar nudge, the tracking error.  Also the guided / unguided tracking."""
    import numpy as np
    import torch
    from vfx.core.fields3d import SIM_DT, Emitter3D, Forces3D, Obstacle3D
    from vfx.datagen3d import DEFAULT_GUIDE, Guide3D, tracking_error
    from vfx.datagen3d import sim as pair

    def plume(k):
        return Forces3D(emitters=[Emitter3D(pos=(0.42, 0.1, 0.5), radius=0.06, density_rate=2.0, temperature=900.0,
                                            velocity=(0.0, 1.2, 0.0))], wind=(0.4, 0.0, 0.0))
    sc = pair.PairScenario(plume, num_steps=warmup + steps, sim=dict(density_dissipation=1.2, temperature_cooling=5.0),
                           obstacles=[Obstacle3D("sphere", (0.5, 0.55, 0.5), (0.12, 0.12, 0.12))])
    run = pair.PairedRun.get(lr_res, hr_res)
    ls, hs = run.lr.solver, run.hr.solver
    lr = run.lr.start(sc.sim, sc.obstacles)
    hr = run.hr.start(sc.sim, sc.obstacles)
    g = Guide3D(DEFAULT_GUIDE, hr.solid, lr_res)
    names = ("guide_velocity", "hr_step", "lr_step", "guide_scalars", "tracking")
    ts = torch.cuda.current_stream()
    rows, tracks = [], []
    t_wall = 0.0
    for k in range(warmup + steps):
        f = sc.forces(k)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev[0].record(ts)
        g.velocity(hr, lr)
        ev[1].record(ts)
        hs.step(hr, f, SIM_DT)
        ev[2].record(ts)
        ls.step(lr, f, SIM_DT)
        ev[3].record(ts)
        g.scalars(hr, lr)
        ev[4].record(ts)
        tr = tracking_error(hr.data, lr.data)
        ev[5].record(ts)
        ev[5].synchronize()
        if k >= warmup:
            t_wall += time.perf_counter() - t0
            rows.append([ev[i].elapsed_time(ev[i + 1]) for i in range(5)])
            tracks.append(tr.cpu().numpy())
    med = {n: round(statistics.median(r[i] for r in rows), 4) for i, n in enumerate(names)}
    ms = 1000.0 * t_wall / steps
    # the guidance kernels' compulsory HBM bytes per call at the HR size: velocity reads the HR
    # velocity for the pool and reads / writes it for the update (3 * (4 + 8) bytes per cell) plus the
    # fluid flag; scalars the same for density and temperature (2 * 12) plus the flag
    h3 = hr_res ** 3
    vb, sb = 37.0 * h3, 25.0 * h3
    vel_gbs = vb / (med["guide_velocity"] * 1e-3) / 1e9
    sc_gbs = sb / (med["guide_scalars"] * 1e-3) / 1e9
    tr = np.array(tracks)
    unguided = []
    pair.free_cache()
    for s_, _, _, t in pair.run_pair(sc, None, hr_res=hr_res, lr_res=lr_res, steps=min(warmup + steps, 60)):
        if s_ >= 10:
            unguided.append(t.cpu().numpy())
    ung = np.array(unguided)
    out = {"metric": f"paired run step ms ({lr_res}^3 runtime + {hr_res}^3 guided)", "value": round(ms, 3),
           "unit": "ms", "higher_is_better": False, "steps": steps, "warmup": warmup,
           "phases_ms": med, "guide_ms": round(med["guide_velocity"] + med["guide_scalars"], 4),
           "guide_velocity_gbs": round(vel_gbs, 1), "guide_scalars_gbs": round(sc_gbs, 1),
           "guide_bytes": {"velocity": int(vb), "scalars": int(sb)},
           "tracking_guided_worst": [round(float(v), 4) for v in tr.max(axis=0)],
           "tracking_guided_mean": [round(float(v), 4) for v in tr.mean(axis=0)],
           "tracking_unguided_worst": [round(float(v), 4) for v in ung.max(axis=0)],
           "config": {"workload": "plume past a sphere (emitter r 0.06 at (0.42, 0.1, 0.5), 900 K, 1.2 m/s up; "
                                  "wind 0.4; density dissipation 1.2, cooling 5), DEFAULT_GUIDE"}}
    print(json.dumps(out))


def render_report(reps: int):
    """The volume renderer (vfx/render3d VolumeRenderer) at 2560 x 1440 on the solver's own 256^3
    fire state (and 64^3, 128^3 states of the same scene), as the reference's feasibility table
    (render3d README: full draw, view-only redraw with reuse_volume, prep / view split), and the
    full draw with return_aux (alpha, depth, motion from the state's velocity and a moved camera),
    and the view-only redraw with the temporal resolve on.
    Times are device events around each draw, median of `reps` draws after 3 warm-up draws."""
    import torch
    from vfx.core.fields3d import Camera, Volume
    from vfx.render3d import RenderParams3D, VolumeRenderer
    from vfx.sim3d import Solver3D

    rows = []
    cam = Camera(position=(0.5, 0.55, -1.4), target=(0.5, 0.45, 0.5))
    prev = Camera(position=(0.49, 0.55, -1.4), target=(0.5, 0.45, 0.5))
    for N in (64, 128, 256):
        s = Solver3D(N, fire_config(0))
        st = s.reset()
        f = fire_forces()
        for _ in range(60):
            s.step(st, f, 1.0 / 60.0)
        vol = Volume(density=st.ch("density"), temperature=st.ch("temperature"))
        for scale, steps in ((0.5, 192.0), (1.0, 64.0), (1.0, 192.0)):
            r = VolumeRenderer((1440, 2560), device="cuda:0", march_scale=scale,
                               params=RenderParams3D(steps=steps), profile=True)
            full, prep, view, reuse, aux = [], [], [], [], []
            for k in range(3 + reps):
                r.draw(vol, cam)
                if k >= 3:
                    t = r.last_timings()
                    full.append(t["total"])
                    prep.append(t["prep"])
                    view.append(t["view"])
                r.draw(vol, cam, reuse_volume=True)
                if k >= 3:
                    reuse.append(r.last_timings()["total"])
                r.draw(vol, cam, return_aux=True, velocity=st.velocity, prev_camera=prev)
                if k >= 3:
                    aux.append(r.last_timings()["total"])
            temporal = []
            r.set_temporal(True)
            for k in range(3 + reps):
                r.draw(vol, cam, reuse_volume=k > 0)
                if k >= 3:
                    temporal.append(r.last_timings()["total"])
            r.set_temporal(False)
            med = lambda v: round(statistics.median(v), 3)
            rows.append({"volume": N, "march_scale": scale, "steps_per_m": steps, "march_px": f"{r.mw}x{r.mh}",
                         "full_ms": med(full), "prep_ms": med(prep), "view_ms": med(view), "reuse_ms": med(reuse),
                         "aux_full_ms": med(aux), "temporal_reuse_ms": med(temporal)})
            del r
        del s, st
        torch.cuda.empty_cache()
    print(json.dumps({"metric": "volume render ms at 2560x1440 (full draw / view-only redraw)", "rows": rows,
                      "config": {"workload": "the solver's fire state after 60 steps, default RenderParams3D "
                                             "(light_res 80), camera (0.5, 0.55, -1.4) -> (0.5, 0.45, 0.5)"}}))


def batch_report(frames: int, warmup: int, seed: int, nsims: int = 32, tune: bool = True):
    """Batched frames (run_batch) as one JSON line."""
    print(json.dumps(run_batch(frames, warmup, seed, nsims, tune, compare=True)))


def run_batch(frames: int, warmup: int, seed: int, nsims: int = 32, tune: bool = True, compare: bool = True):
    """Batched frames: nsims independent simulations with N drawn by seed from 48/64/96/128, seeded
    scenes (1-3 emitters, 0-2 sphere obstacles, beta U[0.5, 2]), MGPCG to rel-res 1e-5 with the
    adaptive iteration budget.  Aggregate cell-steps/s and per-simulation latency of graph frames
    (after the autotuner); with compare, beside the same frames enqueued directly and the
    simulations stepped one by one with Solver3D.step."""
    import numpy as np
    import torch
    from vfx.sim3d import BatchSolver3D
    from vfx.sim3d.learned import random_scene, scene_config
    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in rng.choice(BATCH_SIZES, nsims)]
    scenes = [random_scene(1000 * seed + i) for i in range(nsims)]
    cfgs = [dataclasses.replace(scene_config(beta, 1000 * seed + i), pressure_restarts=0)
            for i, (_, _, beta) in enumerate(scenes)]
    forces = [f for f, _, _ in scenes]
    dt = 1.0 / 60.0
    cells = float(sum(n ** 3 for n in sizes))
    batch = BatchSolver3D(sizes, cfgs, streams=4, graph=True)
    states = batch.reset([o for _, o, _ in scenes])
    print(f"# batch: {nsims} sims, sizes {sorted(sizes)}, {cells / 1e6:.1f} M cells", file=sys.stderr, flush=True)
    for _ in range(warmup):
        batch.step(states, forces, dt)
        batch.wait()
    rep = None
    if tune:
        t0 = time.perf_counter()
        rep = batch.tune(states, forces, dt, reps=2)
        rep["seconds"] = round(time.perf_counter() - t0, 2)
        print(f"# tuned: {rep}", file=sys.stderr, flush=True)
        for _ in range(3):   # budgets settle, graph captured again
            batch.step(states, forces, dt)
            batch.wait()

    def run(graph: bool):
        lat, fms, its, unconv, misses = [], [], [], 0, []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(frames):
            batch.step(states, forces, dt, graph=graph)
            info = batch.wait()
            lat.extend(info["latency_ms"].tolist())
            fms.append(info["frame_ms"])
            its.extend(info["iters"].tolist())
            unconv += int((~info["converged"]).sum())
            for i in np.nonzero(~info["converged"])[0]:
                misses.append((int(i), sizes[int(i)], int(info["iters"][i]), float(f"{info['rel_res'][i]:.3g}")))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return {"cell_steps_per_s": cells * frames / wall, "wall_ms_per_frame": 1000.0 * wall / frames,
                "gpu_ms_per_frame": float(np.median(fms)), "latency_p50_ms": float(np.percentile(lat, 50)),
                "latency_p99_ms": float(np.percentile(lat, 99)), "iters_mean": float(np.mean(its)),
                "unconverged": unconv, "unconverged_first": misses[:6], "budgets": batch.budgets.tolist()}

    g = run(True)
    print(f"# graph frames: {g}", file=sys.stderr, flush=True)
    r3 = lambda d: {k: (round(v, 3) if isinstance(v, float) else v) for k, v in d.items()}
    if not compare:
        out = {"workload": f"{nsims} sims, sizes by seed {seed} from {{48,64,96,128}}^3, MGPCG to rel-res 1e-5, "
                           f"graph frames", "sizes": sorted(sizes), "cells": int(cells), "frames": frames,
               "tuned": tune, **r3(g)}
        del batch, states
        torch.cuda.empty_cache()
        return out
    e = run(False)
    print(f"# direct frames: {e}", file=sys.stderr, flush=True)
    # one by one with the host-polled Solver3D.step
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lat = []
    for _ in range(frames):
        tf = time.perf_counter()
        for s, st, f in zip(batch.solvers, states, forces):
            s.step(st, f, dt)
            torch.cuda.synchronize()
            lat.append(1000.0 * (time.perf_counter() - tf))
    wall = time.perf_counter() - t0
    seq = {"cell_steps_per_s": cells * frames / wall, "wall_ms_per_frame": 1000.0 * wall / frames,
           "latency_p50_ms": float(np.percentile(lat, 50)), "latency_p99_ms": float(np.percentile(lat, 99))}
    print(f"# one by one: {seq}", file=sys.stderr, flush=True)
    out = {"metric": f"aggregate cell-steps/s, {nsims} sims N in {{48,64,96,128}}^3 (seed {seed}), MGPCG to "
                     f"rel-res 1e-5, graph frames",
           "value": round(g["cell_steps_per_s"], 1), "unit": "cell-steps/s", "n_gpus": 1, "frames": frames,
           "warmup": warmup, "higher_is_better": True, "dtype": "fp32", "data": "synthetic",
           "config": {"workload": f"batched frames: {nsims} sims, sizes {sorted(sizes)}", "cells": int(cells)},
           "graph": r3(g), "direct": r3(e), "one_by_one": r3(seq),
           "tune": rep, "vs_one_by_one": round(g["cell_steps_per_s"] / seq["cell_steps_per_s"], 2)}
    return out


def config0_report(steps: int):
    """The 64^3 smoke plume past a sphere (dt 1/60, `steps` steps, seed 0), timed on the GPU and
    on the host cores in the same job: median ms per step (GPU: HIP events around each step after
    20 warm-up steps), median pressure ms (step_profiled, a second run from the same start), PCG
    iterations to rel-res 1e-5; the C oracle (OpenMP) runs the same steps from the same start."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from vfx.sim3d import Solver3D
    from vfx.sim3d.solver import pack_params
    N, dt = 64, 1.0 / 60.0
    cfg = plume_config()
    forces, obst = plume_scene()
    med = lambda v: round(float(statistics.median(v)), 3)
    s = Solver3D(N, cfg)
    st = s.reset(obst)
    data0, solid = st.data.cpu().numpy().copy(), st.solid.cpu().numpy()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    gpu_ms, gpu_it = [], []
    for k in range(steps):
        ev[0].record()
        s.step(st, forces, dt)
        ev[1].record()
        torch.cuda.synchronize()
        gpu_ms.append(ev[0].elapsed_time(ev[1]))
        gpu_it.append(s.pressure_iters())
    s2 = Solver3D(N, cfg)
    st2 = s2.reset(obst)
    prof = [s2.step_profiled(st2, forces, dt) for _ in range(steps)]
    # the same simulation as a one-simulation batched frame: one captured hipGraph per step with a
    # fixed PCG budget (grown when used up), no host round trip inside the step
    from vfx.sim3d import BatchSolver3D
    b = BatchSolver3D([N], [cfg], streams=1, graph=True)
    bst = b.reset([obst])
    graph_ms, graph_it = [], []
    for k in range(steps):
        b.step(bst, [forces], dt)
        w = b.wait()
        graph_ms.append(w["frame_ms"])
        graph_it.append(int(w["iters"][0]))
    warm = min(20, steps // 2)
    cpu_ms, cpu_it = [], []
    data, phi = data0, np.zeros((N,) * 3, np.float32)
    for k in range(steps):
        prm = pack_params(cfg, N, forces, dt, k)
        t0 = time.perf_counter()
        data, it = oracle.step(data, solid, prm, phi, np.float32(N))
        cpu_ms.append(1000.0 * (time.perf_counter() - t0))
        cpu_it.append(int(it))
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    g, c = med(gpu_ms[warm:]), med(cpu_ms)
    print(json.dumps({"metric": "64^3 plume past a sphere: ms per step, GPU and host cores in the same job",
                      "steps": steps, "gpu_ms_per_step": g, "gpu_pressure_ms": med(p["projection"] for p in prof[warm:]),
                      "gpu_iters_median": med(gpu_it), "gpu_iters_max": max(gpu_it),
                      "gpu_graph_ms_per_step": med(graph_ms[warm:]), "gpu_graph_iters_max": max(graph_it),
                      "cpu_ms_per_step": c, "cpu_iters_median": med(cpu_it), "cpu_iters_max": max(cpu_it),
                      "cpu_cores": cores, "cpu_kind": "port (the C oracle, FP32 storage, FP64 dots, OpenMP)",
                      "speedup": round(c / g, 1),
                      "config": {"workload": "64^3 plume past a sphere: emitter sphere r 0.08 at (0.5, 0.1, 0.5), "
                                             "obstacle sphere r 0.15 at the centre, beta 1, eps 0.3, emitter noise "
                                             "U[0.8, 1.2], MGPCG to rel-res 1e-5, dt 1/60, seed 0"}}))


def seeds_report(steps: int, N: int = 256, mg: str = "fp16"):
    """The 256^3 fire for `steps` steps (dt 1/60) from rest with noise seeds 0-3 (multigrid levels
    stored as `mg`): per seed the median ms per step (HIP events around each step, steps 20+),
    PCG iterations (median, max), the worst true relative residual and restarts, and the max
    normalised |div u| after the last step."""
    import torch
    from vfx.sim3d import Solver3D
    med = lambda v: round(float(statistics.median(v)), 3)
    rows = []
    f = fire_forces()
    for seed in range(4):
        s = Solver3D(N, dataclasses.replace(fire_config(seed), mg_storage=mg))
        st = s.reset()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ms, its, worst, restarts = [], [], 0.0, 0
        for k in range(steps):
            ev[0].record()
            s.step(st, f, 1.0 / 60.0)
            ev[1].record()
            torch.cuda.synchronize()
            ms.append(ev[0].elapsed_time(ev[1]))
            d = s.pressure_stats()
            its.append(d["iters"])
            worst = max(worst, d["true_rel_res"])
            restarts += d["restarts"]
        div = float(s.divergence(st).abs().max().item())
        rows.append({"seed": seed, "ms_per_step": med(ms[20:]), "iters_median": med(its), "iters_max": max(its),
                     "worst_true_rel_res": float(f"{worst:.3g}"), "restarts": restarts, "max_div_last": float(f"{div:.3g}")})
        del s, st
        torch.cuda.empty_cache()
    print(json.dumps({"metric": f"{N}^3 fire, {steps} steps from rest per noise seed", "rows": rows,
                      "config": {"workload": f"{N}^3 fire (fuel disc, combustion, buoyancy, confinement), MGPCG "
                                             "rel-res 1e-5 (true residual judged), dt 1/60, seeds 0-3"}}))


def cpu_baseline(solver, st_data, st_solid, cfg, forces, dt, res, step, steps: int = 5):
    """The oracle's FP32 steps on the host cores: `steps` steps (a bounded sample) continuing the
    GPU run's state, the pressure warm-started from the GPU solver's phi as the GPU's next step
    would be; ms per step is the median."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from vfx.sim3d.solver import pack_params
    data = st_data.cpu().numpy()
    solid = st_solid.cpu().numpy()
    phi_d = torch.empty(res, res, res, device=st_data.device)
    solver._chk(solver._api.vfx_get_phi(solver._ctx, ctypes.c_void_p(phi_d.data_ptr()), solver._stream()))
    phi = np.ascontiguousarray(phi_d.cpu().numpy())
    times, its = [], []
    for k in range(steps):
        prm = pack_params(cfg, res, forces, dt, step + k)
        t0 = time.perf_counter()
        data, it = oracle.step(data, solid, prm, phi, np.float32(res))
        times.append(time.perf_counter() - t0)
        its.append(it)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(1000.0 * statistics.median(times), 1), "unit": "ms/step", "cores": cores, "kind": "port",
            "sample": f"{steps} steps of the {res}^3 fire continuing the GPU run (pressure warm-started from the "
                      f"GPU's phi; PCG iterations {its}), C oracle with OpenMP, median"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--presim", type=int, default=60,
                    help="untimed steps before the warm-up, so the timed steps see a developed fire "
                         "(the reference's bench pre-simulates 60 steps, bench3d/run.py:136-156)")
    ap.add_argument("--res", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--profiled", type=int, default=5, help="extra step_profiled steps for the stage split")
    ap.add_argument("--roofline-kernel", default="", help="kernel for the roofline (default: largest step share)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seeds", type=int, default=0, metavar="STEPS",
                    help="the 256^3 fire for STEPS steps from rest with noise seeds 0-3")
    ap.add_argument("--config0", type=int, default=0, metavar="STEPS",
                    help="the 64^3 plume past a sphere for STEPS steps on the GPU and the host cores")
    ap.add_argument("--pressure", default="pcg", choices=("pcg", "fused"),
                    help="projection: MGPCG (the bench line) or the fused stationary multigrid of "
                         "fused_config() (2 V-cycles + 4 Chebyshev sweeps, adaptive to normalised residual 5e-5)")
    ap.add_argument("--scalar-res", type=int, default=0, metavar="S",
                    help="with the finer scalar grid (Solver3D(res, scalar_res=S)): density, temperature and "
                         "fuel carried on S^3 and restricted each step; the fine kernels join the table")
    ap.add_argument("--mg", default="fp16", choices=("fp32", "fp16", "bf16"),
                    help="multigrid level storage: FP16 with FP32 arithmetic (the default; the 256^3 mixed-precision "
                         "test keeps the iteration counts and the FP64 errors), or FP32 as the reference stores it")
    ap.add_argument("--precision-report", type=int, default=0, metavar="STEPS",
                    help="instead of the bench line: FP32 (FP32/FP16/BF16 levels) against the FP64 run of the "
                         "smoke plume past a sphere after STEPS steps at --res")
    ap.add_argument("--batch", type=int, default=0, metavar="FRAMES",
                    help="instead of the bench line: batched frames of 32 simulations of sizes 48-128 "
                         "(--warmup frames, autotuner, FRAMES timed frames)")
    ap.add_argument("--no-tune", action="store_true", help="--batch without the autotuner")
    ap.add_argument("--render", type=int, default=0, metavar="REPS",
                    help="volume renderer at 2560x1440 on 64^3 / 128^3 / 256^3 solver states (render3d)")
    ap.add_argument("--pair", type=int, default=0, metavar="STEPS",
                    help="paired runs (datagen3d run_pair): 64^3 runtime + 256^3 guided, phase split and tracking")
    ap.add_argument("--learned", type=int, default=0, metavar="ITERS",
                    help="instead of the bench line: train the U-Net ITERS Adam steps at --train-res, then "
                         "learned flexible CG vs MGPCG at 64^3, 128^3 and 256^3")
    ap.add_argument("--train-res", type=int, default=128)
    ap.add_argument("--learned-design", default="coarse", choices=("coarse", "full"),
                    help="the coarse correction of the V-cycle (a U-Net on the N/4 grid) or the full-resolution "
                         "U-Net preconditioner")
    ap.add_argument("--no-learned", action="store_true", help="bench line without the learned-vs-MGPCG keys")
    ap.add_argument("--no-batch", action="store_true", help="bench line without the batched-frames keys")
    ap.add_argument("--learned-iters", type=int, default=600,
                    help="training iterations of the coarse correction for the bench line's learned-vs-MGPCG keys")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="short run for counter collection: one 256 MiB device copy first (the scale reference), "
                         "2 steps, then one launch of every timed kernel; prints no bench line")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (the 256^3 step is one "
              "GPU's work; see DESIGN.md for how it would shard)", file=sys.stderr)
        sys.exit(2)

    import torch
    from vfx.sim3d import Solver3D, _lib

    if args.precision_report > 0:
        precision_report(args.res, args.precision_report)
        return
    if args.learned > 0:
        torch.cuda.set_device(torch.device("cuda:0"))
        learned_report(args.train_res, (64, 128, 256), args.learned, args.learned_design)
        return
    if args.batch > 0:
        torch.cuda.set_device(torch.device("cuda:0"))
        batch_report(args.batch, args.warmup, args.seed, tune=not args.no_tune)
        return
    if args.pair > 0:
        torch.cuda.set_device(torch.device("cuda:0"))
        pair_report(args.pair, args.warmup)
        return
    if args.render > 0:
        torch.cuda.set_device(torch.device("cuda:0"))
        render_report(args.render)
        return
    if args.config0 > 0:
        torch.cuda.set_device(torch.device("cuda:0"))
        config0_report(args.config0)
        return
    if args.seeds > 0:
        torch.cuda.set_device(torch.device("cuda:0"))
        seeds_report(args.seeds, args.res, args.mg)
        return
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    N = args.res
    dt = 1.0 / 60.0
    fused = args.pressure == "fused"
    if fused:
        args.mg = "fp32"   # the fused projection keeps its levels in FP32
    cfg = dataclasses.replace(fire_config(args.seed), mg_storage=args.mg)
    mg_bytes = 4 if args.mg == "fp32" else 2
    if fused:
        cfg = dataclasses.replace(cfg, pressure_mode="fused", pressure_tol=5e-5, pressure_max_iters=24)
    forces = fire_forces()
    S = args.scalar_res or None
    if args.pmc_probe:
        # the scale reference is the run's first device copy (the solver's set-up copies come later)
        src = torch.ones(COPY_PROBE_BYTES // 4, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    s = Solver3D(N, cfg, device="cuda:0", scalar_res=S)
    s.set_option(_lib.OPT_TRUE_RESIDUAL, 1)
    st = s.reset()
    if args.pmc_probe:
        for i in range(2):
            s.step(st, forces, dt)
        torch.cuda.synchronize()
        for name in [k for k in _lib.KERNELS if k]:
            t = ctypes.c_float()
            s._lib.vfx_time_kernel(s._ctx, _lib.KERNELS.index(name), 1, s._stream(), ctypes.byref(t))
        torch.cuda.synchronize()
        print(json.dumps({"pmc_probe": True, "res": N, "copy_bytes": COPY_PROBE_BYTES,
                          "kernel_bytes_per_launch": {k: int(kernel_bytes_per_cell(k, mg_bytes) * N ** 3)
                                                      for k in KERNEL_BYTES_PER_CELL}}))
        return
    print(f"# {N}^3 fire: {args.presim} pre-simulation + {args.warmup} warm-up steps", file=sys.stderr, flush=True)
    for i in range(args.presim + args.warmup):
        s.step(st, forces, dt)
    torch.cuda.synchronize()
    print("# timed steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        s.step(st, forces, dt)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = 1000.0 * (t1 - t0) / max(args.steps, 1)
    run_stats = s.pressure_stats()

    # stage split, iterations and residuals of a few more steps
    prof, iters, rel, true_rel = [], [], [], []
    for i in range(args.profiled):
        prof.append(s.step_profiled(st, forces, dt))
        ps = s.pressure_stats()
        iters.append(ps["iters"])
        rel.append(ps["rel_res"])
        true_rel.append(ps["true_rel_res"])
    med = lambda k: round(statistics.median(p[k] for p in prof), 3) if prof else None

    # live device time of every kernel on the solver's stream (HIP events, 20 launches each); the
    # roofline is reported for the kernel with the largest share of the step
    n_it = statistics.median(iters) if iters else 8
    kern = {}
    for name in [k for k in _lib.KERNELS if k]:
        t = ctypes.c_float()
        if s._lib.vfx_time_kernel(s._ctx, _lib.KERNELS.index(name), 20, s._stream(), ctypes.byref(t)) != 0:
            continue
        row = {"us": round(t.value * 1000.0, 1), "per_step_us": round(t.value * 1000.0 * launches_per_step(name, n_it), 1)}
        nbytes = kernel_bytes_per_cell(name, mg_bytes) * N ** 3 if name in KERNEL_BYTES_PER_CELL else \
            fine_kernel_bytes(name, N, S or N)
        if S is None and name.startswith("f"):
            continue
        if S is not None and name in ("adv_sl", "adv_mc", "sources", "diffuse"):
            continue   # with the fine grid the step runs these on the velocity only (3 channels)
        if nbytes:
            gbs = nbytes / (t.value * 1e-3) / 1e9
            row.update(gbs=round(gbs, 1), frac=round(gbs / HBM_PEAK_GBS, 3))
        kern[name] = row
    byte_of = lambda k: kernel_bytes_per_cell(k, mg_bytes) * N ** 3 if k in KERNEL_BYTES_PER_CELL else \
        fine_kernel_bytes(k, N, S)
    kname = args.roofline_kernel or max((k for k in kern if "gbs" in kern[k]), key=lambda k: kern[k]["per_step_us"])
    kbytes = byte_of(kname)
    kms = kern[kname]["us"] / 1000.0
    achieved = kbytes / (kms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": pmc_traffic(kname, N, S), "kernel": kname,
                "kernel_ms": round(kms, 4), "bytes_per_launch": int(kbytes)}
    sb = step_bytes_per_cell(n_it, mg_bytes) * N ** 3
    step_roof = {"bytes_per_step": int(sb), "achieved": round(sb / (ms * 1e-3) / 1e9, 1), "unit": "GB/s",
                 "frac": round(sb / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}

    cpu = None
    if not args.no_cpu and S is None:
        print("# CPU baseline (oracle, 5 steps)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(s, st.data, st.solid, cfg, forces, dt, N, st.meta.step)

    learned = None
    if not args.no_learned and S is None and not fused:
        # learned flexible CG (V-cycle + coarse U-Net correction, MGPCG fallback) against plain MGPCG
        # on the same 256^3 pre-projection state, both to true rel-res 1e-5 from a cold start
        say = lambda *a: print("#", *a, file=sys.stderr, flush=True)
        say(f"learned vs MGPCG: training the coarse correction at {args.train_res}^3 ({args.learned_iters} iterations)")
        params, summary = train_learned("coarse", args.train_res, args.learned_iters, None)
        row = compare_learned(N, params, "coarse")
        say(f"learned {row['learned']['pressure_ms']} ms vs MGPCG {row['mgpcg']['pressure_ms']} ms")
        learned = {"train": summary, **row}
    batch32 = None
    if not args.no_batch and S is None and not fused:
        print("# batched frames: 32 sims of {48,64,96,128}^3", file=sys.stderr, flush=True)
        batch32 = run_batch(10, 3, 0, 32, tune=True, compare=False)
    if fused or S is not None:
        step_roof = None   # the byte formula is the plain MGPCG step's
    solve = ("fused projection (2 V-cycles + 4 Chebyshev sweeps, adaptive to normalised residual 5e-5)" if fused
             else "plain MGPCG to rel-res 1e-5")
    out = {
        "metric": f"{N}^3 fire step ms, {solve}" + ("" if args.mg == "fp32" else f" ({args.mg} levels)")
                  + ("" if S is None else f", scalars on {S}^3"),
        "value": round(ms, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "steps_per_s": round(1000.0 / ms, 2), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic", "presim": args.presim,
        "config": {"workload": f"{N}^3 fire (fuel disc, combustion, buoyancy, confinement), "
                               + ("fused projection" if fused else "MGPCG rel-res 1e-5 (true residual judged)")
                               + f", dt 1/60, seed {args.seed}"
                               + ("" if S is None else f", density / temperature / fuel on a {S}^3 grid"),
                   "res": N, "scalar_res": S},
        "pressure_ms": med("projection"), "advection_ms": med("advection"), "forces_ms": med("forces"),
        "scalars_ms": med("scalars") if S is not None else None,
        "profiled_total_ms": med("total"), "pcg_iters": iters, "rel_res": [float(f"{r:.3g}") for r in rel],
        "true_rel_res": [float(f"{r:.3g}") for r in true_rel],
        "projections": {"sync": cfg.pressure_sync, "unconverged": run_stats["unconverged"],
                        "budget_stops": run_stats["budget_stops"], "budget": run_stats["budget"],
                        "graph_captures": run_stats["graph_captures"], "graph_replays": run_stats["graph_replays"]},
        "learned_pressure_ms": learned["learned"]["pressure_ms"] if learned else None,
        "mgpcg_pressure_ms": learned["mgpcg"]["pressure_ms"] if learned else None,
        "learned": learned,
        "batch32_cell_steps_per_s": round(batch32["cell_steps_per_s"], 1) if batch32 else None,
        "batch32": batch32,
        "roofline": roofline, "step_roofline": step_roof, "cpu_baseline": cpu, "kernels": kern,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

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
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "neural-vfx_amd"))

HBM_PEAK_GBS = 8000.0

# Algorithmic (compulsory) bytes per grid cell of each kernel, per launch: every distinct input
# element read once, every output written once, at its storage type (FP32 fields, uint8 mask,
# FP32 multigrid levels).  Multigrid kernels are per fine-grid cell (level 1 has 8 components on
# (N/2)^3 blocks = N^3 unknowns).
KERNEL_BYTES_PER_CELL = {
    "adv_sl": 24 + 24,            # data (6ch, the velocity is 3 of them) -> hat (6ch)
    "adv_mc": 24 + 24 + 1 + 24,   # data + hat + mask -> adv
    "sources": 24 + 1 + 24,       # adv + mask -> hat
    "diffuse": 24 + 1 + 24,       # hat + mask -> data
    "matvec": 4 + 1 + 4,          # p + mask -> q
    "post_rz": 4 + 4 + 1 + 4,     # zt + r + mask -> z
    "pre": 4 + 4 + 1 + 4 + 4,     # r + q + mask -> r, z0 (k > 0)
    "resid": 4 + 4 + 1 + 4,       # r + z0 + mask -> rr
    "upd_p": 4 + 4 + 4 + 4 + 4,   # phi + p + z -> phi, p
}


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


def cpu_baseline(st_data, st_solid, cfg, forces, dt, res, step):
    """The oracle's FP32 step on the host cores, one step of the same state (a bounded sample)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from vfx.sim3d.solver import pack_params
    data = st_data.cpu().numpy()
    solid = st_solid.cpu().numpy()
    prm = pack_params(cfg, res, forces, dt, step)
    phi = np.zeros((res,) * 3, np.float32)
    t0 = time.perf_counter()
    _, it = oracle.step(data, solid, prm, phi, np.float32(res))
    t = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(1000.0 * t, 1), "unit": "ms/step", "cores": cores, "kind": "port",
            "sample": f"1 step of the {res}^3 fire state after the GPU warm-up (cold pressure start, {it} PCG "
                      f"iterations), C oracle with OpenMP"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--res", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--profiled", type=int, default=5, help="extra step_profiled steps for the stage split")
    ap.add_argument("--roofline-kernel", default="adv_mc")
    ap.add_argument("--time-kernels", default="", help="comma list of extra kernels to time (us)")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (the 256^3 step is one "
              "GPU's work; see DESIGN.md for how it would shard)", file=sys.stderr)
        sys.exit(2)

    import torch
    from vfx.sim3d import Solver3D, _lib

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    N = args.res
    dt = 1.0 / 60.0
    cfg = fire_config(args.seed)
    forces = fire_forces()
    s = Solver3D(N, cfg, device="cuda:0")
    s.set_option(_lib.OPT_TRUE_RESIDUAL, 1)
    st = s.reset()
    print(f"# {N}^3 fire: {args.warmup} warm-up steps", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        s.step(st, forces, dt)
    torch.cuda.synchronize()
    print("# timed steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        s.step(st, forces, dt)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = 1000.0 * (t1 - t0) / max(args.steps, 1)

    # stage split, iterations and residuals of a few more steps
    prof, iters, rel, true_rel = [], [], [], []
    for i in range(args.profiled):
        prof.append(s.step_profiled(st, forces, dt))
        ps = s.pressure_stats()
        iters.append(ps["iters"])
        rel.append(ps["rel_res"])
        true_rel.append(ps["true_rel_res"])
    med = lambda k: round(statistics.median(p[k] for p in prof), 3) if prof else None

    # live roofline of the dominant kernel on the solver's stream
    kname = args.roofline_kernel
    kid = _lib.KERNELS.index(kname)
    kms = ctypes.c_float()
    _lib.check(s._lib.vfx_time_kernel(s._ctx, kid, 20, s._stream(), ctypes.byref(kms)))
    kbytes = KERNEL_BYTES_PER_CELL[kname] * N ** 3
    achieved = kbytes / (kms.value * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kname,
                "kernel_ms": round(kms.value, 4), "bytes_per_launch": kbytes}
    others = {}
    for name in [n for n in args.time_kernels.split(",") if n]:
        t = ctypes.c_float()
        if s._lib.vfx_time_kernel(s._ctx, _lib.KERNELS.index(name), 20, s._stream(), ctypes.byref(t)) == 0:
            others[name] = round(t.value * 1000.0, 1)

    cpu = None
    if not args.no_cpu:
        print("# CPU baseline (oracle, one step)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(st.data, st.solid, cfg, forces, dt, N, st.meta.step)

    out = {
        "metric": f"{N}^3 fire step ms, plain MGPCG to rel-res 1e-5",
        "value": round(ms, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "steps_per_s": round(1000.0 / ms, 2), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{N}^3 fire (fuel disc, combustion, buoyancy, confinement), MGPCG rel-res 1e-5 "
                               f"(true residual judged), dt 1/60, seed {args.seed}", "res": N},
        "pressure_ms": med("projection"), "advection_ms": med("advection"), "forces_ms": med("forces"),
        "profiled_total_ms": med("total"), "pcg_iters": iters, "rel_res": [float(f"{r:.3g}") for r in rel],
        "true_rel_res": [float(f"{r:.3g}") for r in true_rel],
        "roofline": roofline, "cpu_baseline": cpu,
        **({"kernel_us": others} if others else {}),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

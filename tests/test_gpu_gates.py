"""The reference's gate tests (tests3d/test_sim3d_layer1.py) on the HIP path, through the public
Solver3D API at 64^3 (and the reference's 32-128^3 companions), with the reference's thresholds
unchanged.  The scenes, forces and step counts are the reference's; the fields are built with this
repository's helpers (tests/simhelp.py), which follow tests3d/helpers3d.py.

Ported here (the remaining gates are in test_gpu_invariants.py, test_gpu_fused.py and
test_gpu_hires.py): pure-gradient flow into walls (:148-165), mass over 1000 steps in a closed box
and around a sphere (:169-195), integer-shift and analytic-translation advection (:199-245),
advection convergence 32 -> 96 (:258-275), buoyancy (:279-320), ambient rest (:323-330), no leak
through the box walls (:358-371), convergence against a 96^3 reference and 64^3 vs 128^3
agreement (:409-453), the 1000-step fire and impulse stress run (:457-483), the fused-mode gates
(:604-640) and the scalar_res 1000-step stress run (:740-760).
"""
from __future__ import annotations

import hashlib
import math
import sys

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from simhelp import blob, smooth_noise, to_dev, world_grid
from vfx.core.fields import AMBIENT_TEMP, DOMAIN_SIZE, SIM_DT
from vfx.core.fields3d import Emitter3D, Forces3D, Impulse3D, Obstacle3D
from vfx.sim3d import Solver3D, SolverConfig3D, fused_config

pytestmark = pytest.mark.gpu
RES = 64
DIV_TOL = 1e-4    # the reference's spec (test_sim3d_layer1.py:36)
MASS_TOL = 1e-3   # 0.1 % (:37)
DEV = "cuda:0"


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _mass(st):
    d = st.ch("density") * (st.solid < 0.5)
    return float(d.double().sum().item()) * st.meta.dx ** 3


def _finite(st, where=""):
    bad = (~torch.isfinite(st.data)).sum().item()
    assert bad == 0, f"{bad} non-finite values {where}"


def _blob(res, cx, cy, cz, sigma):
    return to_dev(blob(res, cx, cy, cz, sigma))


def _grid(res):
    return [to_dev(g) for g in world_grid(res)]


def _sphere():
    return [Obstacle3D(kind="sphere", pos=(0.5, 0.45, 0.5), size=(0.14, 0.14, 0.14))]


def test_projection_removes_flow_into_walls(cuda):
    """u = grad(x + 0.5 y - 0.3 z) projected in the closed box is zero: residual |u| and the net flux
    through every slab < 1e-2 (:148-165)."""
    s = Solver3D(RES)
    st = s.reset()
    st.velocity[0].fill_(1.0)
    st.velocity[1].fill_(0.5)
    st.velocity[2].fill_(-0.3)
    s.project(st)
    _finite(st)
    u = st.velocity
    resid = u.abs().max().item()
    fx = u[0].mean(dim=(0, 1)).abs().max().item()
    fy = u[1].mean(dim=(0, 2)).abs().max().item()
    fz = u[2].mean(dim=(1, 2)).abs().max().item()
    _say(f"pure gradient: max|u| {resid:.3e}, slab flux {fx:.2e} {fy:.2e} {fz:.2e}")
    assert resid < 1e-2 and max(fx, fy, fz) < 1e-2


def test_mass_conserved_1000_steps_closed_box(cuda):
    """Stirred blobs in a closed box, 1000 steps, dissipation 0: mass within 0.1 % (:169-183)."""
    s = Solver3D(RES, SolverConfig3D(density_dissipation=0.0))
    st = s.reset()
    st.ch("density").copy_(_blob(RES, 0.5, 0.35, 0.5, 0.08) + 0.5 * _blob(RES, 0.3, 0.7, 0.6, 0.06))
    st.ch("temperature").copy_(AMBIENT_TEMP + 400.0 * _blob(RES, 0.5, 0.35, 0.5, 0.08))
    st.velocity.copy_(to_dev(smooth_noise(RES, 3, seed=3, amp=0.5)))
    m0 = _mass(st)
    for _ in range(1000):
        s.step(st, Forces3D(), SIM_DT)
    _finite(st)
    rel = abs(_mass(st) - m0) / m0
    _say(f"closed box 1000 steps: rel mass change {rel:.3e}")
    assert rel < MASS_TOL


def test_mass_conserved_1000_steps_with_obstacle(cuda):
    """A hot blob under a sphere, 1000 steps: mass within 0.1 % (:185-195)."""
    s = Solver3D(RES, SolverConfig3D(density_dissipation=0.0))
    st = s.reset(_sphere())
    st.ch("density").copy_(_blob(RES, 0.5, 0.15, 0.5, 0.07) * (st.solid < 0.5))
    st.ch("temperature").copy_(AMBIENT_TEMP + 500.0 * _blob(RES, 0.5, 0.15, 0.5, 0.07))
    m0 = _mass(st)
    for _ in range(1000):
        s.step(st, Forces3D(), SIM_DT)
    _finite(st)
    rel = abs(_mass(st) - m0) / m0
    _say(f"sphere 1000 steps: rel mass change {rel:.3e}")
    assert rel < MASS_TOL


def test_advect_integer_shift_is_exact(cuda):
    """Whole-cell displacement per step: exact in the interior, max err < 1e-4 (:199-216)."""
    s = Solver3D(RES)
    dx = DOMAIN_SIZE / RES
    f = _blob(RES, 0.3, 0.35, 0.55, 0.07)
    vel = torch.zeros(3, RES, RES, RES, device=DEV)
    vel[0], vel[1], vel[2] = 2 * dx / SIM_DT, 1 * dx / SIM_DT, -1 * dx / SIM_DT
    n = 10
    for _ in range(n):
        f = s.advect(f, vel, SIM_DT)
    expect = _blob(RES, 0.3 + 2 * n * dx, 0.35 + n * dx, 0.55 - n * dx, 0.07)
    err = (f - expect)[: RES - n - 1, n + 1:, 2 * n + 1:].abs().max().item()
    _say(f"integer shift: max err {err:.2e}")
    assert err < 1e-4


def test_advect_matches_analytic_translation(cuda):
    """Fractional-cell constant velocity: centroid within 0.5 cell, rel L2 < 0.15 (:219-245)."""
    s = Solver3D(RES)
    dx = DOMAIN_SIZE / RES
    u = (0.15, 0.08, 0.06)
    sig = 0.09
    c0 = (0.35, 0.38, 0.40)
    f = _blob(RES, *c0, sig)
    vel = torch.zeros(3, RES, RES, RES, device=DEV)
    vel[0], vel[1], vel[2] = u
    n = 45
    for _ in range(n):
        f = s.advect(f, vel, SIM_DT)
    ce = [c0[i] + u[i] * n * SIM_DT for i in range(3)]
    ex = _blob(RES, *ce, sig)
    xx, yy, zz = _grid(RES)
    cen = [((f * g).sum() / f.sum()).item() for g in (xx, yy, zz)]
    cen_ex = [((ex * g).sum() / ex.sum()).item() for g in (xx, yy, zz)]
    cerr = math.dist(cen, cen_ex) / dx
    rel_l2 = ((f - ex).norm() / ex.norm()).item()
    _say(f"translation: centroid err {cerr:.3f} cells, rel L2 {rel_l2:.3f}")
    assert cerr < 0.5 and rel_l2 < 0.15


def test_advection_converges_with_resolution(cuda):
    """The same world-space translation at 32, 48, 64, 96: the error decreases (:258-275)."""
    errs = {}
    u = (0.11, 0.05, 0.07)
    for res in (32, 48, 64, 96):
        s = Solver3D(res)
        vel = torch.zeros(3, res, res, res, device=DEV)
        vel[0], vel[1], vel[2] = u
        f = _blob(res, 0.35, 0.35, 0.35, 0.1)
        n = 30
        for _ in range(n):
            f = s.advect(f, vel, SIM_DT)
        ex = _blob(res, *(0.35 + u[i] * n * SIM_DT for i in range(3)), 0.1)
        errs[res] = ((f - ex).abs().mean() / ex.abs().mean()).item()
    _say(f"advection rel L1: {({k: round(v, 4) for k, v in errs.items()})}")
    r = sorted(errs)
    for a, b in zip(r, r[1:]):
        assert errs[b] < errs[a], errs


def test_buoyancy_lifts_hot_fluid_straight_up(cuda):
    """A hot blob rises along +y with no sideways drift (:279-305)."""
    cfg = SolverConfig3D(buoyancy_alpha=0.0, vorticity_eps=0.0, burn_rate=0.0)
    s = Solver3D(RES, cfg)
    st = s.reset()
    hot = _blob(RES, 0.5, 0.3, 0.5, 0.08)
    st.ch("temperature").copy_(AMBIENT_TEMP + 500.0 * hot)
    st.ch("density").copy_(hot)
    xx, yy, zz = _grid(RES)

    def centroid():
        d = st.ch("density")
        return [((d * g).sum() / d.sum()).item() for g in (xx, yy, zz)]

    c0 = centroid()
    for _ in range(20):
        s.step(st, Forces3D(), SIM_DT)
    c1 = centroid()
    w = hot / hot.sum()
    vx, vy, vz = [(st.velocity[i] * w).sum().item() for i in range(3)]
    _say(f"hot parcel: centroid {c0} -> {c1}, mean v ({vx:.4f}, {vy:.4f}, {vz:.4f})")
    assert vy > 0 and c1[1] > c0[1] + 0.01
    assert abs(vx) < 0.1 * vy and abs(vz) < 0.1 * vy
    assert abs(c1[0] - c0[0]) < 0.1 * (c1[1] - c0[1]) + 1e-3
    assert abs(c1[2] - c0[2]) < 0.1 * (c1[1] - c0[1]) + 1e-3


def test_dense_cold_fluid_sinks(cuda):
    """A dense blob with beta 0 moves down (:308-320)."""
    alpha = SolverConfig3D().buoyancy_alpha
    s = Solver3D(RES, SolverConfig3D(buoyancy_beta=0.0, buoyancy_alpha=alpha, vorticity_eps=0.0, burn_rate=0.0))
    st = s.reset()
    d0 = 2.0 * _blob(RES, 0.5, 0.7, 0.5, 0.08)
    st.ch("density").copy_(d0)
    for _ in range(20):
        s.step(st, Forces3D(), SIM_DT)
    vy = ((st.velocity[1] * d0).sum() / d0.sum()).item()
    _say(f"dense parcel mean vy {vy:.4f}")
    assert vy < 0


def test_ambient_fluid_stays_at_rest(cuda):
    """Ambient fluid around a box, 50 steps: |u| < 1e-5, density < 1e-6 (:323-330)."""
    s = Solver3D(RES)
    st = s.reset([Obstacle3D("box", (0.5, 0.5, 0.5), (0.1, 0.1, 0.1))])
    for _ in range(50):
        s.step(st, Forces3D(), SIM_DT)
    assert st.velocity.abs().max().item() < 1e-5
    assert st.ch("density").abs().max().item() < 1e-6


def test_no_leak_through_box_walls(cuda):
    """Impulses aimed into the walls and corners every 10 steps, 150 steps, dissipation 0: mass within
    0.1 % (:358-371)."""
    s = Solver3D(RES, SolverConfig3D(density_dissipation=0.0))
    st = s.reset()
    st.ch("density").copy_(_blob(RES, 0.85, 0.85, 0.85, 0.06) + _blob(RES, 0.12, 0.1, 0.15, 0.05))
    m0 = _mass(st)
    for i in range(150):
        imp = [Impulse3D(pos=(0.85, 0.85, 0.85), radius=0.1, velocity=(3.0, 3.0, 3.0)),
               Impulse3D(pos=(0.12, 0.1, 0.15), radius=0.1, velocity=(-3.0, -3.0, -3.0))] if i % 10 == 0 else []
        s.step(st, Forces3D(impulses=imp), SIM_DT)
    _finite(st)
    rel = abs(_mass(st) - m0) / m0
    _say(f"wall jets: rel mass change {rel:.3e}")
    assert rel < MASS_TOL


PLUME = Forces3D(emitters=[Emitter3D(pos=(0.5, 0.2, 0.5), radius=0.1, density_rate=1.0, temperature=600.0)])


def _plume(res, steps, forces=PLUME):
    s = Solver3D(res)
    st = s.reset()
    for _ in range(steps):
        s.step(st, forces, SIM_DT)
    _finite(st, f"res {res}")
    return st


def test_convergence_against_fine_reference(cuda):
    """20-step plume: the error against a 96^3 run on a common 16^3 grid shrinks from 32 to 48 to 64
    (:409-419)."""
    pool = lambda st: F.adaptive_avg_pool3d(st.ch("density")[None, None], 16)[0, 0]
    ref = pool(_plume(96, 20))
    errs = {res: ((pool(_plume(res, 20)) - ref).abs().mean() / ref.abs().mean()).item() for res in (32, 48, 64)}
    _say(f"plume rel L1 vs 96^3: {({k: round(v, 4) for k, v in errs.items()})}")
    assert errs[48] < errs[32] and errs[64] < errs[48], errs


def _plume_stats(st):
    d = st.ch("density") * (st.solid < 0.5)
    dv = st.meta.dx ** 3
    _, yy, _ = _grid(d.shape[0])
    smoke = d.double().sum().item() * dv
    heat = ((st.ch("temperature") - AMBIENT_TEMP) * (st.solid < 0.5)).double().sum().item() * dv
    cy = ((d * yy).sum() / d.sum()).item()
    slab = d.sum(dim=(0, 2))
    rows = torch.nonzero(slab > 0.02 * slab.max()).flatten()
    top = (rows.max().item() + 1) * st.meta.dx
    return {"smoke": smoke, "heat": heat, "centroid_y": cy, "top_y": top}


def test_large_scale_agreement_64_vs_128(cuda):
    """The same campfire at 64^3 and 128^3 agrees at large scales: smoke and heat within 10 %,
    centroid height within 0.04 m, plume top within 0.0625 m (:436-453)."""
    scene = Forces3D(emitters=[Emitter3D(pos=(0.5, 0.15, 0.5), radius=0.08, density_rate=1.0, temperature=700.0,
                                         velocity=(0.0, 0.5, 0.0))])
    a = _plume_stats(_plume(64, 60, scene))
    b = _plume_stats(_plume(128, 60, scene))
    rs = abs(a["smoke"] - b["smoke"]) / b["smoke"]
    rh = abs(a["heat"] - b["heat"]) / max(abs(b["heat"]), 1e-12)
    dc = abs(a["centroid_y"] - b["centroid_y"])
    dt_ = abs(a["top_y"] - b["top_y"])
    _say(f"64 vs 128: smoke {rs:.3f} heat {rh:.3f} centroid dy {dc:.4f} top dy {dt_:.4f}")
    assert b["smoke"] > 0 and b["centroid_y"] > 0.15
    assert rs < 0.10 and rh < 0.10 and dc < 0.04 and dt_ < 0.0625


def test_stress_1000_steps_fire_and_impulses(cuda):
    """Two fuel emitters, random impulses every 5 steps and a turning wind around a sphere and a box,
    1000 steps: finite, |u| < 100 m/s, density < 1e3, T < 1e4 K (:457-483)."""
    s = Solver3D(RES)
    st = s.reset([Obstacle3D("sphere", (0.35, 0.6, 0.5), (0.09, 0.09, 0.09)),
                  Obstacle3D("box", (0.7, 0.45, 0.4), (0.06, 0.14, 0.06))])
    g = torch.Generator().manual_seed(0)
    umax = dmax = tmax = float("nan")
    for i in range(1000):
        imp = []
        if i % 5 == 0:
            p = torch.rand(3, generator=g).tolist()
            v = ((torch.rand(3, generator=g) * 2 - 1) * 8.0).tolist()
            imp = [Impulse3D(pos=tuple(p), radius=0.12, velocity=tuple(v))]
        f = Forces3D(emitters=[Emitter3D(pos=(0.5, 0.1, 0.5), radius=0.09, density_rate=5.0, temperature=2000.0,
                                         fuel_rate=5.0, velocity=(0.0, 3.0, 0.0)),
                               Emitter3D(pos=(0.2, 0.15, 0.6), radius=0.06, density_rate=3.0, temperature=1800.0,
                                         fuel_rate=4.0, velocity=(1.5, 2.0, -0.5))],
                     impulses=imp, wind=(0.5 * math.sin(i * 0.01), 0.0, 0.3 * math.cos(i * 0.013)))
        s.step(st, f, SIM_DT)
        if i % 100 == 99:
            _finite(st, f"at step {i + 1}")
            umax = st.velocity.abs().max().item()
            dmax = st.ch("density").max().item()
            tmax = st.ch("temperature").max().item()
            assert umax < 100.0 and dmax < 1e3 and tmax < 1e4, (i, umax, dmax, tmax)
    _say(f"stress 1000 steps: umax {umax:.2f} m/s, dmax {dmax:.2f}, Tmax {tmax:.0f} K")


# ----------------------------------------------------------------------------- fused mode (:604-640)
@pytest.mark.parametrize("obst", [False, True], ids=["open", "sphere"])
def test_fused_divergence_after_projection(cuda, obst):
    """pressure_mode='fused': the project() of a strongly divergent field meets div < 1e-4 (the
    reference accepts a documented lower-quality level instead; this build meets the number)."""
    s = Solver3D(RES, SolverConfig3D(pressure_mode="fused"))
    st = s.reset(_sphere() if obst else [])
    v = smooth_noise(RES, 3, 1, amp=2.0) + np.array([0.8, -0.3, 0.4], np.float32)[:, None, None, None]
    st.velocity.copy_(to_dev(v) * (st.solid < 0.5))
    s.project(st)
    _finite(st)
    div = s.divergence(st)
    m = div.max().item()
    _say(f"fused project ({'sphere' if obst else 'open'}): max normalised div {m:.3e}")
    assert torch.all(div[st.solid > 0.5] == 0)
    assert m < DIV_TOL


@pytest.mark.parametrize("obst", [False, True], ids=["open", "sphere"])
def test_fused_divergence_after_full_steps(cuda, obst):
    s = Solver3D(RES, SolverConfig3D(pressure_mode="fused"))
    st = s.reset(_sphere() if obst else [])
    f = Forces3D(emitters=[Emitter3D(pos=(0.5, 0.15, 0.5), radius=0.07, density_rate=2.0, temperature=900.0,
                                     velocity=(0.0, 1.0, 0.0))])
    for _ in range(60):
        s.step(st, f, SIM_DT)
    _finite(st)
    m = s.divergence(st).max().item()
    _say(f"fused 60 steps ({'sphere' if obst else 'open'}): max normalised div {m:.3e}")
    assert m < DIV_TOL


def _hash_run(cfg, steps=120):
    s = Solver3D(RES, cfg)
    st = s.reset(_sphere())
    for i in range(steps):
        f = Forces3D(emitters=[Emitter3D(pos=(0.5, 0.12, 0.5), radius=0.06, density_rate=2.0, temperature=1300.0,
                                         fuel_rate=1.5, velocity=(0.0, 1.0, 0.0))],
                     impulses=[Impulse3D(pos=(0.3, 0.5, 0.4), radius=0.1, velocity=(2.0, 0.5, -1.0))]
                     if i % 17 == 0 else [], wind=(0.2, 0.0, 0.1))
        s.step(st, f, SIM_DT)
    _finite(st)
    return hashlib.sha256(st.data.cpu().numpy().tobytes()).hexdigest()


def test_fused_deterministic_and_finite(cuda):
    """fused_config(): 120 scripted steps twice give bit-identical states (:633-637)."""
    assert _hash_run(fused_config()) == _hash_run(fused_config())


# ----------------------------------------------------------------------------- scalar_res stress (:740-760)
def test_scalar_res_stress_1000_steps_finite(cuda):
    """scalar_res = 128 with a fuel emitter and random impulses, 1000 steps: coarse and fine fields
    finite, fine density >= -1e-4 and < 1e3 every 100 steps."""
    s = Solver3D(RES, scalar_res=128)
    st = s.reset([Obstacle3D("sphere", (0.35, 0.6, 0.5), (0.09, 0.09, 0.09))])
    g = torch.Generator().manual_seed(0)
    for i in range(1000):
        imp = []
        if i % 5 == 0:
            p = torch.rand(3, generator=g).tolist()
            v = ((torch.rand(3, generator=g) * 2 - 1) * 8.0).tolist()
            imp = [Impulse3D(pos=tuple(p), radius=0.12, velocity=tuple(v))]
        f = Forces3D(emitters=[Emitter3D(pos=(0.5, 0.1, 0.5), radius=0.09, density_rate=5.0, temperature=2000.0,
                                         fuel_rate=5.0, velocity=(0.0, 3.0, 0.0))], impulses=imp)
        s.step(st, f, SIM_DT)
        if i % 100 == 99:
            _finite(st, f"at step {i + 1}")
            v = s.hires_volume()
            nb = (~torch.isfinite(v.density)).sum().item() + (~torch.isfinite(v.temperature)).sum().item()
            assert nb == 0, f"{nb} non-finite fine values at step {i + 1}"
            assert v.density.float().min().item() >= -1e-4
            assert v.density.float().max().item() < 1e3
    _say(f"scalar_res stress: fine density max {s.hires_volume().density.max().item():.3f}")


@pytest.mark.parametrize("obst", [False, True], ids=["open", "sphere"])
def test_scalar_res_fine_mass_conserved_1000_steps(cuda, obst):
    """scalar_res = 128: fill the fine field with an emitter for 30 steps, then 1000 sourceless steps
    with dissipation 0: the fine mass stays within 0.1 % (:644-683)."""
    from vfx.core.fields3d import rasterize_obstacles_3d
    obstacles = _sphere() if obst else []
    s = Solver3D(RES, SolverConfig3D(density_dissipation=0.0), scalar_res=128)
    st = s.reset(obstacles)
    fill = Forces3D(emitters=[Emitter3D(pos=(0.5, 0.2, 0.5), radius=0.08, density_rate=2.0, temperature=700.0,
                                        velocity=(0.0, 0.6, 0.0))])
    for _ in range(30):
        s.step(st, fill, SIM_DT)
    solid_f = rasterize_obstacles_3d(obstacles, 128, DEV)

    def fine_mass():
        d = s.hires_volume().density.float() * (solid_f < 0.5)
        return float(d.double().sum().item()) * (DOMAIN_SIZE / 128) ** 3

    m0 = fine_mass()
    assert m0 > 0
    for _ in range(1000):
        s.step(st, Forces3D(), SIM_DT)
    m1 = fine_mass()
    _finite(st)
    rel = abs(m1 - m0) / m0
    _say(f"scalar_res fine mass ({'sphere' if obst else 'open'}): {m0:.5f} -> {m1:.5f} (rel {rel:.3e})")
    assert rel < MASS_TOL

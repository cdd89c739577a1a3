"""GPU tests of the finer scalar grid (Solver3D(res, scalar_res=S), hires.py): the HIP fine stage,
restriction and whole steps against the oracle (pinned on the CPU by tests/test_oracle_hires.py)
on the same inputs, equality with the plain solver at S == res, and the reference's own
scalar_res gates at a full-size grid (selfcheck.py:262-323: fine mass conservation, restriction
consistency, determinism, emitter footprint).

Bars: the fine stage has exp() in its sources, so the FP32 fields agree with the FP32 oracle to
1e-5 of their max (advection and diffusion alone round identically); whole steps with a fixed 5
PCG iterations to 1e-4 as the plain step parity; the FP64 run to 1e-10.
"""
from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest
import torch
import torch.nn.functional as F

import oracle
from simhelp import rel_err, random_state, smooth_noise, sphere_solid, to_dev
from vfx.core.fields import AMBIENT_TEMP
from vfx.core.fields3d import Emitter3D, Forces3D, Impulse3D, Obstacle3D
from vfx.sim3d import Solver3D, SolverConfig3D, runtime_config
from vfx.sim3d import _lib
from vfx.sim3d.solver import pack_params

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
DT = 1.0 / 60.0

FIRE = Forces3D(emitters=[Emitter3D(pos=(0.5, 0.12, 0.5), radius=0.09, density_rate=2.0, temperature=1300.0,
                                    fuel_rate=1.5, velocity=(0.0, 1.0, 0.0)),
                          Emitter3D(pos=(0.3, 0.2, 0.6), radius=0.07, density_rate=1.0, temperature=900.0)],
                impulses=[Impulse3D(pos=(0.3, 0.5, 0.4), radius=0.1, velocity=(2.0, 0.5, -1.0))],
                wind=(0.2, 0.0, 0.1))
SPHERE = Obstacle3D("sphere", (0.5, 0.45, 0.5), (0.14, 0.14, 0.14))


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _fine_scalars(S, seed, hsolid, dtype=np.float32):
    d = np.abs(smooth_noise(S, 1, seed, amp=0.5))[0]
    T = AMBIENT_TEMP + np.abs(smooth_noise(S, 1, seed + 1, amp=500.0))[0]
    f = np.abs(smooth_noise(S, 1, seed + 2, amp=0.4))[0]
    fl = hsolid < 0.5
    out = np.stack([d * fl, np.where(fl, T, AMBIENT_TEMP), f * fl]).astype(dtype)
    return np.ascontiguousarray(out)


def _setup(N, S, cfg, seed, dtype=torch.float32):
    """A solver with the sphere obstacle at both resolutions, a random coarse state and random
    fine scalars; returns (solver, state, data, solid, hi, hsolid) with numpy copies."""
    npdt = np.float32 if dtype == torch.float32 else np.float64
    s = Solver3D(N, cfg, scalar_res=S, dtype=dtype)
    st = s.reset([SPHERE])
    solid = st.solid.cpu().numpy()
    hsolid = s.hi_solid.cpu().numpy()
    data = random_state(N, seed, solid, vel_amp=1.5).astype(npdt)
    hi = _fine_scalars(S, seed + 10, hsolid, npdt)
    st.data.copy_(to_dev(data))
    s.hi.copy_(to_dev(hi))
    return s, st, data, solid, hi, hsolid


@pytest.mark.parametrize("N,S", [(32, 48), (32, 64), (24, 72), (32, 32)])
def test_fine_stage_and_restriction_parity(cuda, hiplib, N, S):
    """The fine stage (advection, mass ratio, sources, diffusion) and the forces stage with the
    restriction, each from the same inputs as the oracle's."""
    _say(f"fine stage N={N} S={S}")
    cfg = SolverConfig3D(heat_diffusion=2e-4)
    s, st, data, solid, hi, hsolid = _setup(N, S, cfg, 31)
    prm = pack_params(cfg, N, FIRE, DT, 0)
    api = s._api
    for stage, hp in ((_lib.STAGE_ADVECT, prm.ctypes.data_as(ctypes.c_void_p)), (_lib.STAGE_SCALARS, None)):
        s._chk(api.vfx_step_stage(s._ctx, stage, _ptr(st.data), _ptr(st.solid), hp, _stream()))
    got_hi = s.hi.cpu().numpy()
    ref_hi = hi.copy()
    ratio = oracle.hires_stage(ref_hi, hsolid, data[1:4], prm, np.float32(N))
    for c, name in enumerate(("density", "T", "fuel")):
        e = rel_err(got_hi[c], ref_hi[c])
        _say(f"  fine {name}: {e:.3e}")
        assert e <= 1e-5, name
    _say(f"  fine mass ratio {ratio:.9f}")
    s._chk(api.vfx_step_stage(s._ctx, _lib.STAGE_FORCES, _ptr(st.data), _ptr(st.solid), None, _stream()))
    got = st.data.cpu().numpy()
    adv, cr = oracle.stage_advect(data, solid, prm, np.float32(N))
    ref = oracle.stage_forces(adv, solid, prm, cr, np.float32(N), hi=got_hi)
    for c, name in enumerate(("density", "vx", "vy", "vz", "T", "fuel")):
        e = rel_err(got[c], ref[c])
        _say(f"  forces {name}: {e:.3e}")
        assert e <= 1e-5, name
    # the restriction alone (from the GPU's own fine field) rounds like the oracle's
    restr = oracle.hires_restrict(got_hi, solid, prm, got)
    for ch in (0, 4, 5):
        assert rel_err(got[ch], restr[ch]) <= 1e-6


@pytest.mark.parametrize("N,S", [(32, 48), (32, 64)])
def test_hires_steps_parity(cuda, hiplib, N, S):
    """Three whole steps (fixed 5 PCG iterations) against the oracle's steps with the fine grid."""
    _say(f"hires steps N={N} S={S}")
    cfg = runtime_config(pressure_max_iters=5, heat_diffusion=1e-4)
    s, st, data, solid, hi, hsolid = _setup(N, S, cfg, 41)
    phi = np.zeros((N,) * 3, np.float32)
    ref, ref_hi = data.copy(), hi.copy()
    for k in range(3):
        s.step(st, FIRE, DT)
        prm = pack_params(cfg, N, FIRE, DT, k)
        ref, it = oracle.step(ref, solid, prm, phi, np.float32(N), hi=ref_hi, hsolid=hsolid)
        assert s.pressure_iters() == it == 5
    got, got_hi = st.data.cpu().numpy(), s.hi.cpu().numpy()
    for c, name in enumerate(("density", "vx", "vy", "vz", "T", "fuel")):
        e = rel_err(got[c], ref[c])
        _say(f"  {name}: {e:.3e}")
        assert e <= 1e-4, name
    for c, name in enumerate(("fine density", "fine T", "fine fuel")):
        e = rel_err(got_hi[c], ref_hi[c])
        _say(f"  {name}: {e:.3e}")
        assert e <= 1e-4, name


def test_same_size_equals_plain_solver(cuda, hiplib):
    """scalar_res == res: the fine grid is the simulation grid, so every step equals the plain
    solver's bit for bit, and the fine scalars equal the state's."""
    N = 48
    cfg = SolverConfig3D(heat_diffusion=1e-4, emitter_noise=0.3)
    a, b = Solver3D(N, cfg), Solver3D(N, cfg, scalar_res=N)
    sa, sb = a.reset([SPHERE]), b.reset([SPHERE])
    for k in range(5):
        a.step(sa, FIRE, DT)
        b.step(sb, FIRE, DT)
        torch.cuda.synchronize()
        d = (sa.data - sb.data).abs().max().item()
        _say(f"  step {k}: max |plain - hires| {d:.3e}")
        assert torch.equal(sa.data, sb.data), k
        assert torch.equal(b.hi, sb.data[[0, 4, 5]]), k


def test_fp64_fine_stage_parity(cuda, hiplib):
    N, S = 24, 48
    cfg = runtime_config(pressure_max_iters=5, heat_diffusion=1e-4)
    s, st, data, solid, hi, hsolid = _setup(N, S, cfg, 51, dtype=torch.float64)
    s.step(st, FIRE, DT)
    prm = pack_params(cfg, N, FIRE, DT, 0)
    ref, _ = oracle.step(data, solid, prm, np.zeros((N,) * 3), float(N), prec="f64", hi=hi, hsolid=hsolid)
    e = max(rel_err(st.data.cpu().numpy()[c], ref[c]) for c in range(6))
    eh = max(rel_err(s.hi.cpu().numpy()[c], hi[c]) for c in range(3))
    _say(f"  fp64 hires step: state {e:.2e}, fine {eh:.2e}")
    assert e <= 1e-10 and eh <= 1e-10


@pytest.mark.parametrize("N,S", [(32, 48), (32, 64)])
def test_init_from_state_matches_trilinear_resize(cuda, hiplib, N, S):
    """init_hires_from_state equals torch's trilinear resize (align_corners=False, CPU), with
    0 / ambient / 0 in fine solids (solver.py:204-215)."""
    s = Solver3D(N, scalar_res=S)
    st = s.reset([SPHERE])
    data = random_state(N, 61, st.solid.cpu().numpy())
    st.data.copy_(to_dev(data))
    s.init_hires_from_state(st)
    got = s.hi.cpu().numpy()
    src = torch.from_numpy(data[[0, 4, 5]])[None]
    up = F.interpolate(src, size=(S,) * 3, mode="trilinear", align_corners=False)[0].numpy()
    fl = s.hi_solid.cpu().numpy() < 0.5
    want = np.stack([up[0] * fl, np.where(fl, up[1], AMBIENT_TEMP), up[2] * fl])
    for c in range(3):
        e = rel_err(got[c], want[c])
        _say(f"  channel {c}: {e:.2e}")
        assert e <= 1e-5


def test_step_profiled_reports_scalars(cuda, hiplib):
    s = Solver3D(32, scalar_res=48)
    st = s.reset([SPHERE])
    t = s.step_profiled(st, FIRE, DT)
    assert set(t) == {"advection", "scalars", "forces", "projection", "total"}
    assert all(v >= 0 for v in t.values())
    assert st.meta.step == 1 and "+scalars48" in st.meta.solver_version
    v = s.hires_volume()
    assert v.density.data_ptr() == s.hi[0].data_ptr() and v.temperature.shape == (48,) * 3


# ---------------------------------------------------------------- properties at size
def test_fine_mass_conserved_closed_box(cuda, hiplib):
    """Closed box, no sources, dissipation 0, a swirling flow: the fine fluid density total is
    kept by the fine mass ratio (selfcheck.py:263-279; gate there |drift| < 0.1 % in 1000 steps)."""
    N, S = 128, 256
    s = Solver3D(N, SolverConfig3D(density_dissipation=0.0), scalar_res=S)
    st = s.reset([Obstacle3D("sphere", (0.6, 0.6, 0.55), (0.1, 0.1, 0.1))])
    c = (torch.arange(S, device=DEV, dtype=torch.float32) + 0.5) / S
    zz, yy, xx = torch.meshgrid(c, c, c, indexing="ij")
    blob = torch.exp(-((xx - 0.35) ** 2 + (yy - 0.3) ** 2 + (zz - 0.45) ** 2) / 0.01) * (s.hi_solid < 0.5)
    s.hi[0] = blob
    s.hi[1] = AMBIENT_TEMP + 400.0 * blob
    c = (torch.arange(N, device=DEV, dtype=torch.float32) + 0.5) / N
    zz, yy, xx = torch.meshgrid(c, c, c, indexing="ij")
    st.data[1] = 0.5 * torch.sin(3.1 * yy)
    st.data[3] = 0.3 * torch.cos(2.7 * xx)
    s.project(st)
    m0 = float(s.hi[0].double().sum())
    for _ in range(40):
        s.step(st, Forces3D(), 1.0 / 30.0)
    m1 = float(s.hi[0].double().sum())
    _say(f"  fine mass drift over 40 steps: {m1 / m0 - 1:+.3e}")
    assert abs(m1 / m0 - 1.0) < 1e-4
    assert torch.isfinite(s.hi).all() and torch.isfinite(st.data).all()


def test_restriction_consistency_and_determinism(cuda, hiplib):
    """After a plume: the state's density is the 2x2x2 mean of the fine density on coarse fluid
    cells (selfcheck.py:282-291), and a second run from the same start is bit-identical
    (selfcheck.py:294-302)."""
    N, S = 128, 256
    plume = Forces3D(emitters=[Emitter3D(pos=(0.42, 0.1, 0.5), radius=0.06, density_rate=1.0, temperature=900.0,
                                         velocity=(0.0, 1.2, 0.0))], wind=(0.4, 0.0, 0.0))
    outs = []
    for run in range(2):
        s = Solver3D(N, scalar_res=S)
        st = s.reset([SPHERE])
        for _ in range(30):
            s.step(st, plume, DT)
        torch.cuda.synchronize()
        outs.append((st.data.clone(), s.hi.clone()))
        if run == 0:
            fl = st.solid < 0.5
            rd = F.avg_pool3d(s.hi[0][None, None].double(), 2)[0, 0]
            cd = st.data[0].double()
            err = float(((rd - cd) * fl).abs().max() / cd.abs().max())
            mrel = float(abs((rd * fl).sum() - (cd * fl).sum()) / (cd * fl).sum())
            _say(f"  restriction: max rel diff {err:.2e}, mass rel {mrel:.2e}, total {float(cd.sum()):.4f}")
            assert float(cd.sum()) > 0
            assert err < 1e-6 and mrel < 1e-6
    assert torch.equal(outs[0][0], outs[1][0]) and torch.equal(outs[0][1], outs[1][1])


def test_fine_emitter_footprint(cuda, hiplib):
    """One step from rest: the fine density is rate (1 - r^2/R^2)^2 dt at fine-cell centres; fine
    solids hold exactly 0 / ambient (selfcheck.py:305-323)."""
    N, S = 64, 128
    s = Solver3D(N, SolverConfig3D(density_dissipation=0.0), scalar_res=S)
    st = s.reset([SPHERE])
    e = Emitter3D(pos=(0.31, 0.2, 0.47), radius=0.06, density_rate=3.0)
    s.step(st, Forces3D(emitters=[e]), DT)
    c = (torch.arange(S, device=DEV, dtype=torch.float64) + 0.5) / S
    zz, yy, xx = torch.meshgrid(c, c, c, indexing="ij")
    r2 = ((xx - e.pos[0]) ** 2 + (yy - e.pos[1]) ** 2 + (zz - e.pos[2]) ** 2) / e.radius ** 2
    want = torch.where(r2 < 1, (1 - r2) ** 2, torch.zeros_like(r2)) * e.density_rate * DT
    err = float((s.hi[0].double() - want).abs().max() / want.max())
    m = s.hi_solid > 0.5
    clean = float(max(s.hi[0][m].abs().max(), (s.hi[1][m] - AMBIENT_TEMP).abs().max()))
    _say(f"  footprint rel err {err:.2e}, solid clean {clean:.1e}")
    assert err < 1e-5 and clean == 0.0

"""GPU tests of the paired-run guidance (csrc/guide.hip; reference vfx/datagen3d/guide3d.py and
sim.py:52-79) against the guidance oracle (oracle/guide_ref.py, pinned on the CPU by
tests/test_oracle_guide.py), and properties of run_pair.

Bars: the HIP low-pass is an FP32 DCT of the FP64 FFT formulation in the oracle, so fields agree
to 1e-5 of their max; the guided updates (exp, log of the low-passed fields) to 1e-5 as well.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

import guide_ref as G
from simhelp import rel_err, smooth_noise, sphere_solid, to_dev
from vfx.core.fields import AMBIENT_TEMP
from vfx.core.fields3d import Emitter3D, FieldMeta3D, FieldState3D, Forces3D, Obstacle3D
from vfx.datagen3d import DEFAULT_GUIDE, Guide3D, GuideConfig3D, lowpass, tracking_error
from vfx.datagen3d.sim import PairScenario, free_cache, run_pair
from vfx.sim3d import Solver3D, runtime_config

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _cfg(g: GuideConfig3D) -> dict:
    return dict(k_vel=g.k_vel, k_scalar=g.k_scalar, k_fuel=g.k_fuel, cutoff=g.cutoff, eps_density=g.eps_density,
                eps_temp=g.eps_temp, log_clamp=g.log_clamp)


def _state(N, seed, solid=None):
    d = np.zeros((6, N, N, N), np.float32)
    d[0] = np.abs(smooth_noise(N, 1, seed, amp=0.5))[0]
    d[1:4] = smooth_noise(N, 3, seed + 1, amp=1.0)
    d[4] = AMBIENT_TEMP + np.abs(smooth_noise(N, 1, seed + 2, amp=300.0))[0]
    d[5] = np.abs(smooth_noise(N, 1, seed + 3, amp=0.3))[0]
    if solid is not None:
        fl = solid < 0.5
        d[0] *= fl
        d[1:4] *= fl
        d[4] = np.where(fl, d[4], AMBIENT_TEMP)
        d[5] *= fl
    return np.ascontiguousarray(d)


def _fs(data, solid):
    N = data.shape[-1]
    return FieldState3D(data=to_dev(data), solid=to_dev(solid), meta=FieldMeta3D.for_res(N))


@pytest.mark.parametrize("n,cutoff", [(64, 16.0), (32, 6.0), (16, 3.0)])
def test_lowpass_matches_oracle(cuda, hiplib, n, cutoff):
    x = smooth_noise(n, 3, 3, coarse=12, amp=1.0) + 0.2 * np.random.default_rng(1).standard_normal((3, n, n, n))
    x = x.astype(np.float32)
    got = lowpass(to_dev(x), cutoff).cpu().numpy()
    ref = G.lowpass(x.astype(np.float64), cutoff)
    e = rel_err(got, ref)
    _say(f"  lowpass n={n} cutoff={cutoff}: max rel err {e:.2e}")
    assert e <= 1e-5
    # a cosine mode comes back scaled by the mask value at k/2 cycles per domain
    j = np.arange(n)
    k = n // 4
    mode = np.broadcast_to(np.cos(np.pi * k * (j + 0.5) / n)[None, :, None], (n, n, n)).astype(np.float32)
    got = lowpass(to_dev(np.ascontiguousarray(mode))[None], cutoff).cpu().numpy()[0]
    m = np.exp(-((k / 2) / cutoff) ** 8)
    assert np.abs(got - m * mode).max() <= 2e-5


@pytest.mark.parametrize("lr_res,f", [(32, 4), (64, 4), (32, 2)])
def test_velocity_and_scalars_match_oracle(cuda, hiplib, lr_res, f):
    H = lr_res * f
    _say(f"guide ops lr={lr_res} hr={H}")
    solid = sphere_solid(H)
    hr = _state(H, 11, solid)
    lr = _state(lr_res, 21)
    cfg = GuideConfig3D(k_fuel=0.4)
    g = Guide3D(cfg, to_dev(solid), lr_res)
    hs, ls = _fs(hr, solid), _fs(lr, np.zeros((lr_res,) * 3, np.float32))
    g.velocity(hs, ls)
    got_v = hs.data.cpu().numpy()
    ref_v = G.velocity(hr, solid, lr, _cfg(cfg))
    for c in range(1, 4):
        e = rel_err(got_v[c], ref_v[c])
        _say(f"  velocity {c}: {e:.2e}")
        assert e <= 1e-5
    assert np.array_equal(got_v[[0, 4, 5]], hr[[0, 4, 5]])
    g.scalars(hs, ls)
    got_s = hs.data.cpu().numpy()
    ref_s = G.scalars(got_v, solid, lr, _cfg(cfg))
    for c, name in ((0, "density"), (4, "T"), (5, "fuel")):
        e = rel_err(got_s[c], ref_s[c])
        _say(f"  {name}: {e:.2e}")
        assert e <= 1e-5
    m = solid > 0.5
    assert np.array_equal(got_s[:, m], got_v[:, m])


def test_tracking_error_matches_oracle(cuda, hiplib):
    hr = _state(128, 31)
    lr = _state(32, 41)
    got = tracking_error(to_dev(hr), to_dev(lr), 16).cpu().numpy()
    ref = G.tracking_error(hr, lr, 16)
    _say(f"  tracking {got} vs {ref}")
    np.testing.assert_allclose(got, ref, rtol=1e-5)
    got = tracking_error(to_dev(hr), to_dev(lr), 12).cpu().numpy()
    np.testing.assert_allclose(got, G.tracking_error(hr, lr, 12), rtol=1e-5)


def _plume(t_step):
    return Forces3D(emitters=[Emitter3D(pos=(0.42, 0.1, 0.5), radius=0.06, density_rate=2.0, temperature=900.0,
                                        velocity=(0.0, 1.2, 0.0))], wind=(0.4, 0.0, 0.0))


SCENE = PairScenario(_plume, num_steps=60, sim=dict(density_dissipation=1.2, temperature_cooling=5.0),
                     obstacles=[Obstacle3D("sphere", (0.5, 0.55, 0.5), (0.12, 0.12, 0.12))])


def test_run_pair_lr_is_the_runtime_solver(cuda, hiplib):
    """The LR run of run_pair is the runtime-config solver stepped on its own, bit for bit."""
    free_cache()
    steps = 10
    lrs = [lr.data.clone() for _, lr, _, _ in run_pair(SCENE, DEFAULT_GUIDE, hr_res=128, lr_res=32, steps=steps)]
    s = Solver3D(32, runtime_config(**SCENE.sim))
    st = s.reset(SCENE.obstacles)
    for k in range(steps):
        s.step(st, SCENE.forces(k), 1.0 / 30.0)
        assert torch.equal(st.data, lrs[k]), k


def test_guidance_keeps_hr_on_the_lr_run(cuda, hiplib):
    """64^3 LR and 256^3 HR over 60 steps: guided, the HR run tracks the LR one at 16^3 (the
    reference's default, guide3d.py / datagen3d README: worst 0.12 / 0.09 / 0.11 over 150 steps on
    its scenes); unguided it drifts well away."""
    free_cache()
    worst = {}
    for label, guide in (("guided", DEFAULT_GUIDE), ("unguided", None)):
        tr = []
        for s, lr, hr, t in run_pair(SCENE, guide, steps=60):
            if s >= 10:
                tr.append(t)
        tr = torch.stack(tr).cpu().numpy()
        assert np.isfinite(tr).all() and torch.isfinite(hr.data).all()
        worst[label] = tr.max(axis=0)
        _say(f"  {label}: worst tracking density / T / velocity {np.round(worst[label], 3).tolist()}")
    assert np.all(worst["guided"] < 0.25)
    assert np.all(worst["guided"] < 0.6 * worst["unguided"])
    free_cache()

"""Pins of the oracle's sources, diffusion, vorticity, smoothing and body forces
(reference vfx/sim3d/kernels.py:181-457) by closed forms of each term in isolation.

Each test switches every other term off (cooling, dissipation, burn, confinement, wind) so the
expected value follows from the reference's documented formula for one mechanism only.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from simhelp import smooth_noise, sphere_solid
from vfx.core.fields import AMBIENT_TEMP
from vfx.core.fields3d import Emitter3D, Forces3D, Impulse3D
from vfx.sim3d import SolverConfig3D
from vfx.sim3d.solver import pack_params

N = 16
DT = 1.0 / 30.0
QUIET = dict(temperature_cooling=0.0, density_dissipation=0.0, burn_rate=0.0, vorticity_eps=0.0)


def _centres(N):
    c = (np.arange(N, dtype=np.float64) + 0.5) / N
    zz, yy, xx = np.meshgrid(c, c, c, indexing="ij")
    return xx, yy, zz


def _state(T=AMBIENT_TEMP, F=0.0, d=0.0, v=(0.0, 0.0, 0.0)):
    s = np.zeros((6, N, N, N), np.float64)
    s[0], s[1], s[2], s[3], s[4], s[5] = d, v[0], v[1], v[2], T, F
    return s


# ---------------------------------------------------------------- Philox (noise generator)
def test_philox_known_answers():
    """Random123 known-answer vectors for philox4x32-10."""
    assert oracle.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert oracle.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert oracle.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


# ---------------------------------------------------------------- sources
def test_emitter_footprint_closed_form():
    """density += rate * (1 - r^2/R^2)^2 * dt inside the sphere (selfcheck.py:307-323 gate)."""
    em = Emitter3D(pos=(0.47, 0.52, 0.5), radius=0.21, density_rate=3.0, fuel_rate=0.5)
    prm = pack_params(SolverConfig3D(**QUIET), N, Forces3D(emitters=[em]), DT)
    out = oracle.sources(_state(), np.zeros((N,) * 3), prm, 1.0, prec="f64")
    xx, yy, zz = _centres(N)
    f32 = lambda v: float(np.float32(v))      # the parameter buffer holds float32 values
    r2 = ((xx - f32(0.47)) ** 2 + (yy - f32(0.52)) ** 2 + (zz - 0.5) ** 2) / f32(0.21) ** 2
    w = np.where(r2 < 1, (1 - r2) ** 2, 0.0)
    dtf = float(np.float32(DT))
    np.testing.assert_allclose(out[0], 3.0 * w * dtf, rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(out[5], 0.5 * w * dtf, rtol=1e-6, atol=1e-12)
    assert np.count_nonzero(out[0]) == np.count_nonzero(r2 < 1)


def test_emitter_temperature_target_and_velocity_blend():
    em = Emitter3D(pos=(0.5, 0.5, 0.5), radius=0.3, density_rate=0.0, temperature=900.0, velocity=(0.0, 2.0, -1.0))
    prm = pack_params(SolverConfig3D(**QUIET), N, Forces3D(emitters=[em]), DT)
    base = _state(T=AMBIENT_TEMP + 100.0, v=(1.0, 0.0, 0.0))
    out = oracle.sources(base, np.zeros((N,) * 3), prm, 1.0, prec="f64")
    xx, yy, zz = _centres(N)
    r2 = ((xx - 0.5) ** 2 + (yy - 0.5) ** 2 + (zz - 0.5) ** 2) / float(np.float32(0.3)) ** 2
    w = np.where(r2 < 1, (1 - r2) ** 2, 0.0)
    T = np.maximum(AMBIENT_TEMP + 100.0, np.where(r2 < 1, AMBIENT_TEMP + (900.0 - AMBIENT_TEMP) * w, -1))
    np.testing.assert_allclose(out[4], T, rtol=1e-12)
    np.testing.assert_allclose(out[1], 1.0 + (0.0 - 1.0) * w, atol=1e-12)
    np.testing.assert_allclose(out[2], 2.0 * w, atol=1e-12)
    np.testing.assert_allclose(out[3], -1.0 * w, atol=1e-12)


def test_impulse_gaussian_closed_form():
    imp = Impulse3D(pos=(0.3, 0.6, 0.5), radius=0.08, velocity=(1.0, -2.0, 0.5))
    prm = pack_params(SolverConfig3D(**QUIET), N, Forces3D(impulses=[imp]), DT)
    out = oracle.sources(_state(), np.zeros((N,) * 3), prm, 1.0, prec="f64")
    xx, yy, zz = _centres(N)
    f32 = lambda v: float(np.float32(v))
    r2 = ((xx - f32(0.3)) ** 2 + (yy - f32(0.6)) ** 2 + (zz - 0.5) ** 2) / f32(0.08) ** 2
    g = np.where(r2 < 9, np.exp(-r2), 0.0)
    for k, v in enumerate((1.0, -2.0, 0.5)):
        np.testing.assert_allclose(out[1 + k], g * v, rtol=1e-6, atol=1e-12)


@pytest.mark.parametrize("ramp", [100.0, 5.0])
def test_combustion_cooling_dissipation_closed_form(ramp):
    cfg = SolverConfig3D(burn_rate=8.0, ignition_temp=550.0, burn_heat=4000.0, smoke_from_fuel=0.35,
                         temperature_cooling=4.0, density_dissipation=0.05, max_temperature=3000.0,
                         ignition_ramp=ramp)
    prm = pack_params(cfg, N, Forces3D(), DT)
    T0 = np.linspace(500.0, 2900.0, N)[None, None, :] * np.ones((N, N, N))
    s = _state(F=0.8, d=0.2)
    s[4] = T0
    out = oracle.sources(s, np.zeros((N,) * 3), prm, 1.0, prec="f64")
    dt = float(np.float32(DT))
    ign = np.clip((T0 - 550.0) / ramp, 0, 1)
    burned = 0.8 * (1 - math.exp(-8.0 * dt)) * ign
    room = np.clip((3000.0 - T0) / (3000.0 - AMBIENT_TEMP), 0, 1)
    T = T0 + 4000.0 * burned * room
    d = 0.2 + 0.35 * burned
    T = AMBIENT_TEMP + (T - AMBIENT_TEMP) * math.exp(-4.0 * dt)
    d = d * math.exp(-0.05 * dt)
    # parameters travel as float32 (0.35, 0.05 are not exact): 1e-7 relative
    np.testing.assert_allclose(out[5], 0.8 - burned, rtol=1e-7)
    np.testing.assert_allclose(out[4], np.minimum(T, 3000.0), rtol=1e-7)
    np.testing.assert_allclose(out[0], d, rtol=1e-7)


def test_solid_cells_reset_and_mass_ratio_applied():
    solid = sphere_solid(N, r=0.3)
    s = _state(T=800.0, F=1.0, d=2.0, v=(1.0, 2.0, 3.0))
    prm = pack_params(SolverConfig3D(**QUIET), N, Forces3D(), DT)
    out = oracle.sources(s, solid, prm, 0.5, prec="f64")
    m = solid > 0.5
    assert np.all(out[0][m] == 0) and np.all(out[1:4][:, m] == 0) and np.all(out[5][m] == 0)
    assert np.all(out[4][m] == AMBIENT_TEMP)
    assert np.allclose(out[0][~m], 1.0) and np.allclose(out[5][~m], 1.0)


def test_emitter_noise_range_and_determinism():
    em = Emitter3D(pos=(0.5, 0.5, 0.5), radius=0.45, density_rate=1.0)
    cfg = SolverConfig3D(emitter_noise=0.2, noise_seed=3, **QUIET)
    xx, yy, zz = _centres(N)
    r2 = ((xx - 0.5) ** 2 + (yy - 0.5) ** 2 + (zz - 0.5) ** 2) / float(np.float32(0.45)) ** 2
    w = np.where(r2 < 1, (1 - r2) ** 2, 0.0)
    inside = w > 1e-3
    outs = []
    for step in (4, 4, 5):
        prm = pack_params(cfg, N, Forces3D(emitters=[em]), DT, step)
        outs.append(oracle.sources(_state(), np.zeros((N,) * 3), prm, 1.0, prec="f64")[0])
    nu = outs[0][inside] / (w[inside] * float(np.float32(DT)))
    assert nu.min() >= 0.8 - 1e-6 and nu.max() <= 1.2 + 1e-6
    assert abs(nu.mean() - 1.0) < 0.01 and nu.std() > 0.1        # uniform on [0.8, 1.2]: sd 0.115
    assert np.array_equal(outs[0], outs[1]) and not np.array_equal(outs[0], outs[2])
    # noise off reproduces the noiseless formula exactly
    prm = pack_params(SolverConfig3D(**QUIET), N, Forces3D(emitters=[em]), DT, 4)
    np.testing.assert_allclose(oracle.sources(_state(), np.zeros((N,) * 3), prm, 1.0, prec="f64")[0],
                               w * float(np.float32(DT)), rtol=1e-12, atol=1e-15)


# ---------------------------------------------------------------- diffusion
def test_diffusion_conserves_and_spike_stencil():
    solid = sphere_solid(N, r=0.2)
    cfg = SolverConfig3D(heat_diffusion=2e-4)
    prm = pack_params(cfg, N, Forces3D(), DT)
    s = _state()
    s[4] = AMBIENT_TEMP + 100.0 * np.abs(smooth_noise(N, 1, 4))[0] * (solid < 0.5)
    s[5] = np.abs(smooth_noise(N, 1, 5))[0] * (solid < 0.5)
    out = oracle.diffuse(s, solid, prm, 1.0, prec="f64")
    fl = solid < 0.5
    for c in (4, 5):
        assert abs(out[c][fl].sum() - s[c][fl].sum()) <= 1e-9 * abs(s[c][fl].sum())
    assert np.array_equal(out[:4], s[:4])
    # constant field is a fixed point
    s2 = _state(T=500.0)
    assert np.allclose(oracle.diffuse(s2, solid, prm, 1.0, prec="f64")[4], 500.0, rtol=0, atol=1e-12)
    # one spike in an open box: centre S (1 - 6 lam), the six neighbours lam S
    lam = min(float(np.float32(2e-4)) * float(np.float32(DT)) / (float(np.float32(1 / N)) ** 2), 0.16)
    s3 = _state(T=0.0)
    s3[4, 8, 8, 8] = 1.0
    o3 = oracle.diffuse(s3, np.zeros((N,) * 3), prm, 1.0, prec="f64")[4]
    assert abs(o3[8, 8, 8] - (1 - 6 * lam)) < 1e-6
    for nb in ((7, 8, 8), (9, 8, 8), (8, 7, 8), (8, 9, 8), (8, 8, 7), (8, 8, 9)):
        assert abs(o3[nb] - lam) < 1e-6
    # lambda is capped at 0.16
    prm_big = pack_params(SolverConfig3D(heat_diffusion=10.0), N, Forces3D(), DT)
    o4 = oracle.diffuse(s3, np.zeros((N,) * 3), prm_big, 1.0, prec="f64")[4]
    assert abs(o4[8, 8, 8] - (1 - 6 * 0.16)) < 1e-6


# ---------------------------------------------------------------- vorticity, smoothing, forces
def test_curl_of_linear_field():
    a, b, c = 0.7, -0.4, 1.1
    xx, yy, zz = _centres(N)
    s = _state()
    s[1], s[2], s[3] = a * yy, b * zz, c * xx        # omega = (-b, -c, -a)
    w = oracle.curl(s, np.zeros((N,) * 3), float(N) / 2.0, prec="f64")
    inner = (slice(1, -1),) * 3
    np.testing.assert_allclose(w[0][inner], -b, atol=1e-12)
    np.testing.assert_allclose(w[1][inner], -c, atol=1e-12)
    np.testing.assert_allclose(w[2][inner], -a, atol=1e-12)
    # at the x = 0 wall the missing neighbour counts as zero velocity
    np.testing.assert_allclose(w[2][1:-1, 1:-1, 0], (b * 0 - 0) * 0 + (s[2][1:-1, 1:-1, 1] - 0.0) * N / 2.0
                               - (s[1][1:-1, 2:, 0] - s[1][1:-1, :-2, 0]) * N / 2.0, atol=1e-12)


def _binom27_numpy(a, solid):
    """[1,2,1]^3 average over fluid cells, renormalised (scatter form with zero padding)."""
    N = solid.shape[0]
    w = (solid < 0.5).astype(np.float64)
    wp = np.pad(w, 1)
    ap = np.pad(a * w[None], ((0, 0), (1, 1), (1, 1), (1, 1)))
    num = np.zeros_like(a, dtype=np.float64)
    den = np.zeros((N,) * 3)
    for k in (-1, 0, 1):
        for j in (-1, 0, 1):
            for i in (-1, 0, 1):
                q = (2 - abs(i)) * (2 - abs(j)) * (2 - abs(k))
                sl = (slice(1 + k, 1 + k + N), slice(1 + j, 1 + j + N), slice(1 + i, 1 + i + N))
                num += q * ap[(slice(None),) + sl]
                den += q * wp[sl]
    return np.where(den > 0, num / np.where(den > 0, den, 1), 0.0)


def test_separable_smoothing_equals_27_tap():
    solid = sphere_solid(N, r=0.25)
    a = smooth_noise(N, 3, 6).astype(np.float64)
    sep = oracle.smooth3(a, solid, mode=0, prec="f64")
    tap = oracle.binom3(a, solid, prec="f64")
    ref = _binom27_numpy(a, solid)
    np.testing.assert_allclose(sep, ref, atol=1e-12)
    np.testing.assert_allclose(tap, ref, atol=1e-12)
    # mode 1 adds dt * smoothed to the velocity of fluid cells only
    st = np.zeros((6, N, N, N))
    o = oracle.smooth3(a, solid, mode=1, dt_=0.25, out=st, prec="f64")
    fl = solid < 0.5
    np.testing.assert_allclose(o[1:4][:, fl], 0.25 * ref[:, fl], atol=1e-12)
    assert np.all(o[1:4][:, ~fl] == 0)


def test_buoyancy_and_wind_closed_form():
    cfg = SolverConfig3D(buoyancy_alpha=0.6, buoyancy_beta=0.004, vorticity_eps=0.0, wind_marker_temp=300.0)
    prm = pack_params(cfg, N, Forces3D(wind=(0.5, 0.0, -0.25)), DT)
    s = _state()
    s[0] = np.abs(smooth_noise(N, 1, 1))[0]
    s[4] = AMBIENT_TEMP + 200.0 * np.abs(smooth_noise(N, 1, 2))[0]
    s[5] = 0.1
    f = oracle.body_forces(s, np.zeros((N,) * 3), np.zeros((3, N, N, N)), prm, prec="f64")
    m = np.clip(s[0] + s[5] + (s[4] - AMBIENT_TEMP) / 300.0, 0, 1)
    np.testing.assert_allclose(f[1], -float(np.float32(0.6)) * s[0] + float(np.float32(0.004)) * (s[4] - AMBIENT_TEMP),
                               rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(f[0], 0.5 * m, rtol=1e-12)
    np.testing.assert_allclose(f[2], -0.25 * m, rtol=1e-12)


def test_confinement_closed_form():
    """omega_s = (0, 0, w(x)) with |omega| linear in x: grad|omega| = (w', 0, 0) exactly in the
    interior, so f = eps dx / sqrt(w'^2 + (floor w)^2) * (g x omega) = (0, -k w' w, 0)."""
    eps, floor = 6.0, 0.5
    cfg = SolverConfig3D(buoyancy_alpha=0.0, buoyancy_beta=0.0, vorticity_eps=eps, confinement_floor=floor)
    prm = pack_params(cfg, N, Forces3D(), DT)
    xi = np.arange(N, dtype=np.float64)
    w = 1.0 + 0.25 * xi
    om = np.zeros((3, N, N, N))
    om[2] = w[None, None, :]
    f = oracle.body_forces(_state(), np.zeros((N,) * 3), om, prm, prec="f64")
    g = 0.25                       # index-space gradient of |omega| along x
    mag = np.sqrt(g ** 2 + (floor * w) ** 2) + 1e-20
    k = eps * float(np.float32(1.0 / N)) / mag
    inner = (slice(None), slice(None), slice(1, -1))
    np.testing.assert_allclose(f[1][inner], (-k * g * w)[None, None, 1:-1] * np.ones((N, N, 1)), rtol=1e-9)
    assert np.abs(f[0]).max() == 0 and np.abs(f[2]).max() == 0
    # one-sided difference at the walls: (|w|(1) - |w|(0)) / 1
    np.testing.assert_allclose(f[1][:, :, 0], -(eps * float(np.float32(1 / N)) / mag[0]) * 0.25 * w[0], rtol=1e-9)

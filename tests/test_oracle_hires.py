"""Pins of the oracle's finer scalar grid (reference vfx/sim3d/hires.py, solver.py:217-264) by
checks that do not restate its kernels:

  * at S == N the fine grid is the simulation grid (f2c is the identity, fine dx = dx): fine
    advection, sources and diffusion equal the already pinned coarse functions bit for bit on the
    scalar channels;
  * fine advection at S = 2N: zero velocity is the identity, a uniform velocity of one coarse cell
    shifts by exactly two fine cells, a linear field under a uniform velocity is exact (trilinear
    interpolation reproduces it), and at a non-integer ratio the semi-Lagrangian values equal
    scipy's trilinear map_coordinates at the analytically traced points;
  * fine sources: the emitter footprint (1 - r^2/R^2)^2 rate dt at fine-cell centres (the
    reference's hires_raster gate, selfcheck.py:302-323), fine solids hold 0 / ambient / 0;
  * restriction: the block mean at integer ratios, an independent float64 overlap matrix at
    non-integer ratios, conservation of the fine total, coarse solids;
  * the fine stage conserves the fine fluid mass with sources and dissipation off (the mass ratio;
    the reference's hires_mass gate, selfcheck.py:263-279).
"""
from __future__ import annotations

import numpy as np
import pytest
from scipy import ndimage

import oracle
from simhelp import smooth_noise, sphere_solid
from vfx.core.fields import AMBIENT_TEMP
from vfx.core.fields3d import Emitter3D, Forces3D
from vfx.sim3d import SolverConfig3D
from vfx.sim3d.solver import pack_params

DT = 1.0 / 30.0
QUIET = dict(temperature_cooling=0.0, density_dissipation=0.0, burn_rate=0.0, vorticity_eps=0.0)


def _scalars(S, seed, dtype=np.float32):
    f = smooth_noise(S, 3, seed, amp=1.0)
    out = np.empty((3, S, S, S), np.float64)
    out[0] = np.abs(f[0])
    out[1] = AMBIENT_TEMP + 300.0 * np.abs(f[1])
    out[2] = np.abs(f[2])
    return np.ascontiguousarray(out.astype(dtype))


# ---------------------------------------------------------------- advection
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_same_size_equals_coarse_advection(prec):
    N = 20
    dt = np.float32 if prec == "f32" else np.float64
    src = _scalars(N, 3, dt)
    vel = smooth_noise(N, 3, 4, amp=1.5).astype(dt)
    s = 0.9
    fine = oracle.hires_advect(src, vel, s, prec=prec)
    coarse = oracle.advect(src, vel, s, prec=prec)
    assert np.array_equal(fine, coarse)


def test_zero_velocity_is_identity():
    N, S = 12, 24
    src = _scalars(S, 5)
    out = oracle.hires_advect(src, np.zeros((3, N, N, N), np.float32), 1.0)
    assert np.array_equal(out, src)


@pytest.mark.parametrize("axis,sign", [(0, 1.0), (1, -1.0), (2, 1.0)])
def test_uniform_velocity_shifts_two_fine_cells(axis, sign):
    """One coarse cell per step = two fine cells at S = 2N (exact in float: the traced
    coordinates are quarter-integers)."""
    N, S = 12, 24
    src = _scalars(S, 6)
    vel = np.zeros((3, N, N, N), np.float32)
    vel[axis] = sign
    out = oracle.hires_advect(src, vel, 1.0)
    # component `axis` of the velocity moves array axis 2 - axis (x is the last index)
    ax = 2 - axis
    want = np.roll(src, int(2 * sign), axis=1 + ax)
    sl = [slice(None)] * 4
    sl[1 + ax] = slice(4, S - 4)
    assert np.array_equal(out[tuple(sl)], want[tuple(sl)])


def test_linear_field_uniform_velocity_exact():
    N, S = 16, 40      # ratio 2.5
    u = np.array([0.31, -0.22, 0.17])      # coarse cells per unit s
    s = 1.3
    vel = np.empty((3, N, N, N), np.float64)
    for k in range(3):
        vel[k] = u[k]
    g = np.arange(S, dtype=np.float64)
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
    b = np.array([0.05, -0.03, 0.02])
    lin = lambda x, y, z: 1.0 + b[0] * x + b[1] * y + b[2] * z
    f = lin(xx, yy, zz)
    src = np.stack([f, 2.0 * f, f + 3.0])
    out = oracle.hires_advect(src, vel, s, prec="f64")
    d = s * u * S / N          # the displacement in fine cells
    want = lin(xx - d[0], yy - d[1], zz - d[2])
    inner = (slice(5, S - 5),) * 3
    for c, m in enumerate((want, 2.0 * want, want + 3.0)):
        np.testing.assert_allclose(out[c][inner], m[inner], rtol=0, atol=1e-12)


def test_non_integer_ratio_matches_map_coordinates():
    """S / N = 1.5, uniform velocity: the backward point of fine cell x is x - s u S / N; the
    semi-Lagrangian and MacCormack values from scipy's trilinear sampler (clamped to the box)."""
    N, S = 16, 24
    u = np.array([0.43, -0.29, 0.21])
    s = 1.1
    vel = np.empty((3, N, N, N), np.float64)
    for k in range(3):
        vel[k] = u[k]
    src = _scalars(S, 7, np.float64)
    out = oracle.hires_advect(src, vel, s, prec="f64")
    d = s * u * S / N
    g = np.arange(S, dtype=np.float64)
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")

    def sample(f, x, y, z):
        c = [np.clip(z, 0, S - 1), np.clip(y, 0, S - 1), np.clip(x, 0, S - 1)]
        return ndimage.map_coordinates(f, c, order=1, mode="nearest")

    for ch in range(3):
        hat = sample(src[ch], xx - d[0], yy - d[1], zz - d[2])
        bar = sample(hat, xx + d[0], yy + d[1], zz + d[2])      # hat at the forward point
        val = hat + 0.5 * (src[ch] - bar)
        q0 = [np.clip(np.floor(np.clip(p, 0, S - 1)), 0, S - 2).astype(int) for p in (xx - d[0], yy - d[1], zz - d[2])]
        corners = [src[ch][q0[2] + k, q0[1] + j, q0[0] + i] for i in (0, 1) for j in (0, 1) for k in (0, 1)]
        lo, hi = np.min(corners, axis=0), np.max(corners, axis=0)
        want = np.where((val < lo) | (val > hi), hat, val)
        np.testing.assert_allclose(out[ch], want, rtol=0, atol=1e-10)


# ---------------------------------------------------------------- sources and diffusion
def _scene_prm(N, **cfg):
    ems = [Emitter3D(pos=(0.45, 0.3, 0.5), radius=0.2, density_rate=2.0, temperature=1100.0, fuel_rate=1.5),
           Emitter3D(pos=(0.6, 0.55, 0.4), radius=0.15, density_rate=0.5, temperature=0.0, fuel_rate=0.8)]
    return pack_params(SolverConfig3D(**cfg), N, Forces3D(emitters=ems), DT, step=3)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("noise", [0.0, 0.4])
def test_same_size_sources_equal_coarse(prec, noise):
    N = 16
    dt = np.float32 if prec == "f32" else np.float64
    solid = sphere_solid(N).astype(dt)
    sc = _scalars(N, 8, dt)
    sc[1] += 500.0      # above ignition: the combustion terms are active
    adv6 = np.zeros((6, N, N, N), dt)
    adv6[0], adv6[4], adv6[5] = sc
    prm = _scene_prm(N, emitter_noise=noise, noise_seed=5)
    fine = oracle.hires_sources(sc, solid, prm, 0.97, prec=prec)
    coarse = oracle.sources(adv6, solid, prm, 0.97, prec=prec)
    assert np.array_equal(fine, coarse[[0, 4, 5]])


def test_fine_emitter_footprint_closed_form():
    N, S = 16, 32
    em = Emitter3D(pos=(0.31, 0.2, 0.47), radius=0.12, density_rate=3.0, fuel_rate=0.0)
    prm = pack_params(SolverConfig3D(**QUIET), N, Forces3D(emitters=[em]), DT)
    hsolid = sphere_solid(S)
    adv = np.zeros((3, S, S, S), np.float64)
    adv[1] = AMBIENT_TEMP
    out = oracle.hires_sources(adv, hsolid, prm, 1.0, prec="f64")
    c = (np.arange(S, dtype=np.float64) + 0.5) / S
    zz, yy, xx = np.meshgrid(c, c, c, indexing="ij")
    f32 = lambda v: float(np.float32(v))
    r2 = ((xx - f32(0.31)) ** 2 + (yy - f32(0.2)) ** 2 + (zz - f32(0.47)) ** 2) / f32(0.12) ** 2
    want = np.where(r2 < 1, (1 - r2) ** 2, 0.0) * 3.0 * f32(DT)
    want[hsolid > 0.5] = 0.0
    np.testing.assert_allclose(out[0], want, rtol=1e-6, atol=1e-12)
    m = hsolid > 0.5
    assert np.all(out[0][m] == 0) and np.all(out[1][m] == np.float32(AMBIENT_TEMP)) and np.all(out[2][m] == 0)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_same_size_diffusion_equals_coarse(prec):
    N = 16
    dt = np.float32 if prec == "f32" else np.float64
    solid = sphere_solid(N).astype(dt)
    sc = _scalars(N, 9, dt)
    src6 = np.zeros((6, N, N, N), dt)
    src6[0], src6[4], src6[5] = sc
    prm = pack_params(SolverConfig3D(heat_diffusion=2e-4), N, Forces3D(), DT)
    fine = oracle.hires_diffuse(sc, solid, prm, prec=prec)
    coarse = oracle.diffuse(src6, solid, prm, prec=prec)
    assert np.array_equal(fine, coarse[[0, 4, 5]])


def test_fine_diffusion_conserves_and_spike():
    S = 24
    prm = pack_params(SolverConfig3D(heat_diffusion=1e-4), 12, Forces3D(), DT)
    src = np.zeros((3, S, S, S))
    src[1, 10, 11, 12] = 1.0
    out = oracle.hires_diffuse(src, np.zeros((S,) * 3), prm, prec="f64")
    lam = min(float(np.float32(1e-4)) * float(np.float32(DT)) * S * S, float(np.float32(0.16)))
    assert abs(out[1].sum() - 1.0) < 1e-12
    assert abs(out[1, 10, 11, 12] - (1 - 6 * lam)) < 1e-12
    for dz, dy, dx in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        assert abs(out[1, 10 + dz, 11 + dy, 12 + dx] - lam) < 1e-12


# ---------------------------------------------------------------- restriction
def _overlap_matrix(N, S):
    """W[i, j] = |[i S/N, (i+1) S/N) intersected with [j, j+1)| in float64 (fine-cell units)."""
    r = S / N
    W = np.zeros((N, S))
    for i in range(N):
        for j in range(S):
            W[i, j] = max(0.0, min(j + 1, (i + 1) * r) - max(j, i * r))
    return W


@pytest.mark.parametrize("N,S", [(12, 24), (12, 36), (16, 24), (10, 17), (12, 12)])
def test_restriction_overlap_mean(N, S):
    fine = _scalars(S, 10, np.float64)
    csolid = np.zeros((N, N, N))
    prm = pack_params(SolverConfig3D(), N, Forces3D(), DT)
    data = np.full((6, N, N, N), -7.0)
    out = oracle.hires_restrict(fine, csolid, prm, data, prec="f64")
    W = _overlap_matrix(N, S)
    Wn = W / W.sum(axis=1, keepdims=True)
    for c, ch in enumerate((0, 4, 5)):
        want = np.einsum("ai,bj,ck,ijk->abc", Wn, Wn, Wn, fine[c])
        np.testing.assert_allclose(out[ch], want, rtol=1e-12, atol=1e-12)
    assert np.all(out[1:4] == -7.0)
    # conservation: every fine cell is covered exactly once
    vol = (S / N) ** 3
    assert abs(out[0].sum() * vol - fine[0].sum()) < 1e-9 * fine[0].sum()
    if S % N == 0:
        r = S // N
        blk = fine[0].reshape(N, r, N, r, N, r).mean(axis=(1, 3, 5))
        np.testing.assert_allclose(out[0], blk, rtol=1e-13)


def test_restriction_coarse_solids_and_f32():
    N, S = 16, 32
    fine = _scalars(S, 11)
    csolid = sphere_solid(N)
    prm = pack_params(SolverConfig3D(), N, Forces3D(), DT)
    out = oracle.hires_restrict(fine, csolid, prm, np.zeros((6, N, N, N), np.float32))
    m = csolid > 0.5
    assert np.all(out[0][m] == 0) and np.all(out[4][m] == np.float32(AMBIENT_TEMP)) and np.all(out[5][m] == 0)
    blk = fine.astype(np.float64).reshape(3, N, 2, N, 2, N, 2).mean(axis=(2, 4, 6))
    for c, ch in enumerate((0, 4, 5)):
        np.testing.assert_allclose(out[ch][~m], blk[c][~m], rtol=1e-6)


# ---------------------------------------------------------------- the fine stage
def test_stage_is_the_composition():
    N, S = 12, 20
    hsolid = sphere_solid(S)
    hi = _scalars(S, 12)
    vel = smooth_noise(N, 3, 13, amp=1.0)
    prm = _scene_prm(N, heat_diffusion=1e-4)
    inv_dx = np.float32(N)
    s = np.float32(prm[0]) * inv_dx
    adv = oracle.hires_advect(hi, vel, s)
    ratio = oracle.mass_ratio(hi[0], adv[0], hsolid)
    src = oracle.hires_sources(adv, hsolid, prm, ratio)
    want = oracle.hires_diffuse(src, hsolid, prm)
    got = hi.copy()
    r2 = oracle.hires_stage(got, hsolid, vel, prm, inv_dx)
    assert r2 == ratio
    assert np.array_equal(got, want)


def test_stage_conserves_fine_mass():
    """Closed box, no emitters, dissipation 0: the fine fluid density total is kept by the mass
    ratio (selfcheck.py:263-279 gate: |drift| < 0.1 % over 1000 steps)."""
    N, S = 12, 24
    hsolid = sphere_solid(S)
    hi = _scalars(S, 14, np.float64)
    hi[:, hsolid > 0.5] = 0.0
    vel = smooth_noise(N, 3, 15, amp=2.0)
    prm = pack_params(SolverConfig3D(**QUIET), N, Forces3D(), DT)
    m0 = hi[0][hsolid < 0.5].sum()
    for _ in range(5):
        oracle.hires_stage(hi, hsolid, vel, prm, float(N), prec="f64")
    m1 = hi[0][hsolid < 0.5].sum()
    assert abs(m1 / m0 - 1.0) < 1e-12
    assert np.all(hi[0] >= 0)

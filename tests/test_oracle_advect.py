"""Pins of the oracle's advection and mass correction (reference vfx/sim3d/kernels.py:66-178,
solver.py:40-76) by closed forms, not by restating the kernel:

  * affine velocity + linear field: trilinear interpolation is exact for both, so the RK2
    midpoint backtrace, the SL value and the MacCormack value have closed forms (and the 8-corner
    clamp of a linear field is the min / max over the corners of the traced cell);
  * integer-cell shifts are exact in the interior (tests3d/test_sim3d_layer1.py:199-216);
  * a constant field is preserved; the result never leaves [min, max] of the source;
  * C channels at once equal one channel at a time (test_sim3d_layer1.py:248-255);
  * the mass ratio equals the numpy ratio of fluid sums.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from simhelp import smooth_noise, sphere_solid


def _grid(N):
    z, y, x = np.meshgrid(*(np.arange(N, dtype=np.float64),) * 3, indexing="ij")
    return x, y, z


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_affine_velocity_linear_field_closed_form(prec):
    N = 24
    x, y, z = _grid(N)
    A = np.array([[0.02, -0.03, 0.01], [0.04, 0.01, -0.02], [-0.01, 0.02, 0.03]])
    c = np.array([0.3, -0.2, 0.25])
    def vel_at(px, py, pz):
        return [A[k, 0] * px + A[k, 1] * py + A[k, 2] * pz + c[k] for k in range(3)]
    vel = np.stack(vel_at(x, y, z))
    b = np.array([0.7, -1.3, 0.4])
    f = 2.0 + b[0] * x + b[1] * y + b[2] * z
    s = 1.7
    out = oracle.advect(f[None], vel, s, prec=prec)[0]

    def trace(px, py, pz, ss):
        v1 = vel_at(px, py, pz)
        v2 = vel_at(px + 0.5 * ss * v1[0], py + 0.5 * ss * v1[1], pz + 0.5 * ss * v1[2])
        return px + ss * v2[0], py + ss * v2[1], pz + ss * v2[2]

    lin = lambda px, py, pz: 2.0 + b[0] * px + b[1] * py + b[2] * pz
    q = trace(x, y, z, -s)
    hat = lin(*q)
    r = trace(x, y, z, s)
    bar = lin(*trace(*r, -s))
    val = hat + 0.5 * (f - bar)
    # clamp: min / max of the linear field over the corners of q's cell
    q0 = [np.floor(qq) for qq in q]
    corners = [lin(q0[0] + i, q0[1] + j, q0[2] + k) for i in (0, 1) for j in (0, 1) for k in (0, 1)]
    lo, hi = np.min(corners, axis=0), np.max(corners, axis=0)
    expect = np.where((val < lo) | (val > hi), hat, val)
    # interior: every trace, and the backtraces of the grid nodes around the forward point, stay
    # inside the box (the displacement here is below 1.5 cells, hence the 3-cell margin)
    pts = [q, r, trace(*r, -s)]
    ok = np.ones_like(f, dtype=bool)
    for p in pts:
        for comp in p:
            ok &= (comp > 3.0) & (comp < N - 4.0)
    assert ok.sum() > 0.2 * ok.size
    tol = 2e-4 if prec == "f32" else 1e-9
    err = np.abs(out - expect)[ok].max()
    assert err < tol, err
    # and the MacCormack correction is active (not everything clamped back to the SL value); for
    # an affine field it is the O(s^3) defect of the back-and-forth trace
    if prec == "f64":
        assert np.mean(np.abs(expect - hat)[ok] > 5e-6) > 0.1
        assert np.abs(out - hat)[ok].max() > 1e-5


def test_integer_shift_exact():
    N = 32
    f = smooth_noise(N, 1, 3)[0]
    vel = np.zeros((3, N, N, N), np.float32)
    vel[0], vel[1], vel[2] = 2.0, 1.0, -1.0
    out = oracle.advect(f[None], vel, 1.0)[0]
    # out(z, y, x) = f(z + 1, y - 1, x - 2) away from the upstream and downstream walls
    expect = f[2:, :-2, :-4]
    got = out[1:-1, 1:-1, 2:-2]
    assert np.array_equal(got, expect)


def test_constant_preserved_and_bounded():
    N = 24
    vel = smooth_noise(N, 3, 7, amp=4.0)
    c = np.float32(3.75)
    out = oracle.advect(np.full((1, N, N, N), c, np.float32), vel, 1.0)[0]
    assert np.abs(out - c).max() <= 4 * np.spacing(c)
    f = smooth_noise(N, 2, 8)
    out = oracle.advect(f, vel, 1.3)
    for k in range(2):
        assert out[k].min() >= f[k].min() and out[k].max() <= f[k].max()


def test_multichannel_equals_per_channel():
    N = 20
    f = smooth_noise(N, 3, 9)
    vel = smooth_noise(N, 3, 10, amp=2.0)
    all3 = oracle.advect(f, vel, 0.9)
    for k in range(3):
        assert np.array_equal(all3[k], oracle.advect(f[k:k + 1], vel, 0.9)[0])


def test_mass_ratio_closed_form():
    N = 16
    solid = sphere_solid(N, r=0.25)
    pre = np.abs(smooth_noise(N, 1, 1))[0]
    post = np.abs(smooth_noise(N, 1, 2))[0]
    fl = solid < 0.5
    expect = pre[fl].astype(np.float64).sum() / post[fl].astype(np.float64).sum()
    assert abs(oracle.mass_ratio(pre, post, solid) - expect) <= 1e-12 * expect
    assert oracle.mass_ratio(np.zeros_like(pre), post, solid) == 1.0
    assert oracle.mass_ratio(pre, np.zeros_like(post), solid) == 1.0

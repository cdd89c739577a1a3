"""Pins of the guidance oracle (oracle/guide_ref.py, reference vfx/datagen3d/guide3d.py) by closed
forms and a library routine:

  * the mirror-extended spectral low-pass passes a constant unchanged and scales the separable
    cosine mode cos(pi k (j + 1/2) / n) by exactly the mask value at k/2 cycles per domain (the
    half-sample even extension makes the DCT-II modes its eigenfunctions), in any direction, for
    products of modes along the three axes, and is the identity for a huge cutoff;
  * the half-pixel trilinear upsampling equals torch's F.interpolate(trilinear, align_corners=False);
  * consistent fields (LR = pool(HR)) leave the HR state unchanged; uniform fields give the
    closed-form factor exp(k clamp(log ratio)); solid cells are untouched; the density and
    temperature floors hold;
  * tracking_error of uniform fields is the closed-form relative error.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch
import torch.nn.functional as F

import guide_ref as G
from simhelp import smooth_noise, sphere_solid

CFG = dict(k_vel=1.0, k_scalar=0.7, k_fuel=0.0, cutoff=16.0, eps_density=0.01, eps_temp=5.0, log_clamp=1.0)


def _mode(n, k, axis):
    j = np.arange(n, dtype=np.float64)
    c = np.cos(np.pi * k * (j + 0.5) / n)
    shape = [1, 1, 1]
    shape[axis] = n
    return np.broadcast_to(c.reshape(shape), (n, n, n))


def test_lowpass_constant_and_identity():
    n = 16
    x = np.full((2, n, n, n), 3.25)
    np.testing.assert_allclose(G.lowpass(x, 4.0), x, rtol=1e-13)
    y = smooth_noise(n, 2, 3).astype(np.float64)
    np.testing.assert_allclose(G.lowpass(y, 1e6), y, rtol=0, atol=1e-12)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("k", [1, 5, 9, 15])
def test_lowpass_scales_cosine_modes_by_the_mask(axis, k):
    n, cutoff = 16, 3.0
    x = _mode(n, k, axis)[None]
    want = math.exp(-((k / 2.0) / cutoff) ** 8) * x
    np.testing.assert_allclose(G.lowpass(x, cutoff), want, rtol=0, atol=1e-12)


def test_lowpass_product_mode_radial_mask():
    n, cutoff = 16, 4.0
    kx, ky, kz = 3, 6, 5
    x = (_mode(n, kx, 2) * _mode(n, ky, 1) * _mode(n, kz, 0))[None]
    m = math.exp(-(math.sqrt((kx / 2) ** 2 + (ky / 2) ** 2 + (kz / 2) ** 2) / cutoff) ** 8)
    np.testing.assert_allclose(G.lowpass(x, cutoff), m * x, rtol=0, atol=1e-12)


@pytest.mark.parametrize("f", [2, 4])
def test_up_equals_torch_trilinear(f):
    n = 8
    x = smooth_noise(n, 3, 5).astype(np.float64)
    want = F.interpolate(torch.from_numpy(x)[None], scale_factor=f, mode="trilinear", align_corners=False)[0].numpy()
    np.testing.assert_allclose(G.up(x, f), want, rtol=0, atol=1e-12)


def _pair(n, f, seed):
    H = n * f
    hr = np.zeros((6, H, H, H))
    hr[0] = np.abs(smooth_noise(H, 1, seed))[0]
    hr[1:4] = smooth_noise(H, 3, seed + 1)
    hr[4] = G.AMBIENT_TEMP + 200.0 * np.abs(smooth_noise(H, 1, seed + 2))[0]
    hr[5] = np.abs(smooth_noise(H, 1, seed + 3))[0]
    return hr


def test_consistent_fields_leave_hr_unchanged():
    n, f = 8, 2
    hr = _pair(n, f, 7)
    lr = G.pool(hr, f)
    solid = np.zeros((n * f,) * 3)
    v = G.velocity(hr, solid, lr, CFG)
    np.testing.assert_allclose(v, hr, rtol=0, atol=1e-12)
    s = G.scalars(hr, solid, lr, dict(CFG, k_fuel=0.5))
    np.testing.assert_allclose(s, hr, rtol=1e-12, atol=1e-12)


def test_uniform_fields_closed_form_and_solids():
    n, f = 8, 4
    H = n * f
    hr = np.zeros((6, H, H, H))
    hr[0], hr[4], hr[5] = 0.2, G.AMBIENT_TEMP + 50.0, 0.1
    hr[1] = 0.3
    lr = np.zeros((6, n, n, n))
    lr[0], lr[4], lr[5] = 0.5, G.AMBIENT_TEMP + 20.0, 0.4
    lr[1] = -0.1
    solid = sphere_solid(H).astype(np.float64)
    fl = solid < 0.5
    v = G.velocity(hr, solid, lr, CFG)
    np.testing.assert_allclose(v[1][fl], 0.3 + 1.0 * (-0.1 - 0.3), rtol=1e-12)
    assert np.all(v[1][~fl] == 0.3)
    s = G.scalars(hr, solid, lr, dict(CFG, k_fuel=0.5))
    gd = math.exp(0.7 * max(-1.0, min(1.0, math.log((0.5 + 0.01) / (0.2 + 0.01)))))
    gt = math.exp(0.7 * max(-1.0, min(1.0, math.log((20.0 + 5.0) / (50.0 + 5.0)))))
    np.testing.assert_allclose(s[0][fl], 0.2 * gd, rtol=1e-12)
    np.testing.assert_allclose(s[4][fl], 50.0 * gt + G.AMBIENT_TEMP, rtol=1e-12)
    np.testing.assert_allclose(s[5][fl], 0.1 + 0.5 * (0.4 - 0.1), rtol=1e-12)
    assert np.all(s[0][~fl] == 0.2) and np.all(s[4][~fl] == hr[4][~fl]) and np.all(s[5][~fl] == 0.1)


def test_floors():
    """density never below 0; temperature never below min(T_old, ambient)."""
    n, f = 8, 2
    H = n * f
    hr = _pair(n, f, 11)
    hr[4, : H // 2] = G.AMBIENT_TEMP - 30.0      # colder than ambient: the floor is T_old
    lr = G.pool(hr, f)
    lr[0] *= 0.0
    lr[4] = G.AMBIENT_TEMP
    s = G.scalars(hr, np.zeros((H,) * 3), lr, CFG)
    assert np.all(s[0] >= 0)
    assert np.all(s[4] >= np.minimum(hr[4], G.AMBIENT_TEMP) - 1e-9)
    # the LR run is at ambient everywhere: T - ambient shrinks toward 0 from either side
    assert np.all(np.abs(s[4] - G.AMBIENT_TEMP) <= np.abs(hr[4] - G.AMBIENT_TEMP) + 1e-9)


def test_tracking_error_uniform_closed_form():
    H, n = 32, 16
    hr = np.zeros((6, H, H, H))
    lr = np.zeros((6, n, n, n))
    hr[0], lr[0] = 0.3, 0.2
    hr[4], lr[4] = G.AMBIENT_TEMP + 10.0, G.AMBIENT_TEMP + 8.0
    hr[1], lr[1] = 1.0, 0.5
    hr[3], lr[3] = -1.0, 0.0
    e = G.tracking_error(hr, lr, 8)
    np.testing.assert_allclose(e[0], 0.1 / 0.2, rtol=1e-7)
    np.testing.assert_allclose(e[1], 2.0 / 8.0, rtol=1e-7)
    np.testing.assert_allclose(e[2], math.sqrt(0.25 + 1.0) / 0.5, rtol=1e-7)
    # adaptive pooling with a non-divisible size equals torch's
    x = smooth_noise(20, 2, 3).astype(np.float64)
    want = F.adaptive_avg_pool3d(torch.from_numpy(x)[None], 6)[0].numpy()
    np.testing.assert_allclose(G.adaptive_pool(x, 6), want, rtol=1e-12)

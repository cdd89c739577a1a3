"""Pins of the renderer oracle (oracle/render_ref.py + render_oracle.c; reference vfx/render3d) by
closed forms of the physics and the geometry:

  * empty-brick skipping changes nothing beyond single-sample threshold flips (against the march
    that samples every step);
  * a uniform non-scattering medium attenuates a ray by Beer-Lambert, exp(-k_abs d L) over the
    ray's chord L through the unit cube (to one step);
  * an optically thin hot medium emits the blackbody table's colour times the chord length;
  * an opaque sphere hides exactly the rays that pass within its radius (to the smoothing width);
  * the light volume of a uniform medium lit from straight above holds exp(-k_shadow d (1 - y))
    and the sky occlusion exp(-k_amb d * the in-cube length of its 6 upward steps);
  * the blackbody table: the fitted colour-matching functions against the CIE 1931 tabulated values,
    warmer temperatures bluer and brighter.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import render_ref as R
from simhelp import smooth_noise
from vfx.core.fields3d import Camera
from vfx.render.blackbody import blackbody_lut, cie_xyz_bar
from vfx.render3d.renderer import RenderParams3D, camera_basis

CAM = Camera(position=(0.5, 0.5, -1.6), target=(0.5, 0.5, 0.5), fov_y_deg=30.0)


def _rays(cam16, mh, mw):
    """ray origins / directions of the march's pixel grid, float64"""
    o = cam16[0:3].astype(np.float64)
    fw, rt, up = cam16[3:6].astype(np.float64), cam16[6:9].astype(np.float64), cam16[9:12].astype(np.float64)
    th, asp = float(cam16[12]), float(cam16[13])
    c = (np.arange(mw) + 0.5) / mw
    r = (np.arange(mh) + 0.5) / mh
    sx = (2.0 * c - 1.0) * th * asp
    sy = (1.0 - 2.0 * r) * th
    d = fw[None, None] + rt[None, None] * sx[None, :, None] + up[None, None] * sy[:, None, None]
    return o, d / np.linalg.norm(d, axis=-1, keepdims=True)


def _chord(o, d):
    inv = 1.0 / np.where(np.abs(d) < 1e-12, 1e-12, d)
    t0 = (0.0 - o) * inv
    t1 = (1.0 - o) * inv
    tn = np.max(np.minimum(t0, t1), axis=-1)
    tf = np.min(np.maximum(t0, t1), axis=-1)
    return np.maximum(tf - np.maximum(tn, 0.0), 0.0)


def _params(**kw):
    p = RenderParams3D(light_res=16, **kw)
    return p, p.pack(), blackbody_lut(p.t_lo, p.t_hi, 256, t_ref=p.t_ref)


def test_brick_skipping_changes_nothing():
    N = 32
    p, prm, lut = _params()
    f = smooth_noise(N, 2, 7, amp=1.0)
    dens = np.clip(f[0], 0.0, None) * np.linspace(0, 1, N)[:, None, None] ** 4
    temp = 293.0 + 1500.0 * np.clip(f[1], 0.0, None) * (dens > 0.05)
    solid = np.zeros((N, N, N), np.float32)
    solid[20:26, 8:14, 10:16] = 1.0
    pr = R.prep(dens.astype(np.float32), temp.astype(np.float32), solid, prm, lut)
    cam = camera_basis(CAM, 1.25)
    a = R.march(pr, prm, lut, cam, 40, 50, skip=True)
    b = R.march(pr, prm, lut, cam, 40, 50, skip=False)
    assert (pr["occ"] == 0).mean() > 0.3        # some bricks are empty and skipped
    # after a skip the samples sit on the exact lattice t0 + k ds, while the unskipped march adds
    # ds step by step: positions differ by rounding, which can move one sample across a threshold
    # (d_eps, the early-termination limit) on a few rays
    e = np.abs(a - b).max(-1) / np.abs(b).max()
    assert (e > 1e-4).mean() < 0.005 and e.max() < 0.01


@pytest.mark.parametrize("d0", [0.01, 0.05])
def test_beer_lambert_transmittance(d0):
    N = 32      # volumes below the 32^3 brick grid leave the last brick of each axis unmarked (as the reference)
    p, prm, lut = _params(smoke_albedo=(0.0, 0.0, 0.0), floor=0.0)
    dens = np.full((N, N, N), d0, np.float32)
    temp = np.full((N, N, N), 293.0, np.float32)
    pr = R.prep(dens, temp, None, prm, lut)
    cam = camera_basis(CAM, 1.0)
    mh = mw = 24
    hdr = R.march(pr, prm, lut, cam, mh, mw)
    o, d = _rays(cam, mh, mw)
    L = _chord(o, d)
    dd = float(np.float16(d0))
    sigma = p.k_abs * dd
    ds = 1.0 / p.steps
    hit = L > 2 * ds
    assert hit.sum() > 100
    lo, hi = np.exp(-sigma * (L + ds)), np.exp(-sigma * np.maximum(L - ds, 0.0))
    T = hdr[..., 3]
    assert np.all(T[hit] >= lo[hit] * (1 - 1e-4)) and np.all(T[hit] <= hi[hit] * (1 + 1e-4))
    assert np.all(T[~hit] >= lo[~hit] * (1 - 1e-4))
    assert np.abs(hdr[..., :3]).max() == 0.0


def test_thin_hot_medium_emits_blackbody_times_chord():
    N = 32
    p, prm, lut = _params(k_abs=0.0, floor=0.0)
    Tk = 1500.0
    dens = np.full((N, N, N), 0.0, np.float32)
    temp = np.full((N, N, N), Tk, np.float32)
    pr = R.prep(dens, temp, None, prm, lut)
    cam = camera_basis(CAM, 1.0)
    hdr = R.march(pr, prm, lut, cam, 20, 20)
    o, d = _rays(cam, 20, 20)
    L = _chord(o, d)
    f = (float(np.float16(Tk)) - p.t_lo) / (p.t_hi - p.t_lo) * 255.0
    i0 = int(f)
    col = lut[i0] * (1 - (f - i0)) + lut[i0 + 1] * (f - i0)
    e = col * (p.emit_min * p.emit_scale)          # density 0: emissivity emit_min
    ds = 1.0 / p.steps
    hit = L > 4 * ds
    for c in range(3):
        got = hdr[..., c][hit]
        assert np.all(np.abs(got - e[c] * L[hit]) <= e[c] * ds * 1.01 + 1e-5)
    assert np.all(hdr[..., 3] == 1.0)


def test_opaque_sphere_silhouette():
    N = 48
    p, prm, lut = _params(floor=0.0)
    c0, rad = np.array([0.5, 0.5, 0.5]), 0.2
    g = (np.arange(N) + 0.5) / N
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
    solid = (((xx - c0[0]) ** 2 + (yy - c0[1]) ** 2 + (zz - c0[2]) ** 2) <= rad ** 2).astype(np.float32)
    z = np.zeros((N, N, N), np.float32)
    pr = R.prep(z, z + 293.0, solid, prm, lut)
    cam = camera_basis(CAM, 1.0)
    hdr = R.march(pr, prm, lut, cam, 32, 32)
    o, d = _rays(cam, 32, 32)
    v = c0 - o
    dist = np.linalg.norm(v[None, None] - (d @ v)[..., None] * d, axis=-1)
    tol = 2.0 / N                      # the binary mask is smoothed by a Gaussian of one voxel
    inside, outside = dist < rad - tol, dist > rad + tol
    assert inside.sum() > 50 and outside.sum() > 50
    assert np.all(hdr[..., 3][inside] == 0.0)
    assert np.all(hdr[..., 3][outside] == 1.0)
    assert np.all(hdr[..., :3][inside].sum(-1) > 0)      # lit obstacle colour


def test_light_volume_uniform_medium_from_above():
    N, d0 = 16, 0.02
    p, prm, lut = _params(light_dir=(0.0, 1.0, 0.0))
    dens = np.full((N, N, N), d0, np.float32)
    temp = np.full((N, N, N), 293.0, np.float32)
    pr = R.prep(dens, temp, None, prm, lut)
    L = p.light_res
    light = pr["light"].reshape(L, L, L, 2)
    dd = float(np.float16(d0))
    y = (np.arange(L) + 0.5) / L
    ds = 1.5 / L
    lo = np.exp(-p.k_shadow * dd * (1.0 - y + ds))
    hi = np.exp(-p.k_shadow * dd * np.maximum(1.0 - y - ds, 0.0))
    tr = light[..., 0]
    for j in range(L):
        assert np.all(tr[:, j] >= lo[j] * (1 - 2e-3)) and np.all(tr[:, j] <= hi[j] * (1 + 2e-3)), j
    # sky occlusion: 6 steps of 3 / L straight up, the ones below y = 1 count
    da = 3.0 / L
    for j in range(L):
        n_in = sum(1 for k in range(6) if y[j] + (k + 0.5) * da < 1.0)
        want = math.exp(-p.k_amb * dd * da * n_in)
        assert np.allclose(light[:, j, :, 1], want, rtol=2e-3), j


def test_blackbody_table():
    # the fitted colour-matching functions against CIE 1931 2-degree tabulated values
    cie = {450.0: (0.3362, 0.0380, 1.7721), 555.0: (0.5121, 1.0000, 0.0057), 600.0: (1.0622, 0.6310, 0.0008)}
    for lam, want in cie.items():
        got = cie_xyz_bar(np.array([lam]))[0]
        assert np.all(np.abs(got - np.array(want)) <= 0.08 * max(want) + 0.01), (lam, got, want)
    lut = blackbody_lut(700.0, 3000.0, 256, t_ref=1400.0)
    assert np.all(lut[0] == 0)                       # the onset fade starts at 0
    ratio_b = lut[1:, 2] / np.maximum(lut[1:, 0], 1e-9)
    assert np.all(np.diff(ratio_b[20:]) >= -1e-6)    # bluer with temperature
    lum = lut @ np.array([0.2126, 0.7152, 0.0722])
    assert np.all(np.diff(lum) >= 0)                 # brighter with temperature
    assert np.allclose(lut.max(axis=1)[1:], np.max(lut[1:], axis=1))


# ---------------------------------------------------------------------- aux (return_aux)
def _slab_aux(d0, vel=None, cam_prev=None, mh=24, mw=24, vel_dt=1.0 / 30.0):
    N = 32
    p, prm, lut = _params(smoke_albedo=(0.0, 0.0, 0.0), floor=0.0)
    dens = np.full((N, N, N), d0, np.float32)
    temp = np.full((N, N, N), 293.0, np.float32)
    pr = R.prep(dens, temp, None, prm, lut)
    cam = camera_basis(CAM, mw / mh)
    aux = dict(vel=vel, cam_prev=cam if cam_prev is None else cam_prev, vel_dt=vel_dt, out_h=mh, out_w=mw)
    hdr, abuf = R.march(pr, prm, lut, cam, mh, mw, aux=aux)
    return p, pr, prm, lut, cam, hdr, abuf


def test_aux_leaves_colour_unchanged():
    N = 32
    p, prm, lut = _params()
    f = smooth_noise(N, 3, 11, amp=1.0)
    dens = np.clip(f[0], 0.0, None).astype(np.float32)
    temp = (293.0 + 1500.0 * np.clip(f[1], 0.0, None)).astype(np.float32)
    pr = R.prep(dens, temp, None, prm, lut)
    cam = camera_basis(CAM, 1.25)
    a = R.march(pr, prm, lut, cam, 32, 40)
    b, _ = R.march(pr, prm, lut, cam, 32, 40, aux=dict(vel=f, cam_prev=cam, vel_dt=0.1, out_h=64, out_w=80))
    assert np.array_equal(a, b)


@pytest.mark.parametrize("d0", [0.01, 0.05])
def test_aux_alpha_and_depth_of_a_uniform_slab(d0):
    """alpha = 1 - exp(-sigma L) (Beer-Lambert over the chord, to one step); the opacity-weighted
    depth = the mean of the truncated exponential distance: t_near + 1/sigma - L e^(-sigma L) /
    (1 - e^(-sigma L)), to 0.6 steps: the samples sit at t_near + (jitter + k) ds with the jitter in
    [0, 1), and the step sum's mean lies half a step before the integral's (thick media) or halfway
    between the first and last sample (thin media), so the offset stays within half a step"""
    p, pr, prm, lut, cam, hdr, a = _slab_aux(d0)
    o, d = _rays(cam, 24, 24)
    L = _chord(o, d)
    inv = 1.0 / np.where(np.abs(d) < 1e-12, 1e-12, d)
    tn = np.maximum(np.max(np.minimum((0.0 - o) * inv, (1.0 - o) * inv), axis=-1), 0.0)
    sigma = p.k_abs * float(np.float16(d0))
    ds = 1.0 / p.steps
    hit = L > 2 * ds
    assert hit.sum() > 100
    al = a[..., 0]
    # no obstacle, no floor: the medium's transmittance is the march's
    assert np.array_equal(al, 1.0 - hdr[..., 3])
    lo, hi = 1 - np.exp(-sigma * np.maximum(L - ds, 0.0)), 1 - np.exp(-sigma * (L + ds))
    assert np.all(al[hit] >= lo[hit] - 1e-5) and np.all(al[hit] <= hi[hit] + 1e-5)
    dep = a[..., 1][hit] / al[hit]
    Lh = L[hit]
    want = tn[hit] + 1.0 / sigma - Lh * np.exp(-sigma * Lh) / (1.0 - np.exp(-sigma * Lh))
    err = dep - want
    assert np.abs(err).max() <= 0.6 * ds
    assert err.min() < -0.2 * ds and err.max() > 0.2 * ds      # the jitter spreads it both ways
    assert np.all(a[~(L > 0)] == 0.0)                  # rays that miss the cube carry no aux
    assert np.all(a[..., 2:][hit] ** 2 < (1e-3 * al[hit, None]) ** 2 + 1e-12)   # still medium, same camera


def _mean_point_z(cam, a, mh, mw):
    """camera-space depth of each pixel's opacity-weighted mean point"""
    o, d = _rays(cam, mh, mw)
    al = a[..., 0]
    ok = al > 0.05
    dep = np.where(ok, a[..., 1] / np.where(ok, al, 1.0), 0.0)
    q = o[None, None] + d * dep[..., None]
    return ok, (q - cam[0:3].astype(np.float64)) @ cam[3:6].astype(np.float64)


@pytest.mark.parametrize("v", [0.5, -0.75])
def test_aux_motion_of_a_uniform_velocity(v):
    """a medium moving at v along the camera's right (up) axis: the mean point moved back by v dt
    projects v dt out_w / (2 z tan(fov/2) aspect) pixels to the left (v dt out_h / (2 z tan) pixels
    below), so the motion is +that in x (-that in y: rows grow downward)"""
    N, mh, mw, dt = 32, 24, 30, 1.0 / 30.0
    for axis in (0, 1):
        vel = np.zeros((3, N, N, N), np.float32)
        vel[axis] = v
        p, pr, prm, lut, cam, hdr, a = _slab_aux(0.05, vel=vel, mh=mh, mw=mw, vel_dt=dt)
        ok, z = _mean_point_z(cam, a, mh, mw)
        assert ok.sum() > 100
        th, asp = float(cam[12]), float(cam[13])
        m = a[..., 2 + axis][ok] / a[..., 0][ok]
        other = a[..., 3 - axis][ok] / a[..., 0][ok]
        if axis == 0:
            want = v * dt * mw / (2.0 * z[ok] * th * asp)
        else:
            want = -v * dt * mh / (2.0 * z[ok] * th)
        assert np.allclose(m, want, rtol=2e-3, atol=1e-4)
        # rt and up are the world x and y axes here, so the other component stays 0 up to the
        # perspective of the displaced point's depth (unchanged: the motion is parallel to the image)
        assert np.abs(other).max() < 1e-3


def test_aux_motion_of_a_moved_camera():
    """a still medium and a camera that moved by +delta along its right axis since the last sim
    step: every point slides delta out_w / (2 z tan(fov/2) aspect) pixels to the left"""
    mh, mw, delta = 20, 28, 0.02
    cam = camera_basis(CAM, mw / mh)
    prev = cam.copy()
    prev[0:3] = cam[0:3] - delta * cam[6:9]
    p, pr, prm, lut, cam, hdr, a = _slab_aux(0.05, cam_prev=prev, mh=mh, mw=mw)
    ok, z = _mean_point_z(cam, a, mh, mw)
    assert ok.sum() > 100
    th, asp = float(cam[12]), float(cam[13])
    mx = a[..., 2][ok] / a[..., 0][ok]
    my = a[..., 3][ok] / a[..., 0][ok]
    assert np.allclose(mx, -delta * mw / (2.0 * z[ok] * th * asp), rtol=2e-3, atol=1e-4)
    assert np.abs(my).max() < 1e-3


@pytest.mark.parametrize("shape", [(24, 30), (48, 60), (37, 53)])
def test_aux_resolve_against_bilinear(shape):
    """the resolve's upsample is torch's bilinear interpolation (align_corners=False, clamped
    borders), then the division by alpha where alpha >= 1e-3"""
    import torch
    import torch.nn.functional as F
    rng = np.random.default_rng(5)
    mh, mw = 24, 30
    a = rng.random((mh, mw, 4), dtype=np.float32)
    a[..., 0] = np.where(rng.random((mh, mw)) < 0.3, 0.0, a[..., 0])
    a[..., 1:] *= a[..., :1] * 3.0
    got = R.aux_resolve(a, *shape).astype(np.float32)
    up = F.interpolate(torch.from_numpy(a).permute(2, 0, 1)[None].double(), size=shape, mode="bilinear",
                       align_corners=False)[0].numpy()
    al = up[0]
    ok = al >= 1e-3
    want = np.zeros((4,) + shape)
    want[0] = al
    want[1:] = np.where(ok, up[1:] / np.where(ok, al, 1.0), 0.0)
    both = ok | (np.abs(al - 1e-3) > 1e-6)
    assert np.allclose(got[:, both], want[:, both], rtol=2e-3, atol=2e-3)


# ---------------------------------------------------------------------- temporal resolve
def _ramp_bufs(H, W, a=0.2, b=0.02, alpha=1.0, depth=1.8):
    """rgb 0, T = a + b * column; aux = (alpha, alpha depth, 0, 0)"""
    cur = np.zeros((H, W, 4), np.float32)
    cur[..., 3] = a + b * np.arange(W, dtype=np.float32)[None, :]
    aux = np.zeros((H, W, 4), np.float32)
    aux[..., 0] = alpha
    aux[..., 1] = alpha * depth
    return cur, aux


def test_temporal_passes_through_without_history_or_on_disocclusion():
    H, W = 20, 28
    cam = camera_basis(CAM, W / H)
    rng = np.random.default_rng(3)
    cur = rng.random((H, W, 4), dtype=np.float32)
    aux = np.zeros((H, W, 4), np.float32)
    aux[..., 0], aux[..., 1] = 1.0, 1.5
    hist = rng.random((H, W, 4), dtype=np.float32)
    t0 = np.zeros(16, np.float32)                                  # history not valid
    assert np.array_equal(R.temporal_resolve(cur, aux, hist, cam, t0), cur)
    away = camera_basis(Camera(position=(0.5, 0.5, 3.0), target=(0.5, 0.5, 5.0)), W / H)
    away[15] = 1.0                                                 # valid, but every point is behind it
    assert np.array_equal(R.temporal_resolve(cur, aux, hist, cam, away), cur)


@pytest.mark.parametrize("case", ["medium", "floor", "centre"])
def test_temporal_reprojects_a_ramp_under_camera_translation(case):
    """A previous camera moved by -delta along the right axis sees every point delta W / (2 z
    tan(fov/2) aspect) pixels further right (z: the point's depth along the view axis, at the
    medium's mean depth, the floor hit or the distance to the domain centre).  A linear T ramp is
    reproduced exactly by the Catmull-Rom fetch and stays inside the 3x3 variance clip (mean +-
    sqrt(2/3) b), so the resolved T = a + b c + (1 - w) b s with w = alpha + k_t |b s| (1 + b)."""
    H, W, a, b, alpha, k_t = 24, 32, 0.2, 0.02, 0.1, 4.0
    camera = {"medium": CAM, "floor": Camera(position=(0.5, 0.8, -1.0), target=(0.5, 0.0, 0.5), fov_y_deg=30.0),
              "centre": Camera(position=(0.5, 0.2, -1.5), target=(0.5, 0.9, 0.5), fov_y_deg=30.0)}[case]
    cam = camera_basis(camera, W / H)
    o, d = _rays(cam, H, W)
    depth = 1.8
    cur, aux = _ramp_bufs(H, W, a, b, alpha=1.0 if case == "medium" else 0.0, depth=depth)
    if case == "medium":
        dep = np.full((H, W), depth)
    elif case == "floor":
        assert np.all(d[..., 1] < -1e-4)
        dep = -o[1] / d[..., 1]
    else:
        assert np.all(d[..., 1] > 0)
        dep = np.full((H, W), np.linalg.norm(np.array([0.5, 0.5, 0.5]) - o))
    z = dep * (d @ cam[3:6].astype(np.float64))
    th, asp = float(cam[12]), float(cam[13])
    delta = 0.6 * 2.0 * z.min() * th * asp / W                    # |s| <= 0.6 px everywhere
    s = delta * W / (2.0 * z * th * asp)
    prev = cam.copy()
    prev[0:3] = cam[0:3] - delta * cam[6:9]
    prev[14], prev[15] = 0.0, 1.0
    got = R.temporal_resolve(cur, aux, cur.copy(), cam, prev, alpha=alpha, gamma=1.0, k_t=k_t)
    col = np.arange(W)[None, :]
    w = np.clip(alpha + k_t * np.abs(b * s) * (1.0 + b), alpha, 1.0)
    want = a + b * col + (1.0 - w) * b * s
    inner = (slice(2, H - 2), slice(3, W - 3))                     # away from the clamped borders
    assert np.abs(got[..., 3][inner] - want[inner]).max() < 2e-5
    assert np.abs(got[..., :3]).max() == 0.0
    assert s[inner].min() > 0.1                                    # the history really moved


def test_temporal_variance_clip_of_a_checkerboard():
    """T = 0.5 +- e in a checkerboard: the 3x3 neighbourhood holds 5 of the centre's value and 4 of
    the other, mean 0.5 + p e / 9, sigma e sqrt(80) / 9, range [0.5 - e, 0.5 + e]; a history of
    0.95 is clipped to min(0.5 + e, mean + sigma) and blended with w = alpha + k_t dT (1 + e)"""
    H, W, e, alpha, k_t = 16, 20, 0.1, 0.1, 0.5
    cam = camera_basis(CAM, W / H)
    p = np.where((np.arange(H)[:, None] + np.arange(W)[None, :]) % 2 == 0, 1.0, -1.0)
    cur = np.zeros((H, W, 4), np.float32)
    cur[..., 3] = 0.5 + e * p
    aux = np.zeros((H, W, 4), np.float32)
    aux[..., 0], aux[..., 1] = 1.0, 1.6
    hist = np.zeros((H, W, 4), np.float32)
    hist[..., 3] = 0.95
    tc = cam.copy()
    tc[14], tc[15] = 0.0, 1.0
    got = R.temporal_resolve(cur, aux, hist, cam, tc, alpha=alpha, gamma=1.0, k_t=k_t)[..., 3]
    m1 = 0.5 + p * e / 9.0
    clip = np.minimum(0.5 + e, m1 + e * math.sqrt(80.0) / 9.0)
    dT = np.abs(0.5 + e * p - 0.95)
    w = np.clip(alpha + k_t * dT + 0.5 * k_t * (2 * e) * dT, alpha, 1.0)
    want = (1.0 - w) * clip + w * (0.5 + e * p)
    inner = (slice(1, H - 1), slice(1, W - 1))
    assert np.abs(got[inner] - want[inner]).max() < 1e-5

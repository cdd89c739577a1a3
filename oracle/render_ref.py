"""CPU restatement of the reference's volume renderer (vfx/render3d/renderer.py VolumeRenderer.draw,
default path) around the C kernels of oracle/render_oracle.c.

TEST INFRASTRUCTURE ONLY: the oracle of the HIP renderer (neural-vfx_amd/csrc/render.hip).

The reference's Warp kernels (pack, light volume, march, composite) are restated in C; its torch
steps (the solid and glow blurs, the glow pyramid, the bloom levels, renderer.py:227-229, :408-416,
:452-457) run here as the same torch calls on the CPU.  Stored grids are rounded to FP16 where the
reference stores FP16 (packed volume, light volume, glow source, glow volume).

Parity status (tests/test_oracle_render.py): empty-brick skipping against the unskipped march; a
uniform non-scattering slab's transmittance against Beer-Lambert; an optically thin hot slab's
emission against the blackbody table times the path length; an opaque sphere's silhouette against
the ray-sphere intersection; the light volume's transmittance and sky occlusion of a uniform medium
against their closed forms; the background of an empty frame against its gradient.  Aux (return_aux):
a uniform slab's alpha against 1 - Beer-Lambert and its depth against the closed-form weighted mean
distance; zero motion for a still medium and an unmoved camera; uniform-velocity and moved-camera
motion against the projection of the displaced mean point; the colour unchanged by the aux variant;
the resolve against float64 bilinear weights.  Temporal resolve: the first draw and disoccluded pixels
pass the march through; a linear transmittance ramp under a camera translation against the closed
form of the reprojected, Catmull-Rom fetched and blended history (with medium, floor and
domain-centre depths); the variance clip of a checkerboard against its closed form.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np
import torch
import torch.nn.functional as F

import oracle as _sim_oracle

BRICKS = 32


def _lib():
    return _sim_oracle._load()


def _p(x):
    return x.ctypes.data_as(ctypes.c_void_p)


def _f32(x):
    return np.ascontiguousarray(x, dtype=np.float32)


def _h(x):
    """round to FP16 and back (the reference's half storage)"""
    return _f32(np.asarray(x, np.float32).astype(np.float16).astype(np.float32))


def gauss1d(sigma: float) -> torch.Tensor:
    r = max(1, int(math.ceil(2.5 * sigma)))
    t = torch.arange(-r, r + 1, dtype=torch.float32)
    k = torch.exp(-0.5 * (t / sigma) ** 2)
    return k / k.sum()


def blur3(x: torch.Tensor, k: torch.Tensor) -> torch.Tensor:
    """renderer.py:_blur3: separable replicate-padded blur of (1, C, D, H, W), z then y then x."""
    C, r = x.shape[1], k.numel() // 2
    for dim in range(3):
        shp = [1, 1, 1]
        shp[dim] = -1
        kk = k.view(1, 1, *shp).expand(C, 1, *shp)
        pad = [0] * 6
        pad[2 * (2 - dim)] = pad[2 * (2 - dim) + 1] = r
        x = F.conv3d(F.pad(x, pad, mode="replicate"), kk, groups=C)
    return x


def blur2(x: torch.Tensor, k: torch.Tensor) -> torch.Tensor:
    C, r = x.shape[1], k.numel() // 2
    kx = k.view(1, 1, 1, -1).expand(C, 1, 1, -1)
    ky = k.view(1, 1, -1, 1).expand(C, 1, -1, 1)
    x = F.conv2d(F.pad(x, (r, r, 0, 0), mode="replicate"), kx, groups=C)
    return F.conv2d(F.pad(x, (0, 0, r, r), mode="replicate"), ky, groups=C)


def occupancy(a, b, mode, d_eps=0.0, t_hot=0.0):
    a, b = _f32(a), _f32(b)
    D, H, W = a.shape
    occ = np.zeros(BRICKS ** 3, np.uint8)
    _lib().or_render_occ(_p(a), _p(b), D, H, W, int(mode), ctypes.c_float(d_eps), ctypes.c_float(t_hot), _p(occ))
    return occ


def prep(dens, temp, solid, prm, lut):
    """Volume prep: packed (D, H, W, 2), occupancy, solid (values / 255) and its occupancy, light
    (L^3, 2), glow (32^3, 3)."""
    prm = _f32(prm)
    lut = _f32(lut)
    L = int(prm[1])
    d_eps, t_lo = float(prm[23]), float(prm[20])
    vol = _h(np.stack([np.maximum(dens, 0.0), temp], -1))
    occ = occupancy(dens, temp, 0, d_eps, t_lo)
    if solid is not None:
        s = blur3(torch.from_numpy(_f32(solid))[None, None], gauss1d(1.0))[0, 0].numpy()
        su8 = np.clip((s * 255.0 + 0.5).astype(np.int64), 0, 255).astype(np.float32) / 255.0
        su8 = _f32(su8)
        occ_s = occupancy(s, s, 1)
    else:
        su8 = np.zeros((2, 2, 2), np.float32)
        occ_s = np.zeros(BRICKS ** 3, np.uint8)
    D, H, W = dens.shape
    sD, sH, sW = su8.shape
    light = np.empty((L ** 3, 2), np.float32)
    gsrc = np.empty((L ** 3, 3), np.float32)
    _lib().or_render_light(_p(vol), D, H, W, _p(su8), sD, sH, sW, int(solid is not None), _p(lut), _p(prm),
                           _p(light), _p(gsrc))
    light = _h(light)
    gsrc = _h(gsrc)
    # glow pyramid (renderer.py:408-416)
    g = torch.from_numpy(gsrc.reshape(L, L, L, 3)).permute(3, 0, 1, 2)[None]
    g32 = F.adaptive_avg_pool3d(g, 32)
    g3 = gauss1d(1.2)
    g32b = blur3(g32, g3)
    g8 = blur3(F.avg_pool3d(g32, 4), g3)
    glow = g32b + float(prm[51]) * F.interpolate(g8, size=(32, 32, 32), mode="trilinear", align_corners=False)
    glow = _h(glow[0].permute(1, 2, 3, 0).numpy())
    return dict(vol=vol, occ=occ, solid=su8, occ_s=occ_s, has_solid=solid is not None, light=light, glow=glow,
                solid_res=float(max(sD, sH, sW)) if solid is not None else 1.0)


def march(pr, prm, lut, cam, mh, mw, skip=True, aux=None):
    """hdr (mh, mw, 4).  aux = dict(vel (3, D, H, W) or None, cam_prev (16,), vel_dt, out_h, out_w)
    selects the march_aux variant: returns (hdr, aux (mh, mw, 4) premultiplied)."""
    prm, lut, cam = _f32(prm), _f32(lut), _f32(cam)
    D, H, W, _ = pr["vol"].shape
    sD, sH, sW = pr["solid"].shape
    out = np.empty((mh, mw, 4), np.float32)
    vel = cam_prev = abuf = None
    uD = uH = uW = 0
    vel_dt = out_w = out_h = 0.0
    if aux is not None:
        abuf = np.empty((mh, mw, 4), np.float32)
        cam_prev = _f32(aux["cam_prev"])
        vel_dt, out_h, out_w = float(aux["vel_dt"]), float(aux["out_h"]), float(aux["out_w"])
        if aux.get("vel") is not None:
            v = np.asarray(aux["vel"], np.float32)
            _, uD, uH, uW = v.shape
            vel = np.zeros((uD, uH, uW, 4), np.float32)     # the half (D, H, W, 4) texture, 4th channel 0
            vel[..., :3] = _h(np.moveaxis(v, 0, -1))
    nul = ctypes.c_void_p(None)
    _lib().or_render_march(_p(pr["vol"]), D, H, W, _p(pr["light"]), _p(pr["glow"]), _p(pr["solid"]), sD, sH, sW,
                           _p(pr["occ"]), _p(pr["occ_s"]), int(pr["has_solid"]), ctypes.c_float(pr["solid_res"]),
                           _p(lut), _p(prm), _p(cam), mh, mw, int(skip), _p(out),
                           _p(vel) if vel is not None else nul, uD, uH, uW,
                           _p(cam_prev) if cam_prev is not None else nul, ctypes.c_float(vel_dt),
                           ctypes.c_float(out_w), ctypes.c_float(out_h), _p(abuf) if abuf is not None else nul)
    return out if aux is None else (out, abuf)


def aux_resolve(abuf, out_h, out_w, a_min=1.0e-3):
    """kernels.py:603-626: premultiplied (mh, mw, 4) -> (4, out_h, out_w) FP16 alpha, depth, motion_x,
    motion_y (bilinear upsample with clamped taps, then divided by alpha; 0 where alpha < a_min)."""
    a = _f32(abuf)
    mh, mw = a.shape[:2]
    if (mh, mw) == (out_h, out_w):
        x = a
    else:
        u = ((np.arange(out_w, dtype=np.float32) + np.float32(0.5)) / np.float32(out_w))
        v = ((np.arange(out_h, dtype=np.float32) + np.float32(0.5)) / np.float32(out_h))
        fx = u * np.float32(mw) - np.float32(0.5)
        fy = v * np.float32(mh) - np.float32(0.5)
        x0f, y0f = np.floor(fx), np.floor(fy)
        tx, ty = (fx - x0f)[None, :, None], (fy - y0f)[:, None, None]
        x0 = np.clip(x0f.astype(np.int64), 0, mw - 1)
        x1 = np.clip(x0f.astype(np.int64) + 1, 0, mw - 1)
        y0 = np.clip(y0f.astype(np.int64), 0, mh - 1)
        y1 = np.clip(y0f.astype(np.int64) + 1, 0, mh - 1)
        one = np.float32(1.0)
        r0 = a[y0][:, x0] * (one - tx) + a[y0][:, x1] * tx
        r1 = a[y1][:, x0] * (one - tx) + a[y1][:, x1] * tx
        x = r0 * (one - ty) + r1 * ty
    al = x[..., 0]
    ok = al >= np.float32(a_min)
    safe = np.where(ok, al, np.float32(1.0))
    out = np.zeros((4, out_h, out_w), np.float32)
    out[0] = al
    for k in range(1, 4):
        out[k] = np.where(ok, x[..., k] / safe, np.float32(0.0))
    return out.astype(np.float16)


def bloom(hdr, thresh):
    """renderer.py:452-457: (bh, bw, 3)"""
    mh, mw = hdr.shape[:2]
    bh, bw = max(2, mh // 4), max(2, mw // 4)
    x = torch.from_numpy(_f32(hdr[..., :3])).permute(2, 0, 1)[None]
    x = F.relu(x - thresh)
    g2 = gauss1d(2.0)
    b1 = blur2(F.adaptive_avg_pool2d(x, (bh, bw)), g2)
    b2 = blur2(F.avg_pool2d(b1, 4, ceil_mode=True), g2)
    b = b1 + 1.5 * F.interpolate(b2, size=b1.shape[-2:], mode="bilinear", align_corners=False)
    return _f32(b[0].permute(1, 2, 0).numpy())


def composite(hdr, bl, prm, out_h, out_w):
    prm = _f32(prm)
    hdr, bl = _f32(hdr), _f32(bl)
    out = np.empty((out_h, out_w, 3), np.uint8)
    _lib().or_render_composite(_p(hdr), hdr.shape[0], hdr.shape[1], _p(bl), bl.shape[0], bl.shape[1], _p(prm),
                               out_h, out_w, _p(out))
    return out


GOLDEN = 0.6180339887498949


def temporal_resolve(cur, abuf, hist, cam, tcam, alpha=0.1, gamma=1.0, k_t=4.0):
    """temporal.py:70-150: (h, w, 4) march buffer, its premultiplied aux and the history ->
    the resolved (h, w, 4); tcam[14] = minimum current weight, tcam[15] = history valid."""
    cur, abuf = _f32(cur), _f32(abuf)
    hist = _f32(hist if hist is not None else cur)
    H, W = cur.shape[:2]
    out = np.empty((H, W, 4), np.float32)
    _lib().or_render_temporal(_p(cur), _p(abuf), _p(hist), H, W, _p(_f32(cam)), _p(_f32(tcam)),
                              ctypes.c_float(alpha), ctypes.c_float(gamma), ctypes.c_float(k_t), _p(out))
    return out


class TemporalRender:
    """The temporal draw sequence of one renderer (renderer.py:331-346, :465-493): per draw the
    jitter offset (frame * golden ratio) mod 1, the march with aux, the resolve against the
    previous draw's camera, then bloom and composite on the resolved buffer."""

    def __init__(self, prm, lut, out_h, out_w, march_scale=1.0, alpha=0.1, gamma=1.0, k_t=4.0, tick=0.5):
        self.prm, self.lut, self.out_h, self.out_w = _f32(prm), _f32(lut), out_h, out_w
        self.mh = max(8, int(round(out_h * march_scale)))
        self.mw = max(8, int(round(out_w * march_scale)))
        self.alpha, self.gamma, self.k_t, self.tick = alpha, gamma, k_t, tick
        self.frame, self.prev, self.hist = 0, None, None

    def draw(self, pr, cam, reuse_volume=False):
        """(uint8 frame, resolved buffer) for a prepared volume (prep) and a camera basis"""
        self.frame += 1
        cb = _f32(cam).copy()
        cb[14] = np.float32((self.frame * GOLDEN) % 1.0)
        tcam = np.zeros(16, np.float32)
        if self.prev is not None:
            tcam[:] = self.prev
            tcam[14] = 0.0 if reuse_volume else self.tick
            tcam[15] = 1.0
        self.prev = cb.copy()
        hdr, abuf = march(pr, self.prm, self.lut, cb, self.mh, self.mw,
                          aux=dict(vel=None, cam_prev=cb, vel_dt=0.0, out_h=self.out_h, out_w=self.out_w))
        res = temporal_resolve(hdr, abuf, self.hist, cb, tcam, self.alpha, self.gamma, self.k_t)
        self.hist = res
        img = composite(res, bloom(res, float(self.prm[19])), self.prm, self.out_h, self.out_w)
        return img, res


def render(dens, temp, solid, cam, prm, lut, out_h, out_w, march_scale=1.0, aux=None):
    """The whole draw: (uint8 image (out_h, out_w, 3), hdr (mh, mw, 4)); with aux = dict(vel,
    cam_prev, vel_dt) (return_aux) also the resolved (4, out_h, out_w) FP16 aux and the
    premultiplied march-resolution aux buffer."""
    mh = max(8, int(round(out_h * march_scale)))
    mw = max(8, int(round(out_w * march_scale)))
    pr = prep(np.asarray(dens, np.float32), np.asarray(temp, np.float32), solid, prm, lut)
    if aux is None:
        hdr = march(pr, prm, lut, cam, mh, mw)
    else:
        hdr, abuf = march(pr, prm, lut, cam, mh, mw, aux=dict(aux, out_h=out_h, out_w=out_w))
    img = composite(hdr, bloom(hdr, float(_f32(prm)[19])), prm, out_h, out_w)
    if aux is None:
        return img, hdr
    return img, hdr, aux_resolve(abuf, out_h, out_w), abuf

"""Shared fixtures for the sim3d tests: seeded smooth fields, obstacles, scenes, error metrics.

The smooth random field follows the construction of the reference's test helper
(tests3d/helpers3d.py:69-75: coarse 6^3 Gaussian noise, trilinear up-sampling, box blur), on the
CPU with numpy's generator so the same seed gives the same field everywhere.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

from vfx.core.fields import AMBIENT_TEMP, DOMAIN_SIZE


def smooth_noise(res: int, channels: int, seed: int, coarse: int = 6, amp: float = 1.0) -> np.ndarray:
    g = np.random.default_rng(seed)
    z = torch.from_numpy(g.standard_normal((1, channels, coarse, coarse, coarse)).astype(np.float32))
    mid = F.interpolate(z, size=(max(res // 2, 2),) * 3, mode="trilinear", align_corners=False)
    mid = F.avg_pool3d(mid, 3, 1, 1, count_include_pad=False)
    out = amp * F.interpolate(mid, size=(res,) * 3, mode="trilinear", align_corners=False)[0]
    return np.ascontiguousarray(out.numpy(), dtype=np.float32)


def world_grid(res: int):
    c = (np.arange(res, dtype=np.float32) + 0.5) * np.float32(DOMAIN_SIZE / res)
    zz, yy, xx = np.meshgrid(c, c, c, indexing="ij")
    return xx, yy, zz


def blob(res, cx, cy, cz, sigma):
    xx, yy, zz = world_grid(res)
    return np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2 + (zz - cz) ** 2) / (2 * sigma ** 2)).astype(np.float32)


def sphere_solid(res, cx=0.5, cy=0.45, cz=0.5, r=0.14) -> np.ndarray:
    xx, yy, zz = world_grid(res)
    return ((xx - cx) ** 2 + (yy - cy) ** 2 + (zz - cz) ** 2 <= r * r).astype(np.float32)


def sealed_cell_solid(res, c, r=0.2) -> np.ndarray:
    """A centred sphere obstacle plus the six neighbours of cell c, which seal c in (a fluid cell
    with no open face: diag(K) = 0)."""
    solid = sphere_solid(res, cy=0.5, r=r).copy()
    for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
        solid[c[0] + d[0], c[1] + d[1], c[2] + d[2]] = 1.0
    return solid


def random_state(res: int, seed: int, solid: np.ndarray | None = None, vel_amp: float = 1.0) -> np.ndarray:
    """(6,N,N,N) float32: smooth density / velocity / temperature / fuel, zeroed in solids."""
    d = np.abs(smooth_noise(res, 1, seed, amp=0.5))[0]
    v = smooth_noise(res, 3, seed + 1, amp=vel_amp)
    T = AMBIENT_TEMP + np.abs(smooth_noise(res, 1, seed + 2, amp=300.0))[0]
    f = np.abs(smooth_noise(res, 1, seed + 3, amp=0.4))[0]
    data = np.concatenate([d[None], v, T[None], f[None]]).astype(np.float32)
    if solid is not None:
        fl = solid < 0.5
        data[0] *= fl
        data[1:4] *= fl
        data[4] = np.where(fl, data[4], AMBIENT_TEMP)
        data[5] *= fl
    return np.ascontiguousarray(data)


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    """max |a - b| / max |b| (0 if both are zero)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.abs(b).max()
    num = np.abs(a - b).max()
    return float(num / den) if den > 0 else float(num)


def to_dev(x: np.ndarray, dev="cuda:0") -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def soa_from_aos(a: np.ndarray, m: int, ncomp: int) -> np.ndarray:
    """[m^3][ncomp] (component-fastest) -> (ncomp, m, m, m)."""
    return np.ascontiguousarray(a.reshape(m, m, m, ncomp).transpose(3, 0, 1, 2))

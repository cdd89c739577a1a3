"""Blackbody emission colour table (the reference's vfx/render/blackbody.py), computed once on the
host when a renderer is created.

For each temperature, Planck's spectral radiance is integrated against the CIE 1931 2-degree
colour-matching functions, approximated by the multi-lobe piecewise Gaussian fit of Wyman, Sloan
and Shirley, "Simple Analytic Approximations to the CIE XYZ Color Matching Functions" (JCGT 2013),
and converted to linear sRGB (IEC 61966-2-1 matrix).  Each entry is the chromaticity (largest
channel 1) times (Y(T) / Y(t_ref))^compress, faded in by a smoothstep over [t_lo, t_lo + onset].
"""
from __future__ import annotations

import numpy as np

# XYZ -> linear sRGB (D65), IEC 61966-2-1
XYZ_TO_LINEAR_SRGB = np.array([[3.2406, -1.5372, -0.4986],
                               [-0.9689, 1.8758, 0.0415],
                               [0.0557, -0.2040, 1.0570]])

# Wyman-Sloan-Shirley multi-lobe fit: per colour-matching function, lobes (weight, mean nm,
# sigma below the mean, sigma above the mean)
_LOBES = {
    "x": ((1.056, 599.8, 37.9, 31.0), (0.362, 442.0, 16.0, 26.7), (-0.065, 501.1, 20.4, 26.2)),
    "y": ((0.821, 568.8, 46.9, 40.5), (0.286, 530.9, 16.3, 31.1)),
    "z": ((1.217, 437.0, 11.8, 36.0), (0.681, 459.0, 26.0, 13.8)),
}

PLANCK_H = 6.62607015e-34      # J s
LIGHT_C = 2.99792458e8         # m / s
BOLTZMANN_K = 1.380649e-23     # J / K


def cie_xyz_bar(lam_nm: np.ndarray) -> np.ndarray:
    """(n, 3) x-bar, y-bar, z-bar at the wavelengths (nm)."""
    out = []
    for name in ("x", "y", "z"):
        acc = np.zeros_like(lam_nm, dtype=np.float64)
        for w, mu, s_lo, s_hi in _LOBES[name]:
            sig = np.where(lam_nm < mu, s_lo, s_hi)
            acc += w * np.exp(-0.5 * ((lam_nm - mu) / sig) ** 2)
        out.append(acc)
    return np.stack(out, -1)


def planck(lam_nm: np.ndarray, T: np.ndarray) -> np.ndarray:
    """Spectral radiance (len(T), len(lam)) of a blackbody."""
    lam = lam_nm[None, :] * 1e-9
    return (2.0 * PLANCK_H * LIGHT_C ** 2) / (lam ** 5 * np.expm1(PLANCK_H * LIGHT_C / (lam * BOLTZMANN_K * T[:, None])))


def blackbody_lut(t_lo: float = 700.0, t_hi: float = 3000.0, n: int = 256, t_ref: float = 1800.0,
                  compress: float = 0.33, onset: float = 150.0) -> np.ndarray:
    """(n, 3) float32 linear-sRGB emission for temperatures linspace(t_lo, t_hi, n)."""
    lam = np.linspace(380.0, 780.0, 401)
    T = np.linspace(t_lo, t_hi, n)
    xyz = planck(lam, np.append(T, t_ref)) @ cie_xyz_bar(lam)          # (n + 1, 3)
    rgb = np.clip(xyz @ XYZ_TO_LINEAR_SRGB.T, 0.0, None)
    chroma = rgb / rgb.max(axis=1, keepdims=True)
    lum = (xyz[:-1, 1] / xyz[-1, 1]) ** compress
    s = np.clip((T - t_lo) / onset, 0.0, 1.0)
    fade = s * s * (3.0 - 2.0 * s)
    return (chroma[:-1] * (lum * fade)[:, None]).astype(np.float32)

"""numpy/ctypes front end of the C oracle (oracle/sim3d_oracle.c): a CPU restatement of the
reference vfx.sim3d step in FP32 and FP64.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
use it, as the checker (and as the timed CPU path); the product never imports it.

Parity status of each function (how it is pinned, see tests/test_oracle_*.py):
  advect          closed forms (constant field, integer shift, linear field), scipy
                  map_coordinates at the traced points, clamp bounds, multichannel = per-channel
  mass_ratio      closed form (numpy sums)
  sources         per-cell closed forms of each term in isolation (emitter footprint, impulse,
                  combustion, cooling, dissipation, solid reset), Philox known-answer vectors
  diffuse         conservation, constant fixed point, single-spike stencil, lambda cap
  curl            exact curl of linear fields in the interior; solid / wall zero rule
  smooth3         equals the reference's own 27-tap form (binom3) and a numpy scatter form
  body_forces     buoyancy closed form; confinement vs a numpy restatement on a field with
                  known gradient; wind marker closed form
  mask/kmul/div2  K assembled column by column equals D D^T with D assembled independently;
                  symmetric, PSD, constant null space; pure-gradient removal
  probe/galerkin  T K T^T of the assembled K (component-diagonal blocks), Galerkin formula from
                  explicit prolongation matrices
  precond         symmetric M^-1, positive definite, error reduction per application
  project         exact FP64 sparse solve (scipy) of the same system; d(u) = 0 to tolerance
  project_fused   Chebyshev weights closed form, one sweep by hand, convergence to the exact solve
  hires_advect    S == N equals advect bit for bit; constant field; integer fine-cell shift by a
                  uniform velocity; linear field under a uniform velocity (exact for trilinear)
  hires_sources   S == N equals the scalar channels of sources; emitter footprint closed form at
                  fine centres; solid reset
  hires_diffuse   S == N equals diffuse's T / F; conservation; single-spike stencil
  hires_restrict  integer ratio = numpy block mean; any ratio = an independent float64 overlap
                  matrix; mass conservation; coarse solids
  hires_stage     the composition of the above (advect, mass ratio, sources, diffusion)
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build() -> str:
    """Compile the oracle (gcc, OpenMP, no FP contraction)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _dt(prec: str):
    return np.float32 if prec == "f32" else np.float64


def _a(x, dtype):
    return np.ascontiguousarray(x, dtype=dtype)


def _p(x):
    return x.ctypes.data_as(ctypes.c_void_p)


def _fn(name: str, prec: str, restype=None):
    f = getattr(_load(), f"{name}_{prec}" if prec else name)
    f.restype = restype
    return f


def _real(prec):
    return ctypes.c_float if prec == "f32" else ctypes.c_double


def p_size() -> int:
    return int(_load().or_p_size())


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    _load().or_philox4x32_10(c, k, o)
    return list(o)


def advect(src, vel, s, prec="f32"):
    dt = _dt(prec)
    src = _a(src, dt)
    vel = _a(vel, dt)
    C, N = src.shape[0], src.shape[-1]
    hat = np.empty_like(src)
    out = np.empty_like(src)
    _fn("or_advect", prec)(ctypes.c_int(N), ctypes.c_int(C), _p(src), _p(vel), _real(prec)(s), _p(hat), _p(out))
    return out


def mass_ratio(pre, post, solid, prec="f32"):
    dt = _dt(prec)
    pre, post, solid = _a(pre, dt), _a(post, dt), _a(solid, dt)
    return float(_fn("or_mass_ratio", prec, ctypes.c_double)(ctypes.c_int(solid.shape[-1]), _p(pre), _p(post),
                                                              _p(solid)))


def sources(adv, solid, prm, ratio, prec="f32"):
    dt = _dt(prec)
    adv, solid = _a(adv, dt), _a(solid, dt)
    prm = _a(prm, np.float32)
    out = np.empty_like(adv)
    _fn("or_sources", prec)(ctypes.c_int(solid.shape[-1]), _p(adv), _p(solid), _p(prm), _real(prec)(ratio), _p(out))
    return out


def diffuse(src, solid, prm, nsub=1.0, prec="f32"):
    dt = _dt(prec)
    src, solid = _a(src, dt), _a(solid, dt)
    prm = _a(prm, np.float32)
    out = np.empty_like(src)
    _fn("or_diffuse", prec)(ctypes.c_int(solid.shape[-1]), _p(src), _p(solid), _p(prm), _real(prec)(nsub), _p(out))
    return out


def curl(data, solid, inv2dx, prec="f32"):
    dt = _dt(prec)
    data, solid = _a(data, dt), _a(solid, dt)
    N = solid.shape[-1]
    w = np.empty((3, N, N, N), dt)
    _fn("or_curl", prec)(ctypes.c_int(N), _p(data), _p(solid), _real(prec)(inv2dx), _p(w))
    return w


def smooth3(a, solid, mode=0, dt_=0.0, out=None, prec="f32"):
    """mode 0: returns the smoothed 3 channels; mode 1: adds dt * smoothed to out[1:4] (a copy)."""
    dt = _dt(prec)
    a, solid = _a(a, dt), _a(solid, dt)
    N = solid.shape[-1]
    ta = np.empty((4, N, N, N), dt)
    tb = np.empty((4, N, N, N), dt)
    o = np.empty((3, N, N, N), dt) if mode == 0 else _a(out, dt).copy()
    _fn("or_smooth3", prec)(ctypes.c_int(N), _p(a), _p(solid), _p(ta), _p(tb), ctypes.c_int(mode), _real(prec)(dt_),
                            _p(o))
    return o


def binom3(a, solid, prec="f32"):
    dt = _dt(prec)
    a, solid = _a(a, dt), _a(solid, dt)
    N = solid.shape[-1]
    o = np.empty((3, N, N, N), dt)
    _fn("or_binom3", prec)(ctypes.c_int(N), _p(a), _p(solid), _p(o))
    return o


def body_forces(f, solid, om, prm, prec="f32"):
    dt = _dt(prec)
    f, solid, om = _a(f, dt), _a(solid, dt), _a(om, dt)
    prm = _a(prm, np.float32)
    N = solid.shape[-1]
    force = np.empty((3, N, N, N), dt)
    _fn("or_body_forces", prec)(ctypes.c_int(N), _p(f), _p(solid), _p(om), _p(prm), _p(force))
    return force


def mask(solid, prec="f32"):
    solid = _a(solid, _dt(prec))
    fm = np.empty(solid.shape, np.int32)
    _fn("or_mask", prec)(ctypes.c_int(solid.shape[-1]), _p(solid), _p(fm))
    return fm


def kmul(fm, p, prec="f32"):
    fm = _a(fm, np.int32)
    p = _a(p, _dt(prec))
    out = np.empty_like(p)
    _fn("or_kmul", prec)(ctypes.c_int(fm.shape[-1]), _p(fm), _p(p), _p(out))
    return out


def div2(fm, vel3, prec="f32"):
    fm = _a(fm, np.int32)
    vel3 = _a(vel3, _dt(prec))
    out = np.empty(fm.shape, _dt(prec))
    _fn("or_div2", prec)(ctypes.c_int(fm.shape[-1]), _p(fm), _p(vel3), _p(out))
    return out


def apply(fm, phi, vel3, prec="f32"):
    fm = _a(fm, np.int32)
    phi = _a(phi, _dt(prec))
    v = _a(vel3, _dt(prec)).copy()
    _fn("or_apply", prec)(ctypes.c_int(fm.shape[-1]), _p(fm), _p(phi), _p(v))
    return v


def div_abs(fm, vel3, prec="f32"):
    fm = _a(fm, np.int32)
    vel3 = _a(vel3, _dt(prec))
    out = np.empty(fm.shape, _dt(prec))
    _fn("or_div_abs", prec)(ctypes.c_int(fm.shape[-1]), _p(fm), _p(vel3), _p(out))
    return out


def level_sizes(N: int):
    """m of every multigrid level (pressure.py:799-809)."""
    ms = []
    m = N // 2
    while True:
        ms.append(m)
        if m % 2 == 1 or m == 1:
            break
        m //= 2
    return ms


def build_ops(fm, prec="f32"):
    """[(diag (8,m,m,m), off (24,m,m,m)), ...] per level, reference layout."""
    fm = _a(fm, np.int32)
    N = fm.shape[-1]
    ms = level_sizes(N)
    total = sum(32 * m ** 3 for m in ms)
    buf = np.empty(total, _dt(prec))
    n = _fn("or_build_ops", prec, ctypes.c_int)(ctypes.c_int(N), _p(fm), _p(buf))
    assert n == len(ms)
    out, o = [], 0
    for m in ms:
        d = buf[o:o + 8 * m ** 3].reshape(8, m, m, m); o += 8 * m ** 3
        f = buf[o:o + 24 * m ** 3].reshape(24, m, m, m); o += 24 * m ** 3
        out.append((d, f))
    return out


def precond_apply(fm, r, prec="f32"):
    fm = _a(fm, np.int32)
    r = _a(r, _dt(prec))
    z = np.empty_like(r)
    rz = _fn("or_precond_apply", prec, ctypes.c_double)(ctypes.c_int(fm.shape[-1]), _p(fm), _p(r), _p(z))
    return z, float(rz)


MG16 = {"fp32": 0, "fp16": 1, "bf16": 2}


def precond_apply16(fm, r, storage):
    """z = M^-1 r with the multigrid levels stored as `storage` ("fp32", "fp16", "bf16"; FP32
    arithmetic), the level-1 right-hand side scaled by the power of two that brings max|r| into
    [1, 2) (or_precond_apply16)."""
    fm = _a(fm, np.int32)
    r = _a(r, np.float32)
    z = np.empty_like(r)
    rz = _fn("or_precond_apply16", "f32", ctypes.c_double)(ctypes.c_int(fm.shape[-1]), _p(fm), _p(r), _p(z),
                                                          ctypes.c_int(MG16[storage]))
    return z, float(rz)


def round16(values, storage):
    """Each float rounded to the 16-bit storage format and back (or_round16)."""
    fn = _load().or_round16
    fn.argtypes = [ctypes.c_float, ctypes.c_int]
    fn.restype = ctypes.c_float
    return np.array([fn(float(v), MG16[storage]) for v in np.ravel(values)], np.float32).reshape(np.shape(values))


def project(data, solid, prm, phi, prec="f32"):
    """Project data[1:4] in place; phi (warm start) updated in place.  Returns (iters, stats)."""
    assert data.dtype == _dt(prec) and data.flags.c_contiguous
    assert phi.dtype == _dt(prec) and phi.flags.c_contiguous
    solid = _a(solid, _dt(prec))
    prm = _a(prm, np.float32)
    stats = np.zeros(3, np.float64)
    it = _fn("or_project", prec, ctypes.c_int)(ctypes.c_int(solid.shape[-1]), _p(data), _p(solid), _p(prm), _p(phi),
                                               _p(stats))
    return int(it), {"b_norm": stats[0], "r_norm": stats[1], "max_u": stats[2]}


def project_fused(data, solid, prm, phi, cycles, smooth, smoother="chebyshev", prec="f32"):
    """The fused (stationary multigrid) projection of data[1:4] in place, phi updated in place
    (pressure_fused.py:398-489).  Returns (cycles run, {"max_u", "max_r"})."""
    assert data.dtype == _dt(prec) and data.flags.c_contiguous
    assert phi.dtype == _dt(prec) and phi.flags.c_contiguous
    solid = _a(solid, _dt(prec))
    prm = _a(prm, np.float32)
    stats = np.zeros(2, np.float64)
    it = _fn("or_project_fused", prec, ctypes.c_int)(ctypes.c_int(solid.shape[-1]), _p(data), _p(solid), _p(prm),
                                                     _p(phi), ctypes.c_int(int(cycles)), ctypes.c_int(int(smooth)),
                                                     ctypes.c_int(1 if smoother == "jacobi" else 0), _p(stats))
    return int(it), {"max_u": stats[0], "max_r": stats[1]}


def fused_weights(smooth, smoother="chebyshev"):
    """Sweep weights of the fused projection (chebyshev_weights, pressure_fused.py:348-352)."""
    n = max(1, int(smooth))
    if smoother == "jacobi":
        return [0.8] * n
    lo, hi = 0.5, 2.0
    return [1.0 / ((hi + lo) / 2 + (hi - lo) / 2 * math.cos(math.pi * (2 * i + 1) / (2 * n))) for i in range(n)]


def stage_advect(data, solid, prm, inv_dx, prec="f32"):
    dt = _dt(prec)
    data, solid = _a(data, dt), _a(solid, dt)
    prm = _a(prm, np.float32)
    adv = np.empty_like(data)
    ratio = _fn("or_stage_advect", prec, ctypes.c_double)(ctypes.c_int(solid.shape[-1]), _p(data), _p(solid), _p(prm),
                                                           _real(prec)(inv_dx), _p(adv))
    return adv, float(ratio)


def stage_forces(adv, solid, prm, ratio, inv_dx, prec="f32", hi=None):
    """hi: the fine scalars (3, S, S, S) after the fine stage, restricted into the result's
    density, temperature and fuel after the diffusion (scalar_res)."""
    dt = _dt(prec)
    adv, solid = _a(adv, dt), _a(solid, dt)
    prm = _a(prm, np.float32)
    data = np.empty_like(adv)
    r = np.float32(ratio) if prec == "f32" else float(ratio)
    if hi is None:
        _fn("or_stage_forces", prec)(ctypes.c_int(solid.shape[-1]), _p(adv), _p(solid), _p(prm), _real(prec)(r),
                                     _real(prec)(inv_dx), _p(data))
    else:
        hi = _a(hi, dt)
        _fn("or_stage_forces_hires", prec)(ctypes.c_int(solid.shape[-1]), _p(adv), _p(solid), _p(prm),
                                           _real(prec)(r), _real(prec)(inv_dx), ctypes.c_int(hi.shape[-1]), _p(hi),
                                           _p(data))
    return data


# ---------------------------------------------------------------- finer scalar grid (hires.py)
def hires_map(N, S):
    """f2c constants a, b (fine index -> coarse index, p_c = a p_f + b) as float32."""
    return np.float32(N / S), np.float32(0.5 * N / S - 0.5)


def hires_advect(src, vel, s, prec="f32"):
    """MacCormack of fine (3, S^3) scalars by the coarse velocity (3, N^3), s = dt / dx_coarse."""
    dt = _dt(prec)
    src, vel = _a(src, dt), _a(vel, dt)
    S, N = src.shape[-1], vel.shape[-1]
    hat = np.empty_like(src)
    out = np.empty_like(src)
    _fn("or_hires_advect", prec)(ctypes.c_int(S), ctypes.c_int(N), _p(src), _p(vel), _real(prec)(s), _p(hat), _p(out))
    return out


def hires_sources(adv, hsolid, prm, ratio, prec="f32"):
    dt = _dt(prec)
    adv, hsolid = _a(adv, dt), _a(hsolid, dt)
    prm = _a(prm, np.float32)
    S = hsolid.shape[-1]
    out = np.empty_like(adv)
    r = np.float32(ratio) if prec == "f32" else float(ratio)
    dxf = np.float32(1.0 / S) if prec == "f32" else 1.0 / S
    _fn("or_hires_sources", prec)(ctypes.c_int(S), _p(adv), _p(hsolid), _p(prm), _real(prec)(r), _real(prec)(dxf),
                                  _p(out))
    return out


def hires_diffuse(src, hsolid, prm, nsub=1.0, prec="f32"):
    dt = _dt(prec)
    src, hsolid = _a(src, dt), _a(hsolid, dt)
    prm = _a(prm, np.float32)
    S = hsolid.shape[-1]
    out = np.empty_like(src)
    dxf = np.float32(1.0 / S) if prec == "f32" else 1.0 / S
    _fn("or_hires_diffuse", prec)(ctypes.c_int(S), _p(src), _p(hsolid), _p(prm), _real(prec)(nsub), _real(prec)(dxf),
                                  _p(out))
    return out


def hires_restrict(fine, csolid, prm, data, prec="f32"):
    """data (6, N^3) copy with channels 0, 4, 5 replaced by the restriction of fine (3, S^3)."""
    dt = _dt(prec)
    fine, csolid = _a(fine, dt), _a(csolid, dt)
    prm = _a(prm, np.float32)
    out = _a(data, dt).copy()
    _fn("or_hires_restrict", prec)(ctypes.c_int(csolid.shape[-1]), ctypes.c_int(fine.shape[-1]), _p(fine), _p(csolid),
                                   _p(prm), _p(out))
    return out


def hires_stage(hi, hsolid, vel, prm, inv_dx, prec="f32"):
    """The fine stage in place on hi (3, S^3) by the coarse velocity vel (3, N^3); returns the
    fine mass ratio."""
    assert hi.dtype == _dt(prec) and hi.flags.c_contiguous
    dt = _dt(prec)
    hsolid, vel = _a(hsolid, dt), _a(vel, dt)
    prm = _a(prm, np.float32)
    return float(_fn("or_hires_stage", prec, ctypes.c_double)(ctypes.c_int(vel.shape[-1]), ctypes.c_int(hi.shape[-1]),
                                                             _p(hi), _p(hsolid), _p(vel), _p(prm),
                                                             _real(prec)(inv_dx)))


def step(data, solid, prm, phi, inv_dx, prec="f32", hi=None, hsolid=None):
    """One whole step (advect, [fine scalars,] forces, project); returns (new data, iters).
    With hi / hsolid (scalar_res), hi is updated in place."""
    adv, ratio = stage_advect(data, solid, prm, inv_dx, prec)
    if hi is not None:
        hires_stage(hi, hsolid, data[1:4], prm, inv_dx, prec)
    out = stage_forces(adv, solid, prm, ratio, inv_dx, prec, hi=hi)
    it, _ = project(out, solid, prm, phi, prec)
    return out, it

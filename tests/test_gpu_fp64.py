"""GPU tests of the FP64 run of the same solver (libvfxsim3d.so's *_f64 entry points).

The FP64 build is the oracle's FP64 instantiation on the GPU: same kernels, double fields,
levels and constants.  It is checked against the oracle in FP64, then used the way the FP32
path is judged: the FP32 step's error in every field against the FP64 run of the same scene.
"""
from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest
import torch

import oracle
from simhelp import rel_err, random_state, smooth_noise, sphere_solid
from vfx.core.fields3d import Emitter3D, Forces3D, Impulse3D, Obstacle3D
from vfx.sim3d import Solver3D, SolverConfig3D, runtime_config
from vfx.sim3d import _lib
from vfx.sim3d.solver import pack_params

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev64(x):
    return torch.from_numpy(np.ascontiguousarray(x, np.float64)).to(DEV)


FIRE = Forces3D(emitters=[Emitter3D(pos=(0.5, 0.12, 0.5), radius=0.09, density_rate=2.0, temperature=1300.0,
                                    fuel_rate=1.5, velocity=(0.0, 1.0, 0.0))],
                impulses=[Impulse3D(pos=(0.3, 0.5, 0.4), radius=0.1, velocity=(2.0, 0.5, -1.0))],
                wind=(0.2, 0.0, 0.1))


@pytest.mark.parametrize("N", [32, 64])
def test_fp64_precond_parity(cuda, hiplib, N):
    """z = M^-1 r of the FP64 build equals the oracle's FP64 preconditioner."""
    _say(f"fp64 precond N={N}")
    solid = sphere_solid(N)
    fm = oracle.mask(solid)
    r = smooth_noise(N, 1, 5)[0].astype(np.float64) * (fm > 0)
    s = Solver3D(N, dtype=torch.float64)
    sd, rd = _dev64(solid), _dev64(r)
    z = torch.empty(N, N, N, device=DEV, dtype=torch.float64)
    _lib.check(hiplib.vfx_precond_f64(s._ctx, _ptr(sd), _ptr(rd), _ptr(z), _stream()), f64=True)
    torch.cuda.synchronize()
    zref, _ = oracle.precond_apply(fm, r, prec="f64")
    e = rel_err(z.cpu().numpy(), zref)
    _say(f"  max rel err {e:.3e}")
    assert e <= 1e-12


@pytest.mark.parametrize("N", [32, 48])
def test_fp64_step_parity(cuda, hiplib, N):
    """One FP64 step (fixed 5 PCG iterations) equals the oracle's FP64 step."""
    _say(f"fp64 step N={N}")
    solid = sphere_solid(N)
    data = random_state(N, 61, solid, vel_amp=1.5).astype(np.float64)
    cfg = runtime_config(pressure_max_iters=5)
    s = Solver3D(N, cfg, dtype=torch.float64)
    st = s.reset()
    assert st.data.dtype == torch.float64
    st.data.copy_(_dev64(data))
    st.solid.copy_(_dev64(solid))
    s.step(st, FIRE, 1.0 / 60.0)
    got = st.data.cpu().numpy()
    prm = pack_params(cfg, N, FIRE, 1.0 / 60.0, 0)
    ref, it = oracle.step(data, solid.astype(np.float64), prm, np.zeros((N,) * 3, np.float64), float(N), prec="f64")
    assert s.pressure_iters() == it == 5
    for c, name in enumerate(("density", "vx", "vy", "vz", "T", "fuel")):
        e = rel_err(got[c], ref[c])
        _say(f"  {name}: {e:.3e}")
        assert e <= 1e-10, name


def _plume_config(prec_rel_tol=1e-5):
    """The 64^3 smoke plume past a sphere: buoyancy beta 1 per K, confinement 0.3, emitter noise
    U[0.8, 1.2], pressure to relative residual 1e-5."""
    return SolverConfig3D(buoyancy_alpha=0.0, buoyancy_beta=1.0, vorticity_eps=0.3, pressure_rel_tol=prec_rel_tol,
                          pressure_max_iters=200, pressure_restarts=2, emitter_noise=0.2, noise_seed=0)


PLUME = Forces3D(emitters=[Emitter3D(pos=(0.5, 0.1, 0.5), radius=0.08, density_rate=1.0 * 60.0,
                                     temperature=294.0, fuel_rate=0.0, velocity=(0.0, 0.0, 0.0))])


@pytest.mark.parametrize("storage", ["fp32", "fp16", "bf16"])
def test_fp32_against_fp64_run(cuda, hiplib, storage):
    """The FP32 path (multigrid levels stored in `storage`) against the FP64 run of the same 64^3
    plume past a sphere (40 steps): both converge to true relative residual 1e-5 every step; the
    FP32 fields stay close to the FP64 ones."""
    import dataclasses
    N, steps, dt = 64, 40, 1.0 / 60.0
    obst = [Obstacle3D(kind="sphere", pos=(0.5, 0.5, 0.5), size=(0.15, 0.15, 0.15))]
    runs = {}
    for dtype in (torch.float32, torch.float64):
        cfg = _plume_config()
        if dtype == torch.float32:
            cfg = dataclasses.replace(cfg, mg_storage=storage)
        s = Solver3D(N, cfg, dtype=dtype)
        s.set_option(_lib.OPT_TRUE_RESIDUAL, 1)
        st = s.reset(obst)
        worst, iters = 0.0, 0
        for i in range(steps):
            s.step(st, PLUME, dt)
            ps = s.pressure_stats()
            worst = max(worst, ps["true_rel_res"])
            iters += ps["iters"]
        div = s.divergence(st)
        runs[dtype] = (st.data.double().cpu().numpy(), float(div.max()), worst)
        _say(f"  {dtype} (levels {cfg.mg_storage}): worst true rel-res {worst:.2e}, {iters / steps:.2f} PCG "
             f"iterations per step, max normalised |div| {float(div.max()):.3e}")
        assert worst <= 1e-5
    a, b = runs[torch.float32][0], runs[torch.float64][0]
    for c, name in enumerate(("density", "vx", "vy", "vz", "T", "fuel")):
        if np.abs(b[c]).max() == 0.0 and np.abs(a[c]).max() == 0.0:
            continue
        ref = b[c] - (293.0 if name == "T" else 0.0)
        got = a[c] - (293.0 if name == "T" else 0.0)
        emax = rel_err(got, ref)
        el2 = float(np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300))
        _say(f"  FP32 vs FP64 {name}: max {emax:.3e}  L2 {el2:.3e}")
        assert el2 <= 1e-2, name


@pytest.mark.parametrize("N,storage,tol", [(48, "fp16", 2e-3), (48, "bf16", 2e-2), (64, "fp16", 2e-3),
                                           (64, "bf16", 2e-2), (256, "fp16", 2e-3)])
def test_mg16_precond(cuda, hiplib, storage, tol, N):
    """A V-cycle with 16-bit level storage matches the oracle's 16-bit storage path
    (oracle.precond_apply16: every level value rounded where it is stored, the level-1 right-hand
    side scaled by a power of two from max|r|) to 1e-6, stays close to the FP32 one, and its
    standard-stencil path is bit-identical to reading every coefficient from the level arrays.
    N = 48 has even levels of sizes that are not powers of two (24, 12, 6; the item decomposition
    divides) and an odd coarsest level (3); N = 64 has power-of-two levels down to 1; at N = 256
    the level-1 launches stride over their items in two passes."""
    _say(f"precond with {storage} levels N={N}")
    solid = sphere_solid(N)
    fm = oracle.mask(solid)
    r = smooth_noise(N, 1, 9)[0] * (fm > 0)
    sd = torch.from_numpy(solid).to(DEV)
    rd = torch.from_numpy(np.ascontiguousarray(r)).to(DEV)
    outs = {}
    for kind, std in (("fp32", 1), (storage, 1), (storage, 0)):
        s = Solver3D(N)
        s.set_option(_lib.OPT_MG_STORAGE, _lib.MG_STORAGE[kind])
        s.set_option(_lib.OPT_STD_STENCIL, std)
        z = torch.empty(N, N, N, device=DEV)
        _lib.check(hiplib.vfx_precond(s._ctx, _ptr(sd), _ptr(rd), _ptr(z), _stream()))
        torch.cuda.synchronize()
        outs[(kind, std)] = z.cpu()
    assert torch.equal(outs[(storage, 1)], outs[(storage, 0)])
    ref, _ = oracle.precond_apply16(fm, r.astype(np.float32), storage)
    eo = rel_err(outs[(storage, 1)].numpy(), ref)
    e = rel_err(outs[(storage, 1)].numpy(), outs[("fp32", 1)].numpy())
    _say(f"  max rel difference to the oracle's {storage} storage {eo:.3e}, to FP32 levels {e:.3e}")
    assert eo <= 1e-6
    assert e <= tol

"""HIP path vs the oracle, stage by stage, through the C ABI (libvfxsim3d.so).

Bars (written per test): the FP32 HIP kernels and the FP32 oracle are both built without FP
contraction and follow the reference's operation order, so advection, diffusion, curl,
smoothing, operators and the V-cycle agree to round-off (max rel err <= 1e-6); stages with
transcendentals (exp, sqrt) <= 1e-5 (the north-star bar for advection and force fields); the
projection is compared at equal iteration counts (<= 1e-4 of max|u|) and against an FP64 oracle
run, and both reach the same stopping criterion.
"""
from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest
import torch

import oracle
from simhelp import rel_err, random_state, sealed_cell_solid, smooth_noise, soa_from_aos, sphere_solid, to_dev
from vfx.core.fields3d import Emitter3D, Forces3D, Impulse3D
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


FIRE = Forces3D(emitters=[Emitter3D(pos=(0.5, 0.12, 0.5), radius=0.09, density_rate=2.0, temperature=1300.0,
                                    fuel_rate=1.5, velocity=(0.0, 1.0, 0.0)),
                          Emitter3D(pos=(0.3, 0.2, 0.6), radius=0.07, density_rate=1.0, temperature=900.0)],
                impulses=[Impulse3D(pos=(0.3, 0.5, 0.4), radius=0.1, velocity=(2.0, 0.5, -1.0))],
                wind=(0.2, 0.0, 0.1))


@pytest.mark.parametrize("N,C", [(32, 3), (48, 3), (32, 1), (32, 4), (32, 5), (32, 6)])
def test_advect_parity(cuda, hiplib, N, C):
    """every channel count: the backward values travel between the passes as channel pairs, the
    last pair of an odd count padded"""
    _say(f"advect parity N={N} C={C}")
    f = np.concatenate([smooth_noise(N, 2, 11), np.abs(smooth_noise(N, 4, 12))]).astype(np.float32)[:C]
    v = smooth_noise(N, 3, 13, amp=3.0)
    dt = 1.0 / 30.0
    s = Solver3D(N)
    got = s.advect(to_dev(f), to_dev(v), dt).cpu().numpy()
    inv_dx = np.float32(1.0 / (1.0 / N))
    ref = oracle.advect(f, v, np.float32(np.float32(dt) * inv_dx))
    e = rel_err(got, ref)
    _say(f"  max rel err {e:.3e}, mismatched values {(got != ref).sum()} of {got.size}")
    assert e <= 1e-6


def test_advect_single_channel_squeeze(cuda, hiplib):
    N = 32
    f = smooth_noise(N, 2, 3)
    v = smooth_noise(N, 3, 4, amp=2.0)
    s = Solver3D(N)
    both = s.advect(to_dev(f), to_dev(v), 0.05)
    one = s.advect(to_dev(f[1]).contiguous(), to_dev(v), 0.05)
    assert one.shape == (N, N, N)
    assert torch.equal(both[1], one)


@pytest.mark.parametrize("N,obst", [(32, True), (48, False)])
def test_stage_advect_parity(cuda, hiplib, N, obst):
    _say(f"stage advect parity N={N} obstacle={obst}")
    solid = sphere_solid(N) if obst else np.zeros((N,) * 3, np.float32)
    data = random_state(N, 21, solid, vel_amp=2.0)
    cfg = SolverConfig3D()
    prm = pack_params(cfg, N, FIRE, 1.0 / 60.0, 0)
    s = Solver3D(N, cfg)
    d = to_dev(data)
    sd = to_dev(solid)
    _lib.check(hiplib.vfx_step_stage(s._ctx, _lib.STAGE_ADVECT, _ptr(d), _ptr(sd),
                                     prm.ctypes.data_as(ctypes.c_void_p), _stream()))
    adv = torch.empty_like(d)
    ratio = ctypes.c_double()
    _lib.check(hiplib.vfx_get_advected(s._ctx, _ptr(adv), ctypes.byref(ratio), _stream()))
    torch.cuda.synchronize()
    ref_adv, ref_ratio = oracle.stage_advect(data, solid, prm, np.float32(N))
    e = rel_err(adv.cpu().numpy(), ref_adv)
    _say(f"  adv max rel err {e:.3e}; ratio {ratio.value:.12f} vs {ref_ratio:.12f}")
    assert e <= 1e-6
    assert abs(ratio.value - ref_ratio) <= 1e-12 * abs(ref_ratio)


@pytest.mark.parametrize("N,obst,noise", [(32, True, 0.0), (48, True, 0.2), (32, False, 0.2)])
def test_stage_forces_parity(cuda, hiplib, N, obst, noise):
    _say(f"stage forces parity N={N} obstacle={obst} noise={noise}")
    solid = sphere_solid(N) if obst else np.zeros((N,) * 3, np.float32)
    data = random_state(N, 31, solid, vel_amp=2.0)
    cfg = SolverConfig3D(burn_rate=4.0, ignition_temp=450.0, emitter_noise=noise, noise_seed=5)
    prm = pack_params(cfg, N, FIRE, 1.0 / 60.0, 7)
    s = Solver3D(N, cfg)
    d = to_dev(data)
    sd = to_dev(solid)
    for st in (_lib.STAGE_ADVECT, _lib.STAGE_FORCES):
        _lib.check(hiplib.vfx_step_stage(s._ctx, st, _ptr(d), _ptr(sd), prm.ctypes.data_as(ctypes.c_void_p),
                                         _stream()))
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    ref_adv, ratio = oracle.stage_advect(data, solid, prm, np.float32(N))
    ref = oracle.stage_forces(ref_adv, solid, prm, ratio, np.float32(N))
    for c, name in enumerate(("density", "vx", "vy", "vz", "T", "fuel")):
        e = rel_err(got[c], ref[c])
        _say(f"  {name}: max rel err {e:.3e}")
        assert e <= 1e-5, name


@pytest.mark.parametrize("N", [16, 32, 48])
def test_operator_parity(cuda, hiplib, N):
    """Face mask, K p and the level operators (probe + Galerkin) are identical."""
    _say(f"operators N={N}")
    solid = sphere_solid(N, r=0.2)
    s = Solver3D(N)
    sd = to_dev(solid)
    p = smooth_noise(N, 1, 41)[0]
    out = torch.empty(N, N, N, device=DEV)
    pd = to_dev(p)
    _lib.check(hiplib.vfx_kmul(s._ctx, _ptr(sd), _ptr(pd), _ptr(out), _stream()))
    fm = oracle.mask(solid)
    ref = oracle.kmul(fm, p)
    assert rel_err(out.cpu().numpy(), ref) <= 1e-6
    ops = oracle.build_ops(fm)
    nlev = ctypes.c_int()
    m = ctypes.c_int()
    for lvl, (dref, oref) in enumerate(ops):
        mm = dref.shape[1]
        dg = torch.empty(mm ** 3 * 8, device=DEV)
        of = torch.empty(mm ** 3 * 24, device=DEV)
        _lib.check(hiplib.vfx_get_ops(s._ctx, _ptr(sd), lvl, _ptr(dg), _ptr(of), ctypes.byref(nlev), ctypes.byref(m),
                                      _stream()))
        torch.cuda.synchronize()
        assert m.value == mm and nlev.value == len(ops)
        ed = rel_err(soa_from_aos(dg.cpu().numpy(), mm, 8), dref)
        eo = rel_err(soa_from_aos(of.cpu().numpy(), mm, 24), oref)
        _say(f"  level {lvl} m={mm}: diag {ed:.2e} off {eo:.2e}")
        assert ed <= 1e-6 and eo <= 1e-6


@pytest.mark.parametrize("N,coarse_m", [(32, 8), (48, 8), (64, 8), (64, 2)])
def test_precond_parity(cuda, hiplib, N, coarse_m):
    """z = M^-1 r (Jacobi, V(1,1) in the Hadamard basis, Jacobi) equals the oracle's."""
    _say(f"precond N={N} coarse_m={coarse_m}")
    solid = sphere_solid(N)
    s = Solver3D(N)
    s.set_option(_lib.OPT_COARSE_M, coarse_m)
    fm = oracle.mask(solid)
    r = smooth_noise(N, 1, 51)[0] * (fm > 0)
    z = torch.empty(N, N, N, device=DEV)
    sd, rd = to_dev(solid), to_dev(r)     # keep both alive until the copies have run
    _lib.check(hiplib.vfx_precond(s._ctx, _ptr(sd), _ptr(rd), _ptr(z), _stream()))
    torch.cuda.synchronize()
    zref, _ = oracle.precond_apply(fm, r)
    e = rel_err(z.cpu().numpy(), zref)
    _say(f"  max rel err {e:.3e}")
    assert e <= 1e-6


def test_div6_equals_division_on_every_float(cuda, hiplib):
    """The CG kernels divide by diag(K) = 6 at INTERIOR cells with a multiply and one fused
    correction instead of the division sequence: on the device, for all 2^32 single-precision
    inputs, the result equals the IEEE quotient x / 6 bit for bit (NaN for NaN)."""
    cnt = torch.zeros(1, dtype=torch.int64, device=DEV)
    _lib.check(hiplib.vfx_check_div6(_ptr(cnt), _stream()))
    torch.cuda.synchronize()
    _say(f"  div6 mismatches over 2^32 inputs: {int(cnt.item())}")
    assert int(cnt.item()) == 0


def test_sealed_fluid_cell_parity(cuda, hiplib):
    """A fluid cell sealed in by solids (diag(K) = 0): the preconditioner equals the oracle's, with
    0 in that cell, and the projection converges in the oracle's iterations."""
    N, c = 32, (9, 20, 11)
    _say(f"sealed fluid cell N={N} at {c}")
    solid = sealed_cell_solid(N, c)
    fm = oracle.mask(solid)
    assert fm[c] == 64
    r = smooth_noise(N, 1, 52)[0] * (fm > 0)
    s = Solver3D(N)
    z = torch.empty(N, N, N, device=DEV)
    sd, rd = to_dev(solid), to_dev(r)
    _lib.check(hiplib.vfx_precond(s._ctx, _ptr(sd), _ptr(rd), _ptr(z), _stream()))
    torch.cuda.synchronize()
    zg = z.cpu().numpy()
    zref, _ = oracle.precond_apply(fm, r)
    e = rel_err(zg, zref)
    _say(f"  precond max rel err {e:.3e}, z at the cell {zg[c]}")
    assert np.isfinite(zg).all() and zg[c] == 0.0 and e <= 1e-6
    data = _divergent_state(N, solid, seed=5)
    cfg = SolverConfig3D(pressure_rel_tol=1e-5, pressure_max_iters=200, pressure_warm_start=False)
    s2 = Solver3D(N, cfg)
    s2.set_option(_lib.OPT_TRUE_RESIDUAL, 1)
    st = s2.reset()
    st.data.copy_(to_dev(data))
    st.solid.copy_(to_dev(solid))
    s2.project(st)
    stats = s2.pressure_stats()
    ref = data.copy()
    it_ref, _ = oracle.project(ref, solid, pack_params(cfg, N, Forces3D(), 0.0), np.zeros((N,) * 3, np.float32))
    _say(f"  iters gpu {stats['iters']} oracle {it_ref}; true rel-res {stats['true_rel_res']:.2e}")
    assert it_ref > 0 and abs(stats["iters"] - it_ref) <= 1 and stats["true_rel_res"] <= 1e-5
    assert rel_err(st.data.cpu().numpy()[1:4], ref[1:4]) <= 1e-4


@pytest.mark.parametrize("N", [64, 128])
def test_precond_standard_stencil_identical(cuda, hiplib, N):
    """Waves of standard cells take their multigrid coefficients from the per-level table; the
    preconditioner is bit for bit the one that reads every coefficient from the level arrays."""
    _say(f"precond standard-stencil on/off N={N}")
    solid = sphere_solid(N)
    fm = oracle.mask(solid)
    r = smooth_noise(N, 1, 7)[0] * (fm > 0)
    sd, rd = to_dev(solid), to_dev(r)
    outs = []
    for on in (1, 0):
        s = Solver3D(N)
        s.set_option(_lib.OPT_STD_STENCIL, on)
        z = torch.empty(N, N, N, device=DEV)
        _lib.check(hiplib.vfx_precond(s._ctx, _ptr(sd), _ptr(rd), _ptr(z), _stream()))
        torch.cuda.synchronize()
        outs.append(z.cpu())
    assert torch.equal(outs[0], outs[1])
    assert float(outs[0].abs().max()) > 0.0


def _divergent_state(N, solid, seed=1):
    data = random_state(N, seed, solid)
    v = smooth_noise(N, 3, seed, amp=2.0) + np.array([0.8, -0.3, 0.4], np.float32)[:, None, None, None]
    data[1:4] = v * (solid < 0.5)
    return np.ascontiguousarray(data)


@pytest.mark.parametrize("N,obst", [(32, True), (48, True), (64, False)])
def test_project_parity_fixed_iters(cuda, hiplib, N, obst):
    """Same number of PCG iterations on both sides: velocities agree to round-off."""
    _say(f"project fixed iters N={N} obstacle={obst}")
    solid = sphere_solid(N) if obst else np.zeros((N,) * 3, np.float32)
    data = _divergent_state(N, solid)
    cfg = runtime_config(pressure_max_iters=6, pressure_warm_start=False)
    s = Solver3D(N, cfg)
    st = s.reset()
    st.data.copy_(to_dev(data))
    st.solid.copy_(to_dev(solid))
    s.project(st)
    it_gpu = s.pressure_iters()
    got = st.data.cpu().numpy()
    ref = data.copy()
    phi = np.zeros((N,) * 3, np.float32)
    prm = pack_params(cfg, N, Forces3D(), 0.0)
    it_ref, _ = oracle.project(ref, solid, prm, phi)
    e = rel_err(got[1:4], ref[1:4])
    _say(f"  iters gpu {it_gpu} oracle {it_ref}; velocity max rel err {e:.3e}")
    assert it_gpu == it_ref == 6
    assert e <= 1e-4


@pytest.mark.parametrize("N,obst", [(32, True), (64, True)])
def test_project_rel_res_and_fp64(cuda, hiplib, N, obst):
    """Projection to ||r||/||b|| <= 1e-5: both sides converge in about the same number of
    iterations, the true residual of the HIP result meets the tolerance, and the velocity is
    close to the FP64 oracle's."""
    _say(f"project rel-res N={N}")
    solid = sphere_solid(N) if obst else np.zeros((N,) * 3, np.float32)
    data = _divergent_state(N, solid, seed=3)
    cfg = SolverConfig3D(pressure_rel_tol=1e-5, pressure_max_iters=200, pressure_warm_start=False,
                         pressure_restarts=2)
    s = Solver3D(N, cfg)
    s.set_option(_lib.OPT_TRUE_RESIDUAL, 1)
    st = s.reset()
    st.data.copy_(to_dev(data))
    st.solid.copy_(to_dev(solid))
    s.project(st)
    stats = s.pressure_stats()
    got = st.data.cpu().numpy()
    prm = pack_params(cfg, N, Forces3D(), 0.0)
    ref32 = data.copy()
    it32, _ = oracle.project(ref32, solid, prm, np.zeros((N,) * 3, np.float32))
    ref64 = data.astype(np.float64)
    it64, st64 = oracle.project(ref64, solid, prm, np.zeros((N,) * 3, np.float64), prec="f64")
    e64 = rel_err(got[1:4], ref64[1:4])
    _say(f"  iters gpu {stats['iters']} oracle32 {it32} oracle64 {it64}; true rel-res {stats['true_rel_res']:.2e};"
         f" vs fp64 {e64:.2e}; restarts {stats['restarts']}")
    assert abs(stats["iters"] - it32) <= 1
    assert stats["true_rel_res"] <= 1e-5
    assert e64 <= 1e-4


@pytest.mark.parametrize("N", [32, 48])
def test_step_parity(cuda, hiplib, N):
    """One whole step (fixed 5 PCG iterations) equals the oracle's step."""
    _say(f"full step N={N}")
    solid = sphere_solid(N)
    data = random_state(N, 61, solid, vel_amp=1.5)
    cfg = runtime_config(pressure_max_iters=5)
    s = Solver3D(N, cfg)
    st = s.reset()
    st.data.copy_(to_dev(data))
    st.solid.copy_(to_dev(solid))
    s.step(st, FIRE, 1.0 / 60.0)
    got = st.data.cpu().numpy()
    prm = pack_params(cfg, N, FIRE, 1.0 / 60.0, 0)
    ref, it = oracle.step(data, solid, prm, np.zeros((N,) * 3, np.float32), np.float32(N))
    assert s.pressure_iters() == it == 5
    for c, name in enumerate(("density", "vx", "vy", "vz", "T", "fuel")):
        e = rel_err(got[c], ref[c])
        _say(f"  {name}: {e:.3e}")
        assert e <= 1e-4, name


def test_divergence_parity(cuda, hiplib):
    N = 32
    solid = sphere_solid(N)
    data = _divergent_state(N, solid)
    s = Solver3D(N)
    st = s.reset()
    st.data.copy_(to_dev(data))
    st.solid.copy_(to_dev(solid))
    got = s.divergence(st).cpu().numpy()
    fm = oracle.mask(solid)
    da = oracle.div_abs(fm, data[1:4])
    sp = np.sqrt((data[1:4].astype(np.float64) ** 2).sum(0))
    ref = da / (sp[solid < 0.5].max() + 1e-8)
    assert rel_err(got, ref) <= 1e-5
    assert np.all(got[solid > 0.5] == 0)

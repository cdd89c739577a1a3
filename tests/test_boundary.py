"""The drop-in boundary on the CPU: the C-ABI library loads into torch's HIP runtime and exports
every entry point of include/vfxsim3d.h; the Python layer keeps the reference's API, parameter
layout and error behaviour (vfx/sim3d/__init__.py:1-7, solver.py:30-37, :331-359, config.py)."""
from __future__ import annotations

import dataclasses
import os
import re

import numpy as np
import pytest

from vfx.core.fields3d import Emitter3D, Forces3D, Impulse3D, Obstacle3D, rasterize_obstacles_3d
import vfx.sim3d as sim
from vfx.sim3d.solver import P_SIZE, pack_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "vfxsim3d.h")).read()
    return sorted(set(re.findall(r"^int (vfx_\w+)\(", txt, flags=re.M)))


def test_library_exports_every_header_symbol(hiplib):
    syms = _header_symbols()
    assert len(syms) >= 15
    for name in syms:
        assert hasattr(hiplib, name), name
    from vfx.sim3d import _lib
    assert set(syms) == set(_lib.SIGNATURES), "ctypes signatures and the header disagree"


def test_advect_noalloc_rejects_grids_above_its_scratch(hiplib):
    """vfx_advect_noalloc checks the grid against its fixed scratch (2^20 cells) before any HIP call."""
    import ctypes
    from vfx.sim3d import _lib
    fake = ctypes.c_void_p(256)   # never dereferenced: the size check comes first
    rc = hiplib.vfx_advect_noalloc(fake, 6, fake, fake, 128, ctypes.c_float(0.1), None)
    assert rc == _lib.VFX_E_VALUE and "no-allocation scratch" in _lib.last_error()
    rc = hiplib.vfx_advect_noalloc(fake, 6, fake, fake, 7, ctypes.c_float(0.1), None)
    assert rc == _lib.VFX_E_ARG


def test_library_shares_torch_hip_runtime(hiplib):
    maps = open("/proc/self/maps").read()
    rts = sorted(set(re.findall(r"/\S*libamdhip64\S*", maps)))
    assert len(rts) == 1 and "torch" in rts[0], rts


def test_p_size_and_error_string(hiplib):
    assert hiplib.vfx_p_size() == P_SIZE == 576
    import ctypes
    rc = hiplib.vfx_ctx_create(7, ctypes.byref(ctypes.c_void_p()))   # odd size: rejected before any HIP call
    assert rc == 1
    from vfx.sim3d import _lib
    assert "even" in _lib.last_error()
    with pytest.raises(AssertionError):
        _lib.check(rc)
    # above 512 the kernels' 32-bit channel offsets (c N^3 + i) would overflow
    rc = hiplib.vfx_ctx_create(1024, ctypes.byref(ctypes.c_void_p()))
    assert rc == 1 and "512" in _lib.last_error()


def test_api_surface():
    for name in ("Solver3D", "SolverConfig3D", "SOLVER3D_VERSION", "SOLVER3D_REVISION", "runtime_config",
                 "fused_config", "fused_fixed_config", "solver_version_tag"):
        assert hasattr(sim, name), name
    assert sim.SOLVER3D_VERSION == "sim3d-warp-0.1.0"
    cfg = sim.SolverConfig3D()
    ref_defaults = dict(buoyancy_alpha=0.6, buoyancy_beta=0.004, vorticity_eps=6.0, density_dissipation=0.05,
                        temperature_cooling=4.0, burn_rate=8.0, ignition_temp=550.0, burn_heat=4000.0,
                        smoke_from_fuel=0.35, pressure_tol=2e-5, pressure_max_iters=40, pressure_warm_start=True,
                        max_temperature=3000.0, wind_marker_temp=300.0, heat_diffusion=5.0e-5,
                        confinement_floor=0.5, confinement_ref_res=0, pressure_mode="pcg")
    for k, v in ref_defaults.items():
        assert getattr(cfg, k) == v, k
    assert sim.runtime_config().pressure_max_iters == 4 and sim.runtime_config().pressure_tol == 0.0
    assert sim.solver_version_tag(cfg) == "sim3d-warp-0.1.0"
    assert dataclasses.replace(cfg, ignition_ramp=100.0) == cfg


def test_pack_params_layout():
    cfg = sim.SolverConfig3D(confinement_ref_res=128)
    f = Forces3D(emitters=[Emitter3D(pos=(0.1, 0.2, 0.3), radius=0.0, density_rate=2.0, temperature=900.0,
                                     fuel_rate=1.5, velocity=(0.0, 1.0, 2.0))],
                 impulses=[Impulse3D(pos=(0.4, 0.5, 0.6), radius=0.1, velocity=(3.0, 4.0, 5.0))], wind=(0.2, 0.0, 0.1))
    h = pack_params(cfg, 64, f, 1 / 30, 3)
    assert h.dtype == np.float32 and h.shape == (576,)
    assert h[0] == np.float32(1 / 30) and tuple(h[1:4]) == (np.float32(0.2), 0.0, np.float32(0.1))
    assert h[4] == 1 and h[5] == 1 and h[6] == np.float32(1 / 64) and h[7] == 293.0
    assert h[18] == np.float32(6.0 * 64 / 128)          # confinement_ref_res scaling (solver.py:345-346)
    assert tuple(h[32:42]) == tuple(np.float32([0.1, 0.2, 0.3, 1e-6, 2.0, 900.0, 1.5, 0.0, 1.0, 2.0]))
    b = 32 + 32 * 10
    assert tuple(h[b:b + 7]) == tuple(np.float32([0.4, 0.5, 0.6, 0.1, 3.0, 4.0, 5.0]))
    assert h[11] == 100.0 and h[12] == 1.0 and h[13] == 1.0 and h[15] == 3.0   # extension slots: defaults
    with pytest.raises(ValueError):
        pack_params(cfg, 64, Forces3D(emitters=[Emitter3D((0, 0, 0), 0.1)] * 33), 0.1)


def test_rasterize_matches_reference_rule():
    obs = [Obstacle3D("sphere", (0.5, 0.45, 0.5), (0.14, 0.14, 0.14)),
           Obstacle3D("box", (0.3, 0.7, 0.6), (0.1, 0.05, 0.08))]
    s = rasterize_obstacles_3d(obs, 32, device="cpu").numpy()
    c = (np.arange(32) + 0.5) / 32
    zz, yy, xx = np.meshgrid(c, c, c, indexing="ij")
    sph = (xx - 0.5) ** 2 + (yy - 0.45) ** 2 + (zz - 0.5) ** 2 <= 0.14 ** 2
    box = (np.abs(xx - 0.3) <= 0.1) & (np.abs(yy - 0.7) <= 0.05) & (np.abs(zz - 0.6) <= 0.08)
    assert np.array_equal(s > 0.5, sph | box)
    with pytest.raises(ValueError):
        rasterize_obstacles_3d([Obstacle3D("cone", (0, 0, 0), (1, 1, 1))], 8, device="cpu")


def test_bad_modes_fail_loudly():
    # scalar_res below res is the reference's ValueError (solver.py:219-220), raised before any
    # device work
    with pytest.raises(ValueError):
        sim.Solver3D(32, scalar_res=16)
    with pytest.raises(ValueError):
        sim.Solver3D(32, scalar_res=2048)
    with pytest.raises(ValueError):
        sim.Solver3D(32, sim.SolverConfig3D(pressure_mode="bogus"))
    # the fused mode is built; its smoother name is validated before any device work
    with pytest.raises(ValueError):
        sim.Solver3D(32, sim.fused_config(pressure_smoother="bogus"))

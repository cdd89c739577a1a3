"""3D smoke + fire solver on MI355X: the reference's Solver3D API (vfx/sim3d/solver.py:79-459)
over the HIP library libvfxsim3d.so.

Step order (the reference's, solver.py:3-8):
  1. advect density, velocity, temperature, fuel by u^n (MacCormack, RK2 traces, clamped)
  2. density mass correction (global fluid-density ratio, fixed-order FP64 reduction)
     [scalar_res: density, temperature and fuel are advected, corrected, sourced and diffused on
     the finer grid instead, and box-restricted into the state before step 4]
  3. emitters, impulses, combustion, cooling, dissipation; heat + fuel diffusion
  4. buoyancy, vorticity confinement on the smoothed vorticity, wind; the [1,2,1]^3-smoothed
     force is applied
  5. pressure projection (MG-preconditioned CG, exact for the face-averaged central divergence)

The state tensors stay owned by the caller and are updated in place on torch's current stream.
Per-step inputs (dt, forces, config) are packed into the reference's parameter layout and
uploaded through a pinned ring inside the library.  Masks and multigrid operators are rebuilt on
the device only when `solid` changed.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np
import torch

from vfx.core.fields import AMBIENT_TEMP, DOMAIN_SIZE
from vfx.core.fields3d import NUM_CHANNELS_3D, FieldMeta3D, FieldState3D, Forces3D, Volume, rasterize_obstacles_3d

from . import _lib
from .config import SolverConfig3D, solver_version_tag

# parameter-buffer slots (kernels.py:11-42) and the extension slots of this build
P_DT, P_WIND_X, P_WIND_Y, P_WIND_Z, P_N_EM, P_N_IMP, P_DX, P_AMB = 0, 1, 2, 3, 4, 5, 6, 7
P_KAPPA, P_CFLOOR, P_WARM = 8, 9, 10
P_IGN_RAMP, P_NOISE_LO, P_NOISE_HI, P_SEED, P_STEPNO = 11, 12, 13, 14, 15
P_ALPHA, P_BETA, P_EPS, P_DDISS, P_COOL, P_BURN, P_IGN, P_BHEAT, P_SMOKE = 16, 17, 18, 19, 20, 21, 22, 23, 24
P_TOL, P_MAXIT, P_TMAX, P_WMARK, P_REL_TOL = 25, 26, 27, 28, 29
P_EM0, EM_STRIDE, MAX_EMITTERS = 32, 10, 32
P_IMP0 = P_EM0 + MAX_EMITTERS * EM_STRIDE
IMP_STRIDE, MAX_IMPULSES = 7, 32
P_SIZE = P_IMP0 + MAX_IMPULSES * IMP_STRIDE

# config field -> slot (solver.py:30-37)
CFG_SLOTS = {
    "buoyancy_alpha": P_ALPHA, "buoyancy_beta": P_BETA, "vorticity_eps": P_EPS,
    "density_dissipation": P_DDISS, "temperature_cooling": P_COOL, "burn_rate": P_BURN,
    "ignition_temp": P_IGN, "burn_heat": P_BHEAT, "smoke_from_fuel": P_SMOKE,
    "pressure_tol": P_TOL, "pressure_max_iters": P_MAXIT, "max_temperature": P_TMAX,
    "wind_marker_temp": P_WMARK, "heat_diffusion": P_KAPPA, "confinement_floor": P_CFLOOR,
    "pressure_warm_start": P_WARM, "ignition_ramp": P_IGN_RAMP, "pressure_rel_tol": P_REL_TOL,
}


def pack_params(cfg: SolverConfig3D, res: int, forces: Forces3D, dt: float, step: int = 0,
                out: np.ndarray | None = None) -> np.ndarray:
    """The per-step parameter buffer, filled as Solver3D._upload does (solver.py:331-359)."""
    h = np.zeros(P_SIZE, np.float32) if out is None else out
    h[:] = 0.0
    h[P_DT] = dt
    h[P_WIND_X], h[P_WIND_Y], h[P_WIND_Z] = forces.wind
    h[P_DX] = DOMAIN_SIZE / res
    h[P_AMB] = AMBIENT_TEMP
    for name, slot in CFG_SLOTS.items():
        h[slot] = float(getattr(cfg, name))
    if cfg.confinement_ref_res:
        h[P_EPS] = float(cfg.vorticity_eps) * res / float(cfg.confinement_ref_res)
    a = float(cfg.emitter_noise)
    h[P_NOISE_LO], h[P_NOISE_HI] = 1.0 - a, 1.0 + a
    h[P_SEED] = float(int(cfg.noise_seed) & 0xFFFFFF)
    h[P_STEPNO] = float(int(step) & 0xFFFFFF)
    if len(forces.emitters) > MAX_EMITTERS or len(forces.impulses) > MAX_IMPULSES:
        raise ValueError(f"at most {MAX_EMITTERS} emitters / {MAX_IMPULSES} impulses")
    h[P_N_EM] = len(forces.emitters)
    for j, e in enumerate(forces.emitters):
        b = P_EM0 + j * EM_STRIDE
        h[b:b + EM_STRIDE] = (*e.pos, max(e.radius, 1e-6), e.density_rate, e.temperature, e.fuel_rate, *e.velocity)
    h[P_N_IMP] = len(forces.impulses)
    for j, m in enumerate(forces.impulses):
        b = P_IMP0 + j * IMP_STRIDE
        h[b:b + IMP_STRIDE] = (*m.pos, max(m.radius, 1e-6), *m.velocity)
    return h


def _ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


class Solver3D:
    HOST_RING = 16
    # replay each step as a captured hipGraph (device-sequenced projections; the reference's
    # step is one CUDA graph, solver.py:388-420); False: direct launches
    STEP_GRAPH = True

    def __init__(self, res: int, config: SolverConfig3D | None = None, device="cuda", scalar_res: int | None = None,
                 dtype: torch.dtype = torch.float32):
        """scalar_res (opt-in, S >= res): density, temperature and fuel are also carried on a finer
        S^3 grid by the res^3 velocity and box-restricted back each step (hires.py).
        `dtype` torch.float64 selects the FP64 run of the same solver (double state, levels and
        constants); the reference is FP32 only, so float32 is the reference-compatible default."""
        if dtype not in (torch.float32, torch.float64):
            raise ValueError(f"dtype must be torch.float32 or torch.float64, got {dtype}")
        self.dtype = dtype
        self.res = int(res)
        self.config = config if config is not None else SolverConfig3D()
        self.scalar_res = None if scalar_res is None else int(scalar_res)
        if self.scalar_res is not None and self.scalar_res < self.res:
            raise ValueError("scalar_res must be >= res")
        if self.scalar_res is not None and self.scalar_res > _lib.HIRES_MAX:
            raise ValueError(f"scalar_res must be <= {_lib.HIRES_MAX}")
        self._structure()
        self.device = str(device)
        if self.device == "cuda":
            self.device = f"cuda:{torch.cuda.current_device()}"
        self.dx = DOMAIN_SIZE / self.res
        self._lib = _lib.load()
        self._f64 = dtype == torch.float64
        self._api = _lib.Api(self._lib, self._f64)
        self._ctx = ctypes.c_void_p()
        with torch.cuda.device(torch.device(self.device)):
            self._chk(self._api.vfx_ctx_create(self.res, ctypes.byref(self._ctx)), "Solver3D")
        self._h = np.zeros(P_SIZE, np.float32)
        self._net = None
        self._attached = None
        self._apply_options()
        if self.scalar_res is not None:
            self._init_hires()

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx is not None and ctx.value and getattr(self, "_api", None) is not None:
            self._api.vfx_ctx_destroy(ctx)
            self._ctx = None

    def _chk(self, rc: int, what: str = "") -> None:
        _lib.check(rc, what, self._f64)

    def _apply_options(self):
        self._chk(self._api.vfx_set_option(self._ctx, _lib.OPT_RESTARTS, float(self.config.pressure_restarts)))
        sync = _lib.SYNC_MODES.get(self.config.pressure_sync)
        if sync is None:
            raise ValueError(f"pressure_sync must be one of {sorted(_lib.SYNC_MODES)}, got {self.config.pressure_sync!r}")
        self._chk(self._api.vfx_set_option(self._ctx, _lib.OPT_SYNC, float(sync)))
        self._chk(self._api.vfx_set_option(self._ctx, _lib.OPT_GRAPH, 1.0 if self.STEP_GRAPH else 0.0))
        if not self._f64:
            # the network runs only in pressure_mode "learned"; it stays attached on the Python side
            want = self._net if self.config.pressure_mode == "learned" else None
            if want is not self._attached:
                self._chk(self._lib.vfx_ctx_set_net(self._ctx, want._h if want is not None else None))
                self._attached = want
        kind = _lib.MG_STORAGE.get(self.config.mg_storage)
        if kind is None:
            raise ValueError(f"mg_storage must be one of {sorted(_lib.MG_STORAGE)}, got {self.config.mg_storage!r}")
        self._chk(self._api.vfx_set_option(self._ctx, _lib.OPT_MG_STORAGE, float(kind)))
        c = self.config
        fused = c.pressure_mode == "fused"
        self._chk(self._api.vfx_set_option(self._ctx, _lib.OPT_PRESSURE_MODE, 1.0 if fused else 0.0))
        if fused:
            self._chk(self._api.vfx_set_option(self._ctx, _lib.OPT_FUSED_CYCLES, float(max(0, int(c.pressure_cycles)))))
            self._chk(self._api.vfx_set_option(self._ctx, _lib.OPT_FUSED_SMOOTH, float(max(1, int(c.pressure_smooth)))))
            self._chk(self._api.vfx_set_option(self._ctx, _lib.OPT_FUSED_SMOOTHER,
                                               1.0 if c.pressure_smoother == "jacobi" else 0.0))

    def set_option(self, key: int, value: float) -> None:
        """Library option (vfxsim3d.h VFX_OPT_*): true-residual report, restarts, coarse size."""
        self._chk(self._api.vfx_set_option(self._ctx, int(key), float(value)))

    def set_preconditioner(self, net) -> None:
        """Attach a learned preconditioner (vfx.sim3d.learned.UNet of this grid size) and switch
        config.pressure_mode to "learned", or detach it (None) and switch back to "pcg": the
        projection then runs flexible CG with the network, falling back to plain MGPCG when the
        residual reduction stalls (pressure_stats()["fallback"])."""
        if self._f64:
            raise ValueError("the learned preconditioner runs in the FP32 solver")
        if net is not None and net.res not in (self.res, self.res // 4):
            raise ValueError(f"network grid {net.res}: the solver grid {self.res} (the preconditioner) or a "
                             f"quarter of it (the coarse correction of the V-cycle)")
        self._net = net
        if net is not None:
            self.config = dataclasses.replace(self.config, pressure_mode="learned")
        elif self.config.pressure_mode == "learned":
            self.config = dataclasses.replace(self.config, pressure_mode="pcg")
        self._apply_options()

    @property
    def version(self) -> str:
        return solver_version_tag(self.config, self.scalar_res)

    # ------------------------------------------------------------------ public API
    def reset(self, obstacles=()) -> FieldState3D:
        N = self.res
        solid = rasterize_obstacles_3d(list(obstacles), N, device=self.device).to(self.dtype)
        data = torch.zeros(NUM_CHANNELS_3D, N, N, N, device=self.device, dtype=self.dtype)
        data[4] = AMBIENT_TEMP
        meta = FieldMeta3D.for_res(N)
        meta.solver_version = self.version
        if self.scalar_res is not None:
            self.set_hires_obstacles(obstacles)
            self.hi.zero_()
            self.hi[1] = AMBIENT_TEMP
        return FieldState3D(data=data, solid=solid, meta=meta)

    def step(self, state: FieldState3D, forces: Forces3D, dt: float) -> FieldState3D:
        self._check(state)
        self._structure()
        self._apply_options()
        h = pack_params(self.config, self.res, forces, dt, state.meta.step, self._h)
        self._chk(self._api.vfx_step(self._ctx, _ptr(state.data), _ptr(state.solid),
                                      h.ctypes.data_as(ctypes.c_void_p), self._stream()), "step")
        state.meta.step += 1
        state.meta.dt = float(dt)
        state.meta.solver_version = self.version
        return state

    def step_profiled(self, state: FieldState3D, forces: Forces3D, dt: float) -> dict:
        """Same step as step(), as three stages with device events between them (four with
        scalar_res: the fine-scalar stage is timed as `scalars`) (solver.py:157-184).  Returns ms
        per stage {advection, [scalars,] forces, projection, total}."""
        self._check(state)
        self._structure()
        self._apply_options()
        h = pack_params(self.config, self.res, forces, dt, state.meta.step, self._h)
        stages = [(_lib.STAGE_ADVECT, "advection"), (_lib.STAGE_FORCES, "forces"), (_lib.STAGE_PROJECT, "projection")]
        if self.scalar_res is not None:
            stages.insert(1, (_lib.STAGE_SCALARS, "scalars"))
        ts = torch.cuda.current_stream(torch.device(self.device))
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(len(stages) + 1)]
        torch.cuda._sleep(3_000_000)
        ev[0].record(ts)
        hp = h.ctypes.data_as(ctypes.c_void_p)
        for k, (stage, _) in enumerate(stages):
            self._chk(self._api.vfx_step_stage(self._ctx, stage, _ptr(state.data), _ptr(state.solid),
                                                hp if stage == _lib.STAGE_ADVECT else None, self._stream()),
                       "step_profiled")
            ev[k + 1].record(ts)
        ev[-1].synchronize()
        state.meta.step += 1
        state.meta.dt = float(dt)
        state.meta.solver_version = self.version
        t = [ev[i].elapsed_time(ev[i + 1]) for i in range(len(stages))]
        out = dict(zip((name for _, name in stages), t))
        out["total"] = sum(t)
        return out

    # ------------------------------------------------------------------ finer scalar grid (opt-in)
    def hires_volume(self) -> Volume | None:
        """The fine density + temperature (views of the solver's buffer, no copy; valid until the
        next step), or None without scalar_res (solver.py:187-192)."""
        if self.scalar_res is None:
            return None
        return Volume(density=self.hi[0], temperature=self.hi[1])

    def hires_fields(self) -> torch.Tensor | None:
        """(3, S, S, S) density, temperature, fuel (the live buffer), or None."""
        return None if self.scalar_res is None else self.hi

    def set_hires_obstacles(self, obstacles=()) -> None:
        """Rasterise obstacles at scalar_res (reset() does this).  Call it whenever state.solid is
        changed by hand, with the same obstacles: the fine grid reads hi_solid, not state.solid."""
        if self.scalar_res is not None:
            self.hi_solid.copy_(rasterize_obstacles_3d(list(obstacles), self.scalar_res, device=self.device))

    def init_hires_from_state(self, state: FieldState3D) -> None:
        """Fill the fine scalars from the coarse ones (trilinear, cell centred; 0 / ambient in fine
        solids), on the device.  For states built by hand; reset() already starts both grids empty."""
        if self.scalar_res is None:
            return
        self._check(state)
        self._chk(self._api.vfx_hires_init(self._ctx, _ptr(state.data), float(AMBIENT_TEMP), self._stream()),
                  "init_hires_from_state")

    def _init_hires(self):
        S = self.scalar_res
        self.dxf = DOMAIN_SIZE / S
        self.hi = torch.zeros(3, S, S, S, device=self.device, dtype=self.dtype)
        self.hi[1] = AMBIENT_TEMP
        self.hi_solid = torch.zeros(S, S, S, device=self.device, dtype=self.dtype)
        self._chk(self._api.vfx_hires_attach(self._ctx, S, _ptr(self.hi), _ptr(self.hi_solid)), "scalar_res")

    def project(self, state: FieldState3D) -> FieldState3D:
        self._check(state)
        self._structure()
        self._apply_options()
        h = pack_params(self.config, self.res, Forces3D(), 0.0, state.meta.step, self._h)
        self._chk(self._api.vfx_project(self._ctx, _ptr(state.data), _ptr(state.solid),
                                         h.ctypes.data_as(ctypes.c_void_p), self._stream()), "project")
        return state

    def divergence(self, state: FieldState3D) -> torch.Tensor:
        """(D,H,W) |div u| dx / (max|u| + 1e-8), fluid cells only (solver.py:274-287)."""
        self._check(state)
        N = self.res
        out = torch.zeros(N, N, N, device=self.device, dtype=self.dtype)
        self._chk(self._api.vfx_div_abs(self._ctx, _ptr(state.data), _ptr(state.solid), _ptr(out),
                                         self._stream()), "divergence")
        fluid = state.solid < 0.5
        speed = state.velocity.norm(dim=0)
        umax = torch.where(fluid, speed, torch.zeros_like(speed)).max()
        return out / (umax + 1e-8)

    def advect(self, field: torch.Tensor, velocity: torch.Tensor, dt: float) -> torch.Tensor:
        """MacCormack advection without mass correction or obstacles (solver.py:289-308)."""
        squeeze = field.dim() == 3
        f = field.unsqueeze(0) if squeeze else field
        N = self.res
        assert f.shape[-3:] == (N, N, N) and velocity.shape == (3, N, N, N)
        assert f.is_cuda and velocity.is_cuda
        f = f.contiguous().to(self.dtype)
        v = velocity.contiguous().to(self.dtype)
        out = torch.empty_like(f)
        self._chk(self._api.vfx_advect(self._ctx, _ptr(f), int(f.shape[0]), _ptr(v), _ptr(out), float(dt),
                                        self._stream()), "advect")
        return out[0] if squeeze else out

    def pressure_iters(self) -> int:
        """PCG iterations (default) or V-cycles (fused) of the last projection (synchronizes)."""
        n = ctypes.c_int()
        self._chk(self._api.vfx_pressure_iters(self._ctx, ctypes.byref(n)))
        return int(n.value)

    def pressure_stats(self) -> dict:
        """Diagnostics of the last projection (synchronizes): iterations, max|u*|, ||b||, the
        recurrence ||r||, max|r|, the true ||b - K phi|| (if enabled), restarts, whether it
        converged (judged by the true residual when computed) and its first CG run's iteration
        budget; over the solver's life so far: projections that stopped at a budget unconverged,
        unconverged projections, and step graphs captured / replayed."""
        v = (ctypes.c_double * _lib.STATS)()
        self._chk(self._api.vfx_pressure_stats(self._ctx, v, _lib.STATS))
        keys = ("iters", "max_u", "b_norm", "r_norm", "max_r", "true_r_norm", "restarts", "learned_iters", "fallback",
                "converged", "budget", "budget_stops", "unconverged", "graph_captures", "graph_replays")
        d = dict(zip(keys, list(v)))
        for k in ("iters", "restarts", "learned_iters", "fallback", "budget", "budget_stops", "unconverged",
                  "graph_captures", "graph_replays"):
            d[k] = int(d[k])
        d["converged"] = bool(d["converged"])
        d["rel_res"] = d["r_norm"] / d["b_norm"] if d["b_norm"] > 0 else 0.0
        d["true_rel_res"] = d["true_r_norm"] / d["b_norm"] if d["b_norm"] > 0 and d["true_r_norm"] >= 0 else -1.0
        return d

    def fcg_history(self) -> list:
        """||r_k|| of each flexible-CG iteration of the last projection (learned path)."""
        n = ctypes.c_int()
        self._chk(self._lib.vfx_fcg_history(self._ctx, None, 0, ctypes.byref(n)))
        buf = (ctypes.c_double * max(n.value, 1))()
        self._chk(self._lib.vfx_fcg_history(self._ctx, buf, n.value, ctypes.byref(n)))
        return list(buf)[:n.value]

    # ------------------------------------------------------------------ internals
    def _check(self, state: FieldState3D):
        N = self.res
        assert state.data.shape == (NUM_CHANNELS_3D, N, N, N) and state.data.dtype == self.dtype
        assert state.solid.shape == (N, N, N) and state.solid.dtype == self.dtype
        assert state.data.is_cuda and state.data.is_contiguous() and state.solid.is_contiguous()

    def _structure(self) -> tuple:
        c = self.config
        if c.pressure_mode == "pcg":
            return ()
        if c.pressure_mode == "learned":
            if getattr(self, "_net", None) is None and hasattr(self, "_ctx"):
                raise RuntimeError("pressure_mode='learned' needs Solver3D.set_preconditioner(net)")
            return ()
        if c.pressure_mode != "fused":
            raise ValueError(f"pressure_mode must be 'pcg', 'learned' or 'fused', got {c.pressure_mode!r}")
        if getattr(self, "_net", None) is not None:
            raise ValueError("pressure_mode='fused' runs without the learned preconditioner")
        if c.pressure_smoother not in ("chebyshev", "jacobi"):
            raise ValueError(f"pressure_smoother must be 'chebyshev' or 'jacobi', got {c.pressure_smoother!r}")
        return ("fused", int(c.pressure_cycles), int(c.pressure_smooth), str(c.pressure_smoother),
                c.pressure_tol > 0, int(c.pressure_max_iters))

    def _stream(self) -> ctypes.c_void_p:
        return ctypes.c_void_p(torch.cuda.current_stream(torch.device(self.device)).cuda_stream)

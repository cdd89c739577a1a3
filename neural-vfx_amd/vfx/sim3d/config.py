"""Solver configuration: the reference's SolverConfig3D fields and factories
(vfx/sim3d/config.py:13-99), plus opt-in extensions whose defaults keep the reference step.

World units (metres, seconds, kelvin), so grids of different size driven by the same forces
agree at large scales.
"""
from __future__ import annotations

from dataclasses import dataclass

# Version of the default path, as the reference names it (config.py:13).  Opt-in options add
# "+suffix" tags (solver_version_tag).
SOLVER3D_VERSION: str = "sim3d-warp-0.1.0"
SOLVER3D_REVISION: str = "0.2.0"
# Implementation tag of this build (recorded nowhere in the state; for reports).
IMPLEMENTATION: str = "mi355x-hip"


@dataclass
class SolverConfig3D:
    # buoyancy: f_y = -buoyancy_alpha * density + buoyancy_beta * (T - T_ambient)
    buoyancy_alpha: float = 0.6
    buoyancy_beta: float = 0.004
    # vorticity confinement f = vorticity_eps * dx * (N x omega_s)
    vorticity_eps: float = 6.0
    # decay rates (1/s, exponential)
    density_dissipation: float = 0.05
    temperature_cooling: float = 4.0
    # combustion: first order above ignition_temp over an ignition_ramp-wide ramp; heat release
    # saturates at max_temperature
    burn_rate: float = 8.0
    ignition_temp: float = 550.0
    burn_heat: float = 4000.0
    smoke_from_fuel: float = 0.35
    # pressure: MG-preconditioned CG; stops when max|div u| dx / max|u| < pressure_tol or after
    # pressure_max_iters (pressure_tol = 0: fixed iteration count)
    pressure_tol: float = 2e-5
    pressure_max_iters: int = 40
    pressure_warm_start: bool = True
    max_temperature: float = 3000.0
    wind_marker_temp: float = 300.0
    heat_diffusion: float = 5.0e-5
    confinement_floor: float = 0.5
    confinement_ref_res: int = 0
    # "pcg" (MGPCG), "fused" (the reference's opt-in stationary multigrid) or "learned" (flexible
    # CG preconditioned by the network attached with Solver3D.set_preconditioner, MGPCG fallback)
    pressure_mode: str = "pcg"
    pressure_cycles: int = 2
    pressure_smooth: int = 4
    pressure_smoother: str = "chebyshev"
    # ---- extensions (defaults = the reference behaviour)
    # width of the combustion ignition ramp in K (the reference hard-codes 100 K)
    ignition_ramp: float = 100.0
    # > 0: CG stops when ||r||_2 / ||b||_2 <= pressure_rel_tol (instead of the normalised max test)
    pressure_rel_tol: float = 0.0
    # restarts of CG from the true residual b - K phi while it exceeds pressure_rel_tol
    pressure_restarts: int = 0
    # seed of the per-cell emitter noise (Emitter3D.noise / emitter_noise)
    noise_seed: int = 0
    # global emitter noise amplitude: injection scaled by U[1 - a, 1 + a] per cell and step
    emitter_noise: float = 0.0
    # storage of the multigrid levels: "fp32" (the reference), "fp16" or "bf16" (FP32 arithmetic;
    # CG vectors stay FP32, dots and alpha/beta FP64)
    mg_storage: str = "fp32"
    # how the host sequences a projection: "device" (no host synchronisation inside step(), as the
    # reference's on-device loop; iteration budgets adapt to recent projections) or "poll" (the host
    # reads the stop flag every iteration).  Results are the same while the budget suffices.
    pressure_sync: str = "device"


def fused_config(**kw) -> SolverConfig3D:
    """The reference's adaptive stationary-MG preset (config.py:62-68)."""
    base = dict(pressure_mode="fused", pressure_tol=5e-5, pressure_max_iters=24)
    base.update(kw)
    return SolverConfig3D(**base)


def fused_fixed_config(**kw) -> SolverConfig3D:
    """The reference's fixed-cost stationary-MG preset (config.py:71-77)."""
    base = dict(pressure_mode="fused", pressure_tol=0.0)
    base.update(kw)
    return SolverConfig3D(**base)


def solver_version_tag(cfg: SolverConfig3D, scalar_res: int | None = None) -> str:
    """SOLVER3D_VERSION plus one suffix per opt-in option in use (config.py:80-90)."""
    tag = SOLVER3D_VERSION
    if cfg.confinement_ref_res:
        tag += f"+confref{cfg.confinement_ref_res}"
    if cfg.pressure_mode == "fused":
        tag += f"+fused-c{cfg.pressure_cycles}s{cfg.pressure_smooth}{cfg.pressure_smoother[:4]}"
        tag += "-adaptive" if cfg.pressure_tol > 0 else "-fixed"
    elif cfg.pressure_mode == "learned":
        tag += "+learned-fcg"
    if scalar_res is not None:
        tag += f"+scalars{scalar_res}"
    if cfg.ignition_ramp != 100.0:
        tag += f"+ramp{cfg.ignition_ramp:g}"
    if cfg.pressure_rel_tol > 0:
        tag += f"+relres{cfg.pressure_rel_tol:g}"
    if cfg.emitter_noise > 0:
        tag += f"+noise{cfg.emitter_noise:g}s{cfg.noise_seed}"
    if cfg.mg_storage != "fp32":
        tag += f"+mg{cfg.mg_storage}"
    return tag


def runtime_config(**kw) -> SolverConfig3D:
    """Fixed 4 PCG iterations (config.py:93-99)."""
    base = dict(pressure_tol=0.0, pressure_max_iters=4)
    base.update(kw)
    return SolverConfig3D(**base)

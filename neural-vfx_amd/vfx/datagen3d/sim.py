"""Paired 3D simulation: the free 64^3 runtime sim + a 256^3 sim guided to follow it
(the reference's vfx/datagen3d/sim.py:1-85).

LR  = Solver3D(64, runtime_config(**scenario.sim))   the sandbox's solver setting
HR  = Solver3D(256, SolverConfig3D(**scenario.sim))  full-quality pressure
Both get identical Forces3D every step.  Solvers and state buffers are cached per resolution.

A scenario is any object with `sim` (dict of SolverConfig3D overrides), `obstacles` (list of
Obstacle3D), `num_steps` (int) and `forces(step) -> Forces3D` (the reference's Scenario3D
interface); `PairScenario` is a minimal one.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Iterator

import torch

from vfx.core.fields3d import HR_RES_3D, LR_RES_3D, SIM_DT, Forces3D

from .guide3d import Guide3D, GuideConfig3D, tracking_error

_solvers: dict = {}
_states: dict = {}


@dataclass
class PairScenario:
    forces_fn: Callable[[int], Forces3D]
    num_steps: int = 60
    sim: dict = field(default_factory=dict)
    obstacles: list = field(default_factory=list)

    def forces(self, step: int) -> Forces3D:
        return self.forces_fn(step)


def solver(res: int, runtime: bool):
    from vfx.sim3d import Solver3D, SolverConfig3D
    key = (res, runtime)
    if key not in _solvers:
        _solvers[key] = Solver3D(res, SolverConfig3D())
    return _solvers[key]


def configure(s, overrides: dict, runtime: bool) -> None:
    from vfx.sim3d import SolverConfig3D, runtime_config
    s.config = runtime_config(**overrides) if runtime else SolverConfig3D(**overrides)


def fresh_state(s, key, obstacles):
    new = s.reset(obstacles)
    st = _states.get(key)
    if st is None:
        _states[key] = new
        return new
    st.data.copy_(new.data)
    st.solid.copy_(new.solid)
    st.meta = new.meta
    return st


def solver_version() -> str:
    from vfx.sim3d import SOLVER3D_VERSION
    return SOLVER3D_VERSION


def run_pair(sc, guide: GuideConfig3D | None, hr_res: int = HR_RES_3D, lr_res: int = LR_RES_3D,
             steps: int | None = None, with_hr: bool = True) -> Iterator[tuple]:
    """Yield (step, lr_state, hr_state, track) after every solver step (sim.py:52-79): the LR
    runtime step, and (with_hr) the guided HR step with the velocity nudge before it and the
    scalar nudge after it.  track = (3,) device tensor, rel L2 at 16^3 (density, T - ambient,
    velocity) of HR vs LR, or None without HR.  guide=None runs the HR unguided.  States are
    cached buffers: copy what you keep."""
    ls = solver(lr_res, True)
    configure(ls, sc.sim, True)
    lr = fresh_state(ls, ("lr", lr_res), sc.obstacles)
    hs = hr = g = None
    if with_hr:
        hs = solver(hr_res, False)
        configure(hs, sc.sim, False)
        hr = fresh_state(hs, ("hr", hr_res), sc.obstacles)
        g = Guide3D(guide, hr.solid, lr_res) if guide is not None else None
    n = sc.num_steps if steps is None else steps
    for s in range(n):
        f = sc.forces(s)
        if hr is not None:
            if g is not None:
                g.velocity(hr, lr)              # both hold the state after s steps
            hr = hs.step(hr, f, SIM_DT)
        lr = ls.step(lr, f, SIM_DT)
        tr = None
        if hr is not None:
            if g is not None:
                g.scalars(hr, lr)
            tr = tracking_error(hr.data, lr.data)
        yield s, lr, hr, tr


def free_cache() -> None:
    _solvers.clear()
    _states.clear()
    torch.cuda.empty_cache()

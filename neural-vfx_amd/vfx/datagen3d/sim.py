"""Paired 3D runs: a free low-resolution run with the sandbox's runtime solver and a
high-resolution run with the full-quality solver, guided to follow it at large scales.

Interface of the reference's vfx/datagen3d/sim.py:52-85 (`run_pair`, `free_cache`); the two
solvers get the same Forces3D every step.  The design here is a `PairedRun` object per pair of
grid sizes that owns both solvers and both state buffers, so a process that runs many scenarios
keeps one set of device buffers (and the solvers' captured step graphs, which are keyed by the
state pointers) per size pair.

A scenario is any object with `sim` (dict of SolverConfig3D overrides), `obstacles` (list of
Obstacle3D), `num_steps` (int) and `forces(step) -> Forces3D`, the reference's Scenario3D
interface; `PairScenario` is a minimal one.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Iterator, Optional

import torch

from vfx.core.fields3d import HR_RES_3D, LR_RES_3D, SIM_DT, FieldState3D, Forces3D

from .guide3d import Guide3D, GuideConfig3D, tracking_error


@dataclass
class PairScenario:
    forces_fn: Callable[[int], Forces3D]
    num_steps: int = 60
    sim: dict = field(default_factory=dict)
    obstacles: list = field(default_factory=list)

    def forces(self, step: int) -> Forces3D:
        return self.forces_fn(step)


class _Half:
    """One side of a paired run: a solver and the state it steps.  `runtime` selects the
    sandbox's solver setting (runtime_config) instead of the full-quality SolverConfig3D."""

    def __init__(self, res: int, runtime: bool):
        from vfx.sim3d import Solver3D, SolverConfig3D
        self.res = res
        self.runtime = runtime
        self.solver = Solver3D(res, SolverConfig3D())
        self.state: Optional[FieldState3D] = None

    def start(self, overrides: dict, obstacles) -> FieldState3D:
        """Apply the scenario's config overrides and reset the state to the scenario's start.  The
        first start allocates the state; later starts rewrite the same buffers in place."""
        from vfx.sim3d import SolverConfig3D, runtime_config
        make = runtime_config if self.runtime else SolverConfig3D
        self.solver.config = make(**overrides)
        start = self.solver.reset(obstacles)
        if self.state is None:
            self.state = start
        else:
            self.state.data.copy_(start.data)
            self.state.solid.copy_(start.solid)
            self.state.meta = start.meta
        return self.state

    def step(self, forces: Forces3D, dt: float) -> FieldState3D:
        self.state = self.solver.step(self.state, forces, dt)
        return self.state


class PairedRun:
    """The LR runtime run and the (optional) guided HR run of one pair of grid sizes."""

    _live: dict = {}

    def __init__(self, lr_res: int, hr_res: int):
        self.lr_res, self.hr_res = lr_res, hr_res
        self.lr = _Half(lr_res, runtime=True)
        self._hr: Optional[_Half] = None

    @classmethod
    def get(cls, lr_res: int = LR_RES_3D, hr_res: int = HR_RES_3D) -> "PairedRun":
        """The process-wide run of this size pair, created on first use."""
        key = (lr_res, hr_res)
        run = cls._live.get(key)
        if run is None:
            run = cls._live[key] = cls(lr_res, hr_res)
        return run

    @property
    def hr(self) -> _Half:
        if self._hr is None:
            self._hr = _Half(self.hr_res, runtime=False)
        return self._hr

    def frames(self, sc, guide: GuideConfig3D | None, steps: int | None, with_hr: bool) -> Iterator[tuple]:
        """Reset both runs to the scenario and yield (step, lr_state, hr_state, track) after every
        step.  Per step: the velocity nudge of HR toward LR (both hold the state after `step`
        steps), the HR step, the LR step, then the scalar nudge and the tracking error."""
        lr = self.lr.start(sc.sim, sc.obstacles)
        hr = self.hr.start(sc.sim, sc.obstacles) if with_hr else None
        g = Guide3D(guide, hr.solid, self.lr_res) if (with_hr and guide is not None) else None
        total = sc.num_steps if steps is None else steps
        for k in range(total):
            f = sc.forces(k)
            track = None
            if hr is not None:
                if g is not None:
                    g.velocity(hr, lr)
                hr = self.hr.step(f, SIM_DT)
            lr = self.lr.step(f, SIM_DT)
            if hr is not None:
                if g is not None:
                    g.scalars(hr, lr)
                track = tracking_error(hr.data, lr.data)
            yield k, lr, hr, track

    @classmethod
    def release_all(cls) -> None:
        cls._live.clear()


def run_pair(sc, guide: GuideConfig3D | None, hr_res: int = HR_RES_3D, lr_res: int = LR_RES_3D,
             steps: int | None = None, with_hr: bool = True) -> Iterator[tuple]:
    """Yield (step, lr_state, hr_state, track) after every solver step (reference sim.py:52-79).
    track is a (3,) device tensor: the rel L2 error at 16^3 of HR against LR for density,
    T - ambient and velocity; None without HR.  guide=None runs HR unguided.  The states are the
    run's own buffers, rewritten by the next step: copy what you keep."""
    return PairedRun.get(lr_res, hr_res).frames(sc, guide, steps, with_hr)


def solver_version() -> str:
    from vfx.sim3d import SOLVER3D_VERSION
    return SOLVER3D_VERSION


def free_cache() -> None:
    """Drop every paired run (solvers and states) and return the memory to the device."""
    PairedRun.release_all()
    torch.cuda.empty_cache()

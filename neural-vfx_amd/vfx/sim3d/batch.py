"""Batched frames: many independent sim3d simulations of any sizes stepped as one frame.

Each frame runs every simulation's Solver3D.step (vfx/sim3d/solver.py:146-155) on a few HIP
streams, directly or as one captured hipGraph replayed frame after frame (libvfxsim3d.so
vfx_batch_*).  Projections are MGPCG with a per-simulation iteration budget that grows when a
simulation uses it up, so a frame never waits on the host; wait() returns per-simulation
completion times measured on the GPU's wall clock.  tune() times the per-size launch shapes and
the stream count on snapshots of the simulations' own states and keeps the fastest.

The reference has no batched solver: it steps one Solver3D per call.  FP32 only.  A simulation
with a coarse-correction network (set_preconditioner) runs the learned flexible CG with its MGPCG
fallback, sequenced on the device like MGPCG; the full-resolution network, whose flexible-CG loop
waits on the host, is not available inside frames.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from vfx.core.fields3d import Forces3D

from . import _lib
from .config import SolverConfig3D
from .solver import P_SIZE, Solver3D, pack_params


class BatchSolver3D:
    def __init__(self, sizes, configs=None, streams: int = 4, graph: bool = True, device="cuda"):
        sizes = [int(n) for n in sizes]
        if not sizes:
            raise ValueError("at least one simulation")
        if configs is None:
            configs = [SolverConfig3D() for _ in sizes]
        elif isinstance(configs, SolverConfig3D):
            configs = [configs for _ in sizes]
        if len(configs) != len(sizes):
            raise ValueError("one config per simulation")
        self.solvers = [Solver3D(n, cfg, device=device) for n, cfg in zip(sizes, configs)]
        self.n = len(sizes)
        self.graph = bool(graph)
        self._lib = _lib.load()
        ctxs = (ctypes.c_void_p * self.n)(*[s._ctx.value for s in self.solvers])
        self._h = ctypes.c_void_p()
        _lib.check(self._lib.vfx_batch_create(ctxs, self.n, int(streams), ctypes.byref(self._h)), "BatchSolver3D")
        self._prm = np.zeros((self.n, P_SIZE), np.float32)
        self._pending = False

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and getattr(self, "_lib", None) is not None:
            self._lib.vfx_batch_destroy(h)
            self._h = None

    @property
    def sizes(self):
        return [s.res for s in self.solvers]

    def reset(self, obstacles=None) -> list:
        obstacles = obstacles if obstacles is not None else [()] * self.n
        return [s.reset(o) for s, o in zip(self.solvers, obstacles)]

    def set_preconditioner(self, params) -> None:
        """Attach the learned coarse correction to every simulation: params are the trained weights
        of a vfx.sim3d.learned.UNet (they do not depend on the grid); each simulation gets its own
        network instance on its quarter-resolution grid.  None detaches."""
        from .learned import UNet
        for s in self.solvers:
            s.set_preconditioner(None if params is None else UNet(s.res // 4, params=params))

    # ------------------------------------------------------------------ frames
    def _arrays(self, states, forces, dt):
        if len(states) != self.n or len(forces) != self.n:
            raise ValueError("one state and one Forces3D per simulation")
        for s, st in zip(self.solvers, states):
            s._check(st)
            s._structure()
            if s._net is not None and s._net.res * 4 != s.res:
                raise ValueError("inside batched frames the learned preconditioner is the coarse correction "
                                 "(a network on the quarter-resolution grid)")
            s._apply_options()
        for i, (s, st, f) in enumerate(zip(self.solvers, states, forces)):
            pack_params(s.config, s.res, f if f is not None else Forces3D(), dt, st.meta.step, self._prm[i])
        data = (ctypes.c_void_p * self.n)(*[st.data.data_ptr() for st in states])
        solid = (ctypes.c_void_p * self.n)(*[st.solid.data_ptr() for st in states])
        return data, solid

    def step(self, states, forces, dt: float, graph: bool | None = None) -> list:
        """Enqueue one frame on torch's current stream (returns without waiting)."""
        data, solid = self._arrays(states, forces, dt)
        g = self.graph if graph is None else bool(graph)
        _lib.check(self._lib.vfx_batch_step(self._h, data, solid, self._prm.ctypes.data_as(ctypes.c_void_p),
                                            1 if g else 0, self._stream()), "BatchSolver3D.step")
        self._pending = True
        for s, st in zip(self.solvers, states):
            st.meta.step += 1
            st.meta.dt = float(dt)
            st.meta.solver_version = s.version
        return states

    def wait(self) -> dict:
        """Wait for the last frame: per-simulation latency (ms from the frame's start to the end of
        the simulation's step), PCG iterations, convergence flags, recurrence ||r||/||b||, and the
        frame's GPU time."""
        lat = np.zeros(self.n, np.float32)
        its = np.zeros(self.n, np.int32)
        conv = np.zeros(self.n, np.int32)
        rr = np.zeros(self.n, np.float64)
        fm = ctypes.c_float()
        _lib.check(self._lib.vfx_batch_wait(self._h, lat.ctypes.data_as(ctypes.c_void_p),
                                            its.ctypes.data_as(ctypes.c_void_p), conv.ctypes.data_as(ctypes.c_void_p),
                                            rr.ctypes.data_as(ctypes.c_void_p), ctypes.byref(fm)),
                   "BatchSolver3D.wait")
        self._pending = False
        return {"latency_ms": lat, "iters": its, "converged": conv.astype(bool), "rel_res": rr,
                "frame_ms": float(fm.value)}

    @property
    def budgets(self) -> np.ndarray:
        b = np.zeros(self.n, np.int32)
        _lib.check(self._lib.vfx_batch_get_budget(self._h, b.ctypes.data_as(ctypes.c_void_p)))
        return b

    def truncated(self) -> int:
        """Simulation-frames, among the frames wait() has read, whose projection stopped at its
        iteration budget unconverged (no wait; wait() doubles that budget)."""
        n = ctypes.c_longlong()
        _lib.check(self._lib.vfx_batch_truncated(self._h, ctypes.byref(n)))
        return int(n.value)

    def set_budget(self, budget: int, sim: int = -1) -> None:
        _lib.check(self._lib.vfx_batch_set_budget(self._h, int(sim), int(budget)))

    def tune(self, states, forces, dt: float, reps: int = 3) -> dict:
        """Autotune the launch shapes per grid size and the stream count on the current states
        (which are left as they were)."""
        data, solid = self._arrays(states, forces, dt)
        rep = np.zeros(_lib.TUNE_REPORT, np.float64)
        _lib.check(self._lib.vfx_batch_tune(self._h, data, solid, self._prm.ctypes.data_as(ctypes.c_void_p), int(reps),
                                            self._stream(), rep.ctypes.data_as(ctypes.c_void_p), rep.size),
                   "BatchSolver3D.tune")
        sizes = {}
        for k in range(int(rep[0])):
            N, zc, cm, best, dflt = rep[1 + 5 * k:6 + 5 * k]
            sizes[int(N)] = {"cells_per_thread": int(zc), "coarse_m": int(cm), "ms": float(best),
                             "ms_before": float(dflt)}
        o = 1 + 5 * _lib.TUNE_MAX_SIZES
        return {"sizes": sizes, "streams": int(rep[o]),
                "frame_ms_by_streams": dict(zip((1, 2, 4, 8), [float(x) for x in rep[o + 1:o + 5]]))}

    def _stream(self) -> ctypes.c_void_p:
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

"""Learned pressure preconditioner: a 3-level 3D U-Net on the library's BF16 MFMA kernels.

The network maps a residual field to a pressure correction: z = rms(r) * net(r / rms(r), fluid),
zero in solids.  It is trained on (divergence, pressure) pairs that the solver itself produces,
in the energy norm of the solver's operator K (loss = e.K e / phi.K phi with e = net(b) - phi),
with Adam on FP32 master weights; the convolutions run in BF16 with FP32 accumulation
(libvfxsim3d.so, vfx_net_*).  Flexible CG uses it as its preconditioner, with plain MGPCG as the
fallback (Solver3D with pressure_mode="learned").
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib

# (cin, cout, relu) of the 11 convolutions in execution order: encoder 2->16->16 | pool |
# 16->32->32 | pool | 32->64->64 | up, concat (64 + 32) | 96->32->32 | up, concat (32 + 16) |
# 48->16->16 -> 1
LAYERS = ((2, 16, True), (16, 16, True), (16, 32, True), (32, 32, True), (32, 64, True), (64, 64, True),
          (96, 32, True), (32, 32, True), (48, 16, True), (16, 16, True), (16, 1, False))


def param_count() -> int:
    return sum(co * ci * 27 + co for ci, co, _ in LAYERS)


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights for the ReLU layers (std sqrt(2 / fan_in)), a small last layer, zero
    biases; flat FP32 in the library's order (per layer: weight [cout][cin][3][3][3], bias)."""
    g = torch.Generator().manual_seed(int(seed))
    out = []
    for i, (ci, co, relu) in enumerate(LAYERS):
        std = math.sqrt(2.0 / (ci * 27)) if relu else 0.1 * math.sqrt(1.0 / (ci * 27))
        out.append((torch.randn(co * ci * 27, generator=g, dtype=torch.float64) * std).numpy())
        out.append(np.zeros(co))
    return np.concatenate(out).astype(np.float32)


def split_params(flat: np.ndarray):
    """flat parameters -> [(weight (cout, cin, 3, 3, 3), bias (cout,)), ...]"""
    res, k = [], 0
    for ci, co, _ in LAYERS:
        w = flat[k:k + co * ci * 27].reshape(co, ci, 3, 3, 3)
        k += co * ci * 27
        b = flat[k:k + co]
        k += co
        res.append((w, b))
    return res


class UNet:
    """One network instance for grids of size res (res % 4 == 0)."""

    def __init__(self, res: int, params: np.ndarray | None = None, seed: int = 0):
        self.res = int(res)
        self._lib = _lib.load()
        self._h = ctypes.c_void_p()
        _lib.check(self._lib.vfx_net_create(self.res, ctypes.byref(self._h)), "UNet")
        n = ctypes.c_int()
        _lib.check(self._lib.vfx_net_param_count(self._h, ctypes.byref(n)))
        self.nparam = int(n.value)
        assert self.nparam == param_count()
        self.set_params(init_params(seed) if params is None else params)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and getattr(self, "_lib", None) is not None:
            self._lib.vfx_net_destroy(h)
            self._h = None

    # ---------------------------------------------------------------- parameters
    def params(self) -> np.ndarray:
        out = np.empty(self.nparam, np.float32)
        _lib.check(self._lib.vfx_net_get_params(self._h, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (self.nparam,)
        _lib.check(self._lib.vfx_net_set_params(self._h, p.ctypes.data_as(ctypes.c_void_p)))

    def save(self, path: str) -> None:
        np.savez(path, params=self.params(), layers=np.array(LAYERS, dtype=np.int64))

    @classmethod
    def load(cls, path: str, res: int) -> "UNet":
        with np.load(path, allow_pickle=False) as f:
            return cls(res, params=f["params"])

    # ---------------------------------------------------------------- raw network (tests)
    @staticmethod
    def _stream():
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def forward(self, inp: torch.Tensor) -> torch.Tensor:
        """inp (2, res, res, res) float32 on the GPU -> (res, res, res) float32."""
        N = self.res
        assert inp.shape == (2, N, N, N) and inp.dtype == torch.float32 and inp.is_cuda
        inp = inp.contiguous()
        out = torch.empty(N, N, N, device=inp.device)
        _lib.check(self._lib.vfx_net_forward(self._h, ctypes.c_void_p(inp.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                             self._stream()), "UNet.forward")
        return out

    def backward(self, grad_out: torch.Tensor) -> torch.Tensor:
        """After forward(): d loss / d out -> flat parameter gradient (float32, GPU)."""
        g = grad_out.contiguous().float()
        grad = torch.empty(self.nparam, device=g.device)
        _lib.check(self._lib.vfx_net_backward(self._h, ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(grad.data_ptr()),
                                              self._stream()), "UNet.backward")
        return grad

    def adam(self, grad: torch.Tensor | None, lr: float, beta1=0.9, beta2=0.999, eps=1e-8) -> None:
        gp = ctypes.c_void_p(grad.contiguous().data_ptr()) if grad is not None else None
        _lib.check(self._lib.vfx_net_adam(self._h, gp, float(lr), float(beta1), float(beta2), float(eps),
                                          self._stream()), "UNet.adam")

    # ---------------------------------------------------------------- with a solver
    def train_step(self, solver, solid: torch.Tensor, b: torch.Tensor, phi: torch.Tensor, lr: float) -> float:
        """One Adam step on the pair (b, phi) of `solver`'s grid; returns the energy-norm loss."""
        loss = ctypes.c_double()
        _lib.check(self._lib.vfx_net_train_step(self._h, solver._ctx, ctypes.c_void_p(solid.data_ptr()),
                                                ctypes.c_void_p(b.data_ptr()), ctypes.c_void_p(phi.data_ptr()),
                                                float(lr), ctypes.byref(loss), self._stream()), "UNet.train_step")
        return float(loss.value)

    def train_step_coarse(self, solver, solid: torch.Tensor, rv: torch.Tensor, ev: torch.Tensor, lr: float) -> float:
        """One Adam step of the coarse correction (this network's grid is a quarter of `solver`'s) on
        the residual rv and error ev a V-cycle left; returns the fine energy-norm loss."""
        loss = ctypes.c_double()
        _lib.check(self._lib.vfx_net_train_step_coarse(self._h, solver._ctx, ctypes.c_void_p(solid.data_ptr()),
                                                       ctypes.c_void_p(rv.data_ptr()), ctypes.c_void_p(ev.data_ptr()),
                                                       float(lr), ctypes.byref(loss), self._stream()),
                   "UNet.train_step_coarse")
        return float(loss.value)

    def apply(self, solver, solid: torch.Tensor, r: torch.Tensor) -> torch.Tensor:
        """The preconditioner z = M(r) on `solver`'s grid."""
        z = torch.empty_like(r)
        _lib.check(self._lib.vfx_net_apply(self._h, solver._ctx, ctypes.c_void_p(solid.data_ptr()),
                                           ctypes.c_void_p(r.contiguous().data_ptr()), ctypes.c_void_p(z.data_ptr()),
                                           self._stream()), "UNet.apply")
        return z


# ---------------------------------------------------------------------------- data and training
def random_scene(seed: int):
    """A seeded training / evaluation scene: 1-3 spherical emitters at uniform positions in the
    lower half, 0-2 sphere obstacles of radius U[0.05, 0.2] m, buoyancy beta U[0.5, 2] per K.
    Returns (Forces3D, obstacles, beta)."""
    from vfx.core.fields3d import Emitter3D, Forces3D, Obstacle3D
    g = np.random.default_rng(int(seed))
    ems = []
    for _ in range(int(g.integers(1, 4))):
        pos = (float(g.uniform(0.15, 0.85)), float(g.uniform(0.05, 0.45)), float(g.uniform(0.15, 0.85)))
        ems.append(Emitter3D(pos=pos, radius=float(g.uniform(0.05, 0.1)), density_rate=60.0,
                             temperature=float(293.0 + g.uniform(1.0, 5.0)), fuel_rate=0.0,
                             velocity=(0.0, float(g.uniform(0.0, 1.0)), 0.0)))
    obst = []
    for _ in range(int(g.integers(0, 3))):
        r = float(g.uniform(0.05, 0.2))
        obst.append(Obstacle3D(kind="sphere", pos=(float(g.uniform(0.25, 0.75)), float(g.uniform(0.45, 0.8)),
                                                   float(g.uniform(0.25, 0.75))), size=(r, r, r)))
    return Forces3D(emitters=ems), obst, float(g.uniform(0.5, 2.0))


def scene_config(beta: float, seed: int, rel_tol: float = 1e-5):
    from .config import SolverConfig3D
    return SolverConfig3D(buoyancy_alpha=0.0, buoyancy_beta=beta, vorticity_eps=0.3, pressure_rel_tol=rel_tol,
                          pressure_max_iters=200, pressure_restarts=2, emitter_noise=0.2, noise_seed=int(seed))


def generate_pairs(res: int, seeds, steps: int, every: int = 1, warmup: int = 3):
    """(solid, b, phi) training pairs from the plain solver: phi is the converged pressure of a
    step (relative residual 1e-6) of a random scene, b = K phi (the right-hand side phi solves
    exactly).  Returns a list of GPU tensor triples."""
    from .solver import Solver3D
    pairs = []
    for sd in seeds:
        forces, obst, beta = random_scene(sd)
        s = Solver3D(res, scene_config(beta, sd, rel_tol=1e-6))
        st = s.reset(obst)
        for i in range(steps):
            s.step(st, forces, 1.0 / 60.0)
            if i >= warmup and (i - warmup) % every == 0:
                phi = torch.empty(res, res, res, device=st.data.device)
                s._chk(s._api.vfx_get_phi(s._ctx, ctypes.c_void_p(phi.data_ptr()), s._stream()))
                b = torch.empty_like(phi)
                s._chk(s._api.vfx_kmul(s._ctx, ctypes.c_void_p(st.solid.data_ptr()), ctypes.c_void_p(phi.data_ptr()),
                                       ctypes.c_void_p(b.data_ptr()), s._stream()))
                pairs.append((st.solid.clone(), b, phi))
        torch.cuda.synchronize()
    return pairs


def noise_pairs(res: int, solids, count: int, seed: int = 0):
    """(solid, K v, v) pairs with v multi-scale random noise in the fluid cells (octaves 4^3 to
    res^3 with equal weight): right-hand sides whose content reaches the high frequencies that
    the residuals of later CG iterations carry.  Exactly consistent, like generate_pairs."""
    import torch.nn.functional as F
    from .solver import Solver3D
    s = Solver3D(res)
    g = torch.Generator(device="cpu").manual_seed(int(seed))
    out = []
    for i in range(int(count)):
        solid = solids[i % len(solids)]
        v = torch.zeros(1, 1, res, res, res)
        size = 4
        while size <= res:
            z = torch.randn(1, 1, size, size, size, generator=g)
            v += F.interpolate(z, size=(res,) * 3, mode="trilinear", align_corners=False)
            size *= 2
        v = (v[0, 0].to(solid.device) * (solid < 0.5)).contiguous()
        b = torch.empty_like(v)
        s._chk(s._api.vfx_kmul(s._ctx, ctypes.c_void_p(solid.data_ptr()), ctypes.c_void_p(v.data_ptr()),
                               ctypes.c_void_p(b.data_ptr()), s._stream()))
        out.append((solid, b, v))
    torch.cuda.synchronize()
    return out


def vcycle_pairs(res: int, pairs):
    """(solid, r_V, e_V) from (solid, b, phi) pairs of grid res: z_V = M_MG(b) (the MG
    preconditioner), r_V = b - K z_V and e_V = phi - z_V, so that K e_V = r_V exactly as K phi = b:
    what the V-cycle leaves for the coarse correction to remove."""
    from .solver import Solver3D
    s = Solver3D(res)
    out = []
    for solid, b, phi in pairs:
        z = torch.empty_like(b)
        s._chk(s._api.vfx_precond(s._ctx, ctypes.c_void_p(solid.data_ptr()), ctypes.c_void_p(b.data_ptr()),
                                  ctypes.c_void_p(z.data_ptr()), s._stream()))
        kz = torch.empty_like(b)
        s._chk(s._api.vfx_kmul(s._ctx, ctypes.c_void_p(solid.data_ptr()), ctypes.c_void_p(z.data_ptr()),
                               ctypes.c_void_p(kz.data_ptr()), s._stream()))
        out.append((solid, (b - kz).contiguous(), (phi - z).contiguous()))
    torch.cuda.synchronize()
    return out


def train_coarse(net: UNet, vpairs, iters: int, lr: float = 1e-3, seed: int = 0, lr_decay: float = 0.3, log=None):
    """Adam on the coarse correction's fine energy-norm loss over (solid, r_V, e_V) pairs of grid
    4 net.res (shuffled each epoch; the same learning-rate schedule as train).  Returns the loss per
    iteration."""
    from .solver import Solver3D
    solver = Solver3D(4 * net.res)
    g = np.random.default_rng(int(seed))
    order, losses = [], []
    for it in range(int(iters)):
        if not order:
            order = list(g.permutation(len(vpairs)))
        solid, rv, ev = vpairs[order.pop()]
        rate = lr * (lr_decay if it >= 0.6 * iters else 1.0) * (lr_decay if it >= 0.85 * iters else 1.0)
        losses.append(net.train_step_coarse(solver, solid, rv, ev, rate))
        if log is not None and (it % 100 == 0 or it == iters - 1):
            log(f"train {it}: loss {float(np.mean(losses[-50:])):.4f}")
    return losses


def train(net: UNet, pairs, iters: int, lr: float = 1e-3, seed: int = 0, lr_decay: float = 0.3, log=None):
    """Adam on the energy-norm loss over `pairs` (shuffled each epoch); the learning rate decays
    by lr_decay at 60 % and 85 % of the iterations.  Returns the loss per iteration."""
    from .solver import Solver3D
    solver = Solver3D(net.res)
    g = np.random.default_rng(int(seed))
    order = []
    losses = []
    for it in range(int(iters)):
        if not order:
            order = list(g.permutation(len(pairs)))
        solid, b, phi = pairs[order.pop()]
        rate = lr * (lr_decay if it >= 0.6 * iters else 1.0) * (lr_decay if it >= 0.85 * iters else 1.0)
        losses.append(net.train_step(solver, solid, b, phi, rate))
        if log is not None and (it % 100 == 0 or it == iters - 1):
            log(f"train {it}: loss {float(np.mean(losses[-50:])):.4f}")
    return losses

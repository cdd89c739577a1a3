"""Guided HR 3D simulation: a 256^3 run follows an independent 64^3 run at large scales
(the reference's vfx/datagen3d/guide3d.py), on the HIP guidance kernels of libvfxsim3d.so.

Every step (guide3d.py:1-22):

    L(x)  = spectral low-pass exp(-(k/cutoff)^8), k in cycles/domain, of the mirror-extended LR
            grid, then trilinear upsampling by the integer factor HR / LR
    vel  += k_vel * L(vel_lr - avgpool(vel_hr))                   BEFORE the HR step
    s    *= exp(k_scalar * L(clamp(log((lp(s_lr)+eps) / (lp(avgpool(s_hr))+eps)), -1, 1)))
            AFTER the HR step, for s in {density, temperature - ambient}
    fuel += k_fuel * L(fuel_lr - avgpool(fuel_hr))                (default 0: off)

Solid cells are never touched; density stays >= 0, temperature >= min(T_old, ambient).  The
updates edit `state.data` in place on torch's current stream.  The low-pass runs as a DCT
(the mirror extension's FFT restricted to the grid), see neural-vfx_amd/csrc/guide.hip.
"""
from __future__ import annotations

import ctypes
import dataclasses
from dataclasses import dataclass

import torch

from vfx.core.fields3d import NUM_CHANNELS_3D
from vfx.sim3d import _lib


@dataclass(frozen=True)
class GuideConfig3D:
    k_vel: float = 1.0          # per-step additive relaxation of the velocity large scales
    k_scalar: float = 0.7       # per-step multiplicative (log-ratio) relaxation of density, T - ambient
    k_fuel: float = 0.0         # per-step additive relaxation of fuel
    cutoff: float = 16.0        # low-pass cutoff, cycles per domain (64^3 Nyquist = 32)
    eps_density: float = 0.01   # log-ratio regularizers
    eps_temp: float = 5.0       # K
    log_clamp: float = 1.0      # |log ratio| cap per step

    def asdict(self) -> dict:
        d = dataclasses.asdict(self)
        d["scheme"] = ("before the HR step vel += k_vel*L(vel_lr - pool(vel_hr)); after it density, T-ambient "
                       "*= exp(k_scalar*L(clamp(log-ratio))); fuel += k_fuel*L(dfuel); L = exp(-(k/cutoff)^8) "
                       "mirror-extended spectral low-pass on the LR grid (as a DCT) + trilinear upsampling")
        return d


DEFAULT_GUIDE = GuideConfig3D()


def _ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _stream(dev) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _chk(rc: int, what: str) -> None:
    _lib.check(rc, what)


class Guide3D:
    """Applies the large-scale relaxation of an HR FieldState3D toward an LR one, in place
    (guide3d.py:91-130).  HR / LR must be an integer factor; LR a multiple of 16 up to 128."""

    def __init__(self, cfg: GuideConfig3D, hr_solid: torch.Tensor, lr_res: int):
        self.cfg = cfg
        self.lib = _lib.load()
        hr = hr_solid.shape[-1]
        assert hr_solid.shape == (hr, hr, hr) and hr_solid.is_cuda and hr_solid.dtype == torch.float32
        self.hr_res, self.lr_res = int(hr), int(lr_res)
        self.factor = self.hr_res // self.lr_res
        self.device = hr_solid.device
        vals = (ctypes.c_float * 7)(*[float(getattr(cfg, k)) for k in _lib.GUIDE_CFG])
        self._h = ctypes.c_void_p()
        solid = hr_solid.contiguous()
        _chk(self.lib.vfx_guide_create(self.hr_res, self.lr_res, vals, _ptr(solid), _stream(self.device),
                                       ctypes.byref(self._h)), "Guide3D")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self.lib.vfx_guide_destroy(h)
            self._h = None

    def _check(self, hr, lr):
        assert hr.data.shape == (NUM_CHANNELS_3D,) + (self.hr_res,) * 3 and hr.data.is_contiguous()
        assert lr.data.shape == (NUM_CHANNELS_3D,) + (self.lr_res,) * 3 and lr.data.is_contiguous()

    @torch.no_grad()
    def velocity(self, hr, lr) -> None:
        """Before the HR step: hr velocity += k_vel up(L(lr velocity - pool(hr velocity)))."""
        self._check(hr, lr)
        _chk(self.lib.vfx_guide_velocity(self._h, _ptr(hr.data), _ptr(lr.data), _stream(self.device)),
             "Guide3D.velocity")

    @torch.no_grad()
    def scalars(self, hr, lr) -> None:
        """After the HR step: multiplicative density / temperature relaxation (and fuel if k_fuel)."""
        self._check(hr, lr)
        _chk(self.lib.vfx_guide_scalars(self._h, _ptr(hr.data), _ptr(lr.data), _stream(self.device)),
             "Guide3D.scalars")

    @torch.no_grad()
    def lowpass(self, x: torch.Tensor) -> torch.Tensor:
        """L on (C, n, n, n) at the guide's LR size, C <= 3 (guide3d.py:60-68)."""
        assert x.shape[-3:] == (self.lr_res,) * 3 and x.dim() == 4 and x.shape[0] <= 3
        x = x.contiguous().float()
        out = torch.empty_like(x)
        _chk(self.lib.vfx_guide_lowpass(self._h, _ptr(x), int(x.shape[0]), _ptr(out), _stream(self.device)),
             "lowpass")
        return out


def lowpass(x: torch.Tensor, cutoff: float) -> torch.Tensor:
    """The spectral low-pass exp(-(k/cutoff)^8) of (C, n, n, n), C <= 3, through a temporary guide
    (allocates; the per-step path reuses the Guide3D's buffers)."""
    n = x.shape[-1]
    g = Guide3D(GuideConfig3D(cutoff=cutoff), torch.zeros(n, n, n, device=x.device), n)
    return g.lowpass(x)


_scratch: dict = {}


def tracking_error(hr_data: torch.Tensor, lr_data: torch.Tensor, res: int = 16) -> torch.Tensor:
    """Relative L2 at res^3 (adaptive average pool) of HR vs LR: (density, T - ambient, velocity),
    a (3,) tensor on the device, no host synchronisation (guide3d.py:133-145)."""
    assert hr_data.is_cuda and lr_data.is_cuda and hr_data.shape[0] == lr_data.shape[0] == NUM_CHANNELS_3D
    key = (hr_data.device, res)
    buf = _scratch.get(key)
    if buf is None:
        buf = _scratch[key] = torch.empty(2 * 6 * res ** 3, device=hr_data.device)
    out = torch.empty(3, device=hr_data.device)
    lib = _lib.load()
    _chk(lib.vfx_tracking_error(_ptr(hr_data.contiguous()), hr_data.shape[-1], _ptr(lr_data.contiguous()),
                                lr_data.shape[-1], int(res), _ptr(buf), _ptr(out), _stream(hr_data.device)),
         "tracking_error")
    return out

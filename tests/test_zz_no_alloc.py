"""GPU test of vfx_advect_noalloc, the advection entry point that allocates nothing: through the C
ABI, two calls in a row (the second on a smaller grid, so it runs on what the first left in the
shared scratch) equal vfx_advect on a context of the same size bit for bit.  (The rejection of a
grid above the scratch's maximum is a CPU test: it happens before any HIP call.)"""
from __future__ import annotations

import ctypes
import sys

import pytest
import torch

from vfx.sim3d import _lib

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _ctx_advect(lib, N, field, vel, dt, stream):
    ctx = ctypes.c_void_p()
    _lib.check(lib.vfx_ctx_create(N, ctypes.byref(ctx)), "ctx")
    try:
        out = torch.empty_like(field)
        _lib.check(lib.vfx_advect(ctx, _p(field), int(field.shape[0]), _p(vel), _p(out), ctypes.c_float(dt), stream),
                   "vfx_advect")
        torch.cuda.synchronize()
        return out
    finally:
        lib.vfx_ctx_destroy(ctx)


def test_advect_noalloc_matches_advect(cuda, hiplib):
    lib = hiplib
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    g = torch.Generator(device="cuda").manual_seed(5)
    for N, C, dt in ((96, 6, 1.0 / 30.0), (48, 5, 1.0 / 60.0)):
        _say(f"advect_noalloc {N}^3, {C} channels")
        field = torch.rand(C, N, N, N, device="cuda", generator=g)
        vel = (torch.rand(3, N, N, N, device="cuda", generator=g) - 0.5) * 2.0
        out = torch.full_like(field, float("nan"))
        rc = lib.vfx_advect_noalloc(_p(field), C, _p(vel), _p(out), N, ctypes.c_float(dt), stream)
        assert rc == 0, _lib.last_error()
        torch.cuda.synchronize()
        ref = _ctx_advect(lib, N, field, vel, dt, stream)
        _say(f"  max |diff| {(out - ref).abs().max().item():.3e}")
        assert torch.equal(out, ref)

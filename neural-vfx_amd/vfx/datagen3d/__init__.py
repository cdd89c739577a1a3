"""Paired 3D simulations: a free low-resolution run and a high-resolution run guided to follow it
at large scales (the reference's vfx/datagen3d sim.py + guide3d.py, on the MI355X solver and
guidance kernels)."""
from .guide3d import DEFAULT_GUIDE, Guide3D, GuideConfig3D, lowpass, tracking_error

__all__ = ["DEFAULT_GUIDE", "Guide3D", "GuideConfig3D", "lowpass", "tracking_error"]

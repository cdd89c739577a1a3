"""3D smoke + fire volume renderer on the MI355X (the reference's vfx/render3d)."""
from .renderer import RenderParams3D, VolumeRenderer, camera_basis

__all__ = ["RenderParams3D", "VolumeRenderer", "camera_basis"]

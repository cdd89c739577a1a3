"""3D state schema at the solver boundary (the data contract of vfx/core/fields3d.py:1-116).

Layout (C, D, H, W), index [c, z, y, x]; x right, y up (buoyancy +y), z depth.  The domain is a
cube of DOMAIN_SIZE metres, dx = DOMAIN_SIZE / res, world units everywhere.  Velocity is
cell-centred in m/s.  The box faces are walls; `solid` marks obstacles only.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Tuple

import torch

from .fields import AMBIENT_TEMP, DOMAIN_SIZE, SIM_DT

LR_RES_3D: int = 64                 # the runtime (low-resolution) grid of the paired runs
UPSCALE_3D: int = 4
HR_RES_3D: int = LR_RES_3D * UPSCALE_3D   # 256

CHANNELS_3D: Tuple[str, ...] = ("density", "vel_x", "vel_y", "vel_z", "temperature", "fuel")
CH3 = {name: i for i, name in enumerate(CHANNELS_3D)}
NUM_CHANNELS_3D = len(CHANNELS_3D)

Vec3 = Tuple[float, float, float]

__all__ = ["AMBIENT_TEMP", "DOMAIN_SIZE", "SIM_DT", "LR_RES_3D", "UPSCALE_3D", "HR_RES_3D", "CHANNELS_3D", "CH3",
           "NUM_CHANNELS_3D", "FieldMeta3D", "FieldState3D", "Volume", "Camera", "Emitter3D", "Impulse3D", "Forces3D",
           "Obstacle3D", "rasterize_obstacles_3d"]


@dataclass
class FieldMeta3D:
    res: Tuple[int, int, int]
    dx: float
    dt: float
    channels: Tuple[str, ...] = CHANNELS_3D
    ambient_temp: float = AMBIENT_TEMP
    step: int = 0
    solver_version: str = "unknown"

    @staticmethod
    def for_res(res: int, dt: float = SIM_DT) -> "FieldMeta3D":
        return FieldMeta3D(res=(res, res, res), dx=DOMAIN_SIZE / res, dt=dt)


@dataclass
class FieldState3D:
    """data: (6, D, H, W) float32 on the GPU; solid: (D, H, W) float32, 1 inside obstacles."""
    data: torch.Tensor
    solid: torch.Tensor
    meta: FieldMeta3D

    def ch(self, name: str) -> torch.Tensor:
        return self.data[CH3[name]]

    @property
    def velocity(self) -> torch.Tensor:
        return self.data[CH3["vel_x"]:CH3["vel_z"] + 1]


@dataclass
class Volume:
    density: torch.Tensor
    temperature: torch.Tensor


@dataclass
class Emitter3D:
    """Sphere source with falloff (1 - r^2/R^2)^2.  `temperature` is a target (K) inside the
    sphere, 0 = no heat; rates are per second."""
    pos: Vec3
    radius: float
    density_rate: float = 1.0
    temperature: float = 0.0
    fuel_rate: float = 0.0
    velocity: Vec3 = (0.0, 0.0, 0.0)


@dataclass
class Impulse3D:
    pos: Vec3
    radius: float
    velocity: Vec3


@dataclass
class Forces3D:
    emitters: List[Emitter3D] = field(default_factory=list)
    impulses: List[Impulse3D] = field(default_factory=list)
    wind: Vec3 = (0.0, 0.0, 0.0)


@dataclass
class Obstacle3D:
    """kind 'sphere' (size[0] = radius) or 'box' (size = half extents)."""
    kind: str
    pos: Vec3
    size: Vec3


def rasterize_obstacles_3d(obstacles: List[Obstacle3D], res: int, device: str = "cuda") -> torch.Tensor:
    """(res,)^3 float32, 1 where a cell centre lies inside an obstacle (inclusive test)."""
    dx = DOMAIN_SIZE / res
    c = (torch.arange(res, device=device, dtype=torch.float32) + 0.5) * dx
    zz, yy, xx = torch.meshgrid(c, c, c, indexing="ij")
    solid = torch.zeros(res, res, res, device=device, dtype=torch.float32)
    for o in obstacles:
        if o.kind == "sphere":
            inside = (xx - o.pos[0]) ** 2 + (yy - o.pos[1]) ** 2 + (zz - o.pos[2]) ** 2 <= o.size[0] ** 2
        elif o.kind == "box":
            inside = (((xx - o.pos[0]).abs() <= o.size[0]) & ((yy - o.pos[1]).abs() <= o.size[1])
                      & ((zz - o.pos[2]).abs() <= o.size[2]))
        else:
            raise ValueError(o.kind)
        solid[inside] = 1.0
    return solid


@dataclass
class Camera:
    """Perspective camera in world space; the domain cube spans [0, DOMAIN_SIZE]^3 (the reference's
    vfx/core/fields3d.py:119-125)."""
    position: Vec3 = (0.5, 0.55, -1.4)
    target: Vec3 = (0.5, 0.45, 0.5)
    up: Vec3 = (0.0, 1.0, 0.0)
    fov_y_deg: float = 40.0

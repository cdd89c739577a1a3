"""Shared constants of the solver boundary (the values of vfx/core/fields.py:25-30)."""
from __future__ import annotations

DOMAIN_SIZE: float = 1.0          # metres, cube side
AMBIENT_TEMP: float = 293.0       # kelvin
SIM_DT: float = 1.0 / 30.0        # seconds per solver step (30 Hz)

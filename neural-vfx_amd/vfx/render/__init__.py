"""Rendering helpers shared by the volume renderer (the blackbody emission table)."""

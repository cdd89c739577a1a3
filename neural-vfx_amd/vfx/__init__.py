"""MI355X-native rebuild of the vfx.sim3d smoke/fire solver step (JohnRiceML/neural-vfx API)."""

"""Host side of the MI355X DRMLT mutation loop: procedural scenes, the ctypes
binding of libdrmlt_hip.so and an integrator that mirrors the reference's
DRIntegrator (PrePass, DirectLighting, Render)."""
from . import scenes  # noqa: F401

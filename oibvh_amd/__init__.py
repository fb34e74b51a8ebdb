"""oibvh_amd: MI355X-native (gfx950 HIP) oibvh build / refit / broad / narrow phase.

The product is the C-ABI library oibvh_amd/liboibvh_hip.so (include/oibvh.h);
this package holds its sources (csrc/), the build driver, a ctypes mirror of
the reference's OibvhTree / Scene interface (api.py) and the seeded workload
generators (meshgen.py).
"""
from . import meshgen  # noqa: F401

__all__ = ["meshgen", "api", "build"]

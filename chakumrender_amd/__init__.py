"""chakumrender_amd: MI355X-native core of ChakumRender's per-sample radiance path.

The product is libchakum.so (HIP for gfx950, C ABI in include/chakum_c.h).
This package holds its ctypes binding (`abi`), the host-side mirror of the
reference's Accel / material / light / shader interfaces (`host`), the
seeded scene generators (`scenes`) and the PPM writer.
"""
from . import abi, scenes  # noqa: F401

__all__ = ["abi", "scenes"]

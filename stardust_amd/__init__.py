"""stardust_amd — MI355X (gfx950) implementation of stardust's per-timestep
DEM / rigid-body step path.

The compute path is libstardust_hip.so (hand-written HIP kernels behind the C
ABI in include/sd_api.h); this package only drives it.
"""
from . import _lib
from ._lib import ContactParams, DeviceBuffer, Grid, SdError, SphereParams
from .scene import SphereScene

__all__ = ["_lib", "ContactParams", "DeviceBuffer", "Grid", "SdError", "SphereParams", "SphereScene"]

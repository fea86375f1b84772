"""MI355X-native RMPCDMD solvent/colloid time step.

The arithmetic lives in librmpcd_amd.so (HIP, gfx950) behind the C ABI of
include/rmpcd_capi.h; ``capi`` binds it, ``system`` holds device-resident particle and
cell sets, ``sim`` runs whole time steps.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]

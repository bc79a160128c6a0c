"""MI355X-native Grains3D DEM time step (PacIFiC's granular solver path).

The compute path is libgrains_gpu.so (hand-written HIP for gfx950) behind the C
ABI in include/grains_gpu.h; :mod:`pacific_amd.grains` is its Python binding.
"""
from .grains import (ADAMS_BASHFORTH2, FIRST_ORDER_EXPLICIT, FP64, LEAPFROG, MIXED, SECOND_ORDER_EXPLICIT,  # noqa: F401
                     GrainsBatch, GrainsError, GrainsGPU, lib)

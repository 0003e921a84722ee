"""cfd_cavity_amd — MI355X (gfx950) drop-in for the uniform-grid thermal LBM
time step of bbrangeo/cfd_cavity (src/uniform).

The compute runs in libcavlbm.so (hand-written HIP kernels behind the C-ABI in
include/cavlbm.h); this package is its host-side mirror of the reference's
``CavitySolver`` API.  There is no CPU fallback.
"""
from .solver import CavitySolver, DELTA_T, T_COLD, T_HOT  # noqa: F401
from . import run, snapshot  # noqa: F401
from .batch import CavityBatch  # noqa: F401
from .octree import OctreeCavitySolver  # noqa: F401

__all__ = ["CavitySolver", "CavityBatch", "OctreeCavitySolver", "T_HOT", "T_COLD", "DELTA_T",
           "run", "snapshot"]

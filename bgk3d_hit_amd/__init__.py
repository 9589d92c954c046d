"""MI355X (gfx950) implementation of the BGK3D_HIT time-step hot path.

The compute path is libbgk3d_hit.so (hand-written HIP kernels behind the
C-ABI in include/bgk3d_hit.h).  This package is its Python host binding;
the Fortran host driver lives in bgk3d_hit_amd/fortran/.
"""
from ._lib import BGKError, load  # noqa: F401
from .lattice import (Lattice, device_count, hit_velocity, omega_from_svisc,  # noqa: F401
                      svisc_from_omega)

__all__ = ["Lattice", "BGKError", "hit_velocity", "device_count", "omega_from_svisc",
           "svisc_from_omega", "load"]

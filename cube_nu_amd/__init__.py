"""cube_nu_amd -- MI355X-native P3M force-and-update step of CUBE.

The compute path is the HIP library ``lib/libcube_p3m.so`` (C-ABI in
``include/cube_p3m.h``); :mod:`cube_nu_amd.p3m` mirrors the reference's
``particle_mesh`` / ``kick`` / ``drift`` call surface over it.
"""
from .p3m import (CUBE_OK, CubeError, Geometry, P3M, RATIO_SF, load_library, LIB_PATH, ABI_FUNCTIONS)  # noqa: F401

__all__ = ["P3M", "Geometry", "CubeError", "load_library", "LIB_PATH", "ABI_FUNCTIONS", "RATIO_SF"]

"""MI355X (gfx950) non-rigid fusion step: GN warp solve, DQ-warped TSDF
integration and marching cubes behind a C ABI (include/nrf.h)."""
from ._lib import (CORR, DUALQUAT, NODE, VOXEL, NrfError, load)  # noqa: F401

__all__ = ["CORR", "DUALQUAT", "NODE", "VOXEL", "NrfError", "load"]

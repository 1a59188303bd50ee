"""MI355X-native D2Q9 lattice Boltzmann step (the lbm-wgpu crate's simulation
core) on gfx950: hand-written HIP kernels behind a C ABI (include/lbm_hip.h),
with this package as the host-side mirror of the reference's ``LBM`` type."""
from ._native import (BOUNDARY_REFERENCE, BOUNDARY_ZERO_GRADIENT, PRECISION_F16E, PRECISION_F16S,
                      PRECISION_F32, LbmError, load)
from .batch import Batch, gen_batch, gen_shapes
from .f64 import LBM64
from . import shapes
from .lbm import LBM, ColorMap, SummaryStat, gen_circles, points_to_pairs, save_ppm, splitmix64, to_rgb8

__all__ = [
    "LBM", "LBM64", "shapes", "Batch", "ColorMap", "save_ppm", "to_rgb8", "gen_batch", "gen_shapes", "SummaryStat", "LbmError", "gen_circles", "points_to_pairs", "splitmix64", "load",
    "PRECISION_F32", "PRECISION_F16S", "PRECISION_F16E", "BOUNDARY_REFERENCE", "BOUNDARY_ZERO_GRADIENT",
]

"""lvr2_amd: MI355X-native AdaptiveKSearchSurface path of lvr2 (LBVH, exact kNN, PCA normals,
normal interpolation, Hoppe signed distance) behind a CudaSurface-shaped API."""
from ._lib import EXPORTED, LvrHipError, lib
from .surface import HipSurface

__all__ = ["HipSurface", "LvrHipError", "lib", "EXPORTED"]

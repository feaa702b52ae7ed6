"""MI355X-native (gfx950) implementation of SimpleWindErosion's erosion hot path.

World::erode -> Wind::fly -> World::cascade run as hand-written HIP kernels in
libswe.so behind the C-ABI of include/swe.h; this package is the Python mirror
of the reference's World/Wind interface over that library.
"""
from . import capi
from .world import Batch, SweError, Wind, World, generate_terrain, glm_normalize

__all__ = ["capi", "Batch", "SweError", "Wind", "World", "generate_terrain", "glm_normalize"]

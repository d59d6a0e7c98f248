"""MI355X-native differentiable Gaussian tile rasterizer (gfx950 HIP kernels
behind a C ABI), a drop-in for the rasterizer that Word2World's Renderer.py
calls (cudaForwardBackwardRenderer/_C)."""
from . import _lib, camera, raster, scene  # noqa: F401

__all__ = ["_lib", "camera", "raster", "scene"]

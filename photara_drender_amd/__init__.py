"""photara_drender for AMD Instinct MI355X: the differentiable raster / interpolate /
texture chain and the atlas refiner of photara_drender, as HIP kernels behind the
reference's API (python/photara_drender/__init__.py:10-60 for the exported names).

Importing the package loads the native module ``_photara_drender`` (pybind11 over
libpdr.so); it fails loudly if the native build is missing.
"""
try:  # torch first: libpdr then binds to torch's HIP runtime (one runtime per process)
    import torch as _torch  # noqa: F401
except ImportError:
    _torch = None

from ._photara_drender import DeviceInfo, Rasterizer, TextureBaker, enumerate_devices
from .bake import pad_texture_atlas

try:
    from .torch import interpolate, rasterize, texture
    from .atlas import (
        AtlasOptimizationOptions,
        TexturedRender,
        atlas_total_variation,
        masked_charbonnier_loss,
        optimize_texture_atlas,
        render_textured_mesh,
        seam_consistency_loss,
    )
except ImportError:  # torch missing: numpy API only, as in the reference
    AtlasOptimizationOptions = None
    TexturedRender = None
    atlas_total_variation = None
    interpolate = None
    masked_charbonnier_loss = None
    optimize_texture_atlas = None
    rasterize = None
    render_textured_mesh = None
    seam_consistency_loss = None
    texture = None

__all__ = [
    "AtlasOptimizationOptions",
    "DeviceInfo",
    "Rasterizer",
    "TextureBaker",
    "TexturedRender",
    "atlas_total_variation",
    "enumerate_devices",
    "interpolate",
    "masked_charbonnier_loss",
    "optimize_texture_atlas",
    "pad_texture_atlas",
    "rasterize",
    "render_textured_mesh",
    "seam_consistency_loss",
    "texture",
]
__version__ = "0.4.0"

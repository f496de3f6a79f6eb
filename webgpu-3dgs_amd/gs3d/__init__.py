"""gs3d — MI355X (gfx950) rebin-then-rasterize 3D Gaussian Splatting.

Python host over libgs3d.so (include/gs3d.h). The renderer, camera and scene
modules mirror the reference's GaussianSplatRenderer, Camera and loader.
"""
from . import camera, scene  # noqa: F401
from ._lib import Gs3dError, lib  # noqa: F401


def renderer(device: int = 0):
    from .renderer import GaussianSplatRenderer

    return GaussianSplatRenderer(device)

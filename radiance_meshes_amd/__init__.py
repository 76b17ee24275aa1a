"""MI355X (gfx950) implementation of Radiance Meshes' differentiable tetrahedral tile
rasterizer and hash-grid field, behind the reference's Python interface.

`install_reference_aliases()` registers this package's `delaunay_rasterization` mirror (and its
Delaunay step as `gdel3d`) under the reference's import paths, so `train.py` / `evaluate.py` style
callers run on it unchanged.
"""
from __future__ import annotations

import sys

__all__ = ["install_reference_aliases"]


def install_reference_aliases():
    """Make `import delaunay_rasterization...` resolve to radiance_meshes_amd's mirror."""
    import importlib
    names = [
        "delaunay_rasterization",
        "delaunay_rasterization.internal",
        "delaunay_rasterization.internal.render_grid",
        "delaunay_rasterization.internal.util",
        "delaunay_rasterization.internal.tile_shader_slang",
        "delaunay_rasterization.internal.alphablend_tiled_slang_interp",
        "delaunay_rasterization.internal.sort_by_keys",
        "delaunay_rasterization.internal.render_err",
    ]
    for n in names:
        sys.modules[n] = importlib.import_module("radiance_meshes_amd." + n)
    # gdel3d's Del (models/ingp_color.py:5) -> the GPU Delaunay
    sys.modules["gdel3d"] = importlib.import_module("radiance_meshes_amd.delaunay")

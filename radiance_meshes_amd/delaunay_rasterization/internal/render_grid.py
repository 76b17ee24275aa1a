"""delaunay_rasterization/internal/render_grid.py:20-27"""
from radiance_meshes_amd.rasterizer import RenderGrid, ceil_div  # noqa: F401

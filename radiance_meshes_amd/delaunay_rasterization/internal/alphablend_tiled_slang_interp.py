"""delaunay_rasterization/internal/alphablend_tiled_slang_interp.py: AlphaBlendTiledRender."""
from radiance_meshes_amd.rasterizer import AlphaBlendTiledRender, RenderGrid, vertex_and_tile_shader  # noqa: F401

"""delaunay_rasterization/internal/tile_shader_slang.py: vertex_and_tile_shader, VertexShader."""
from radiance_meshes_amd.rasterizer import VertexShader, ceil_div, vertex_and_tile_shader  # noqa: F401

"""delaunay_rasterization/internal/util.py:3-23"""
from radiance_meshes_amd.rasterizer import recombine_tensors, split_tensors  # noqa: F401

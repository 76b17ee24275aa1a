"""delaunay_rasterization/internal/sort_by_keys: `sort_by_keys_cub.sort_by_keys(keys, values, msb)`."""
from radiance_meshes_amd.rasterizer import sort_by_keys_cub  # noqa: F401

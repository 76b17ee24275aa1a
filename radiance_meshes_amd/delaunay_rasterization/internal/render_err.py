"""delaunay_rasterization/internal/render_err.py: render_err, gaussian_blur (votes via calc_tet_err)."""
from radiance_meshes_amd.render_err import calc_tet_err, gaussian_blur, render_err, ssim  # noqa: F401

"""Mirror of the reference package `delaunay_rasterization` backed by the HIP kernels."""

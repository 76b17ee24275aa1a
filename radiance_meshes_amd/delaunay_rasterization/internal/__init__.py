"""Mirror of `delaunay_rasterization.internal` (same module names and entry points)."""

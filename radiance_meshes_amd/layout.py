"""Spatial layout of a tetrahedral mesh in HBM.

The reference keeps vertices and tets in whatever order the Delaunay step produced
(models/ingp_color.py:522-558). Reordering both along a Morton (Z-order) curve makes the
per-tet vertex gathers of projection and compositing hit neighbouring cache lines and lets the
hash-grid backward find runs of equal table slots within a wavefront. Call after every
re-triangulation; the permutations map new ids to old ones.
"""
from __future__ import annotations

import numpy as np


def _spread10(v):
    v = v.astype(np.uint64) & np.uint64(0x3FF)
    v = (v | (v << np.uint64(16))) & np.uint64(0x030000FF)
    v = (v | (v << np.uint64(8))) & np.uint64(0x0300F00F)
    v = (v | (v << np.uint64(4))) & np.uint64(0x030C30C3)
    v = (v | (v << np.uint64(2))) & np.uint64(0x09249249)
    return v


def morton_codes(p: np.ndarray, lo=None, hi=None) -> np.ndarray:
    """30-bit Morton codes of points quantised to a 1024^3 grid over [lo, hi]."""
    p = np.asarray(p, np.float64)
    lo = p.min(0) if lo is None else np.asarray(lo, np.float64)
    hi = p.max(0) if hi is None else np.asarray(hi, np.float64)
    q = np.clip((p - lo) / np.maximum(hi - lo, 1e-30) * 1023.0, 0, 1023).astype(np.uint64)
    return _spread10(q[:, 0]) | (_spread10(q[:, 1]) << np.uint64(1)) | (_spread10(q[:, 2]) << np.uint64(2))


def spatial_layout(vertices: np.ndarray, indices: np.ndarray):
    """Returns (vertices', indices', vertex_perm, tet_perm): vertices sorted by Morton code of their
    position, tets by the Morton code of their centroid (over the bounding box of the vertices
    between the 1st and 99th percentile, so far-away bounding vertices do not flatten the curve)."""
    v = np.asarray(vertices)
    lo, hi = np.percentile(v, 1, axis=0), np.percentile(v, 99, axis=0)
    vperm = np.argsort(morton_codes(v, lo, hi), kind="stable")
    inv = np.empty_like(vperm)
    inv[vperm] = np.arange(len(vperm))
    idx = inv[np.asarray(indices)].astype(np.int32)
    cen = v[vperm][idx].mean(1)
    tperm = np.argsort(morton_codes(cen, lo, hi), kind="stable")
    return np.ascontiguousarray(v[vperm]), np.ascontiguousarray(idx[tperm]), vperm, tperm

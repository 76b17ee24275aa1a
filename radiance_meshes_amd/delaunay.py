"""Delaunay tetrahedralisation on the GPU (rm_delaunay), behind the reference's triangulation API.

models/ingp_color.py:523-544 (`update_triangulation`) triangulates the model's vertices with
gdel3d's `Del(V).compute(points)` -- whose indices may name the point at infinity, dropped by the
caller -- or with `scipy.spatial.Delaunay` when `high_precision`, then swaps the first two columns
of negatively oriented tets (`topo_utils.tet_volumes`, utils/topo_utils.py:9-24).
`delaunay_tets` does all of that on the device in one call. `Del` mirrors the gdel3d object the
reference imports (`from gdel3d import Del`, models/ingp_color.py:5; used as in tests/test_del.py:
19-27), so that import resolves here after `install_reference_aliases()`.

Coordinates are taken as float32, the dtype of the model's vertices (the reference hands gdel3d
their float64 copy, which is exact). Duplicate points keep one cell: the copy with the lowest index.
"""
from __future__ import annotations

import numpy as np
import torch

from ._capi import check, lib, ptr, stream_of

INFO_FIELDS = ("n_tets", "voronoi_vertices", "overflow_cells", "duplicates", "exact_evaluations",
               "exact_out_of_range", "retried_cells", "max_cell_vertices")


def delaunay_tets(points: torch.Tensor, capacity: int | None = None, return_info: bool = False):
    """Delaunay tets of points [N, 3] as int32 [T, 4] on the points' GPU, positively oriented.

    Raises RuntimeError when a cell exceeded its capacity or the cells disagree (a tet found by
    fewer than all four of its vertices), never returning a partial triangulation."""
    if points.dim() != 2 or points.shape[1] != 3:
        raise ValueError(f"points must be [N, 3], got {tuple(points.shape)}")
    dev = points.device if points.is_cuda else torch.device("cuda")
    pts = points.detach().to(device=dev, dtype=torch.float32).contiguous()
    if not torch.isfinite(pts).all():
        raise ValueError("points must be finite")
    n = int(pts.shape[0])
    cap = int(capacity) if capacity is not None else 8 * n + 1024
    L = lib()
    info = torch.zeros(len(INFO_FIELDS), dtype=torch.int64, device=dev)
    while True:
        ws = torch.empty(max(1, L.rm_delaunay_workspace_size(n, cap)), dtype=torch.uint8, device=dev)
        tets = torch.empty((max(cap, 1), 4), dtype=torch.int32, device=dev)
        check(L.rm_delaunay(ptr(pts), n, ptr(tets), cap, ptr(info), ptr(ws), ws.numel(), stream_of(dev)),
              "rm_delaunay")
        vals = info.tolist()
        if vals[0] <= cap:
            break
        cap = int(vals[0])
    d = dict(zip(INFO_FIELDS, vals))
    if d["overflow_cells"] or d["voronoi_vertices"] != 4 * d["n_tets"]:
        raise RuntimeError(f"rm_delaunay: cells disagree or overflowed: {d}")
    out = tets[:d["n_tets"]]
    return (out, d) if return_info else out


def update_triangulation(vertices: torch.Tensor) -> torch.Tensor:
    """The triangulation part of models/ingp_color.py:523-544 (both branches): int32 [T, 4] tets
    of the vertices, on their device, with positive volume and no tet at infinity."""
    return delaunay_tets(vertices)


class _Prev:
    """Stand-in for gdel3d's previous-triangulation state (returned by Del.compute)."""

    def __init__(self, tets: torch.Tensor):
        self.tets = tets

    def check_correctness(self, points) -> bool:
        """Positive orientation of every tet and every triangle shared by at most two tets."""
        p = np.asarray(points.detach().cpu() if hasattr(points, "detach") else points, np.float64)
        t = self.tets.cpu().numpy().astype(np.int64)
        if len(t) == 0:
            return True
        a, b, c, d = (p[t[:, k]] for k in range(4))
        if not (np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a) > 0).all():
            raise AssertionError("negatively oriented tet")
        faces = np.sort(np.concatenate([t[:, [1, 2, 3]], t[:, [0, 2, 3]], t[:, [0, 1, 3]], t[:, [0, 1, 2]]]), axis=1)
        _, counts = np.unique(faces, axis=0, return_counts=True)
        if counts.max() > 2:
            raise AssertionError("triangle shared by more than two tets")
        return True


class Del:
    """gdel3d.Del as the reference calls it: `indices, prev = Del(V).compute(points)`, indices a CPU
    int32 tensor [T, 4] (here without tets at infinity and already positively oriented)."""

    def __init__(self, n_points: int):
        self.n_points = int(n_points)

    def compute(self, points):
        pts = torch.as_tensor(points)
        if pts.shape[0] != self.n_points:
            raise ValueError(f"Del({self.n_points}) given {pts.shape[0]} points")
        tets = delaunay_tets(pts)
        return tets.cpu(), _Prev(tets)

    def compute_from_prev(self, points, prev):
        """The reference's warm-started variant; the GPU construction does not need the warm start."""
        return self.compute(points)

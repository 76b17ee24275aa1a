"""Writes the golden fixtures tests/golden/*.npz from the CPU oracle.

The reference ships no fixtures for this path (its tests compare JAX quadrature against the Slang
renderer at 1e-1 on the fly). These fixtures restage the scenes of its hot-path tests with the
oracle's FP64 outputs:
  single_tet: one tet, vertices U(-r, r) (r = 0.2), camera inside the tet at a random barycentric
              point with a random rotation (tests/single_tet_test.py:63-82), features
              [10.0, 0.25, 0.5, 0.1, -0.15, 0.25, 0.1] (single_tet_test.py:43-47), 32x32, 16x16 tiles
  multi_tet:  20 points in a ball of radius 10, scipy Delaunay, camera at the origin looking down
              a random direction (tests/multi_tet_test.py:24-47,90-110), density 1, 16x16 image,
              8x8 tiles
Inputs are seeded with the reference tests' seed 189710234. Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402
from radiance_meshes_amd import scenes  # noqa: E402

SEED = 189710234


def view_from(origin, R):
    """w2v = inverse([R | origin]) as the reference tests build viewmat (single_tet_test.py:70-76)."""
    c2w = np.eye(4)
    c2w[:3, :3] = R
    c2w[:3, 3] = origin
    return np.linalg.inv(c2w).astype(np.float32)


def random_rotation(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def camera(w2v, H, W, fov_deg):
    f = W / (2 * np.tan(np.radians(fov_deg) / 2))
    K = np.array([[f, 0, W / 2], [0, f, H / 2], [0, 0, 1]], np.float32)
    cam_pos = np.linalg.inv(w2v.astype(np.float64))[:3, 3].astype(np.float32)
    fov = np.radians(fov_deg)
    return scenes.Camera(w2v, K, cam_pos, H, W, fov, fov)


def render(name, V, idx, F, cam, tile, min_t, seed):
    V = V.astype(np.float32).astype(np.float64)   # the stored (float32) inputs are the rendered ones
    F = F.astype(np.float32).astype(np.float64)
    ocam = O.make_camera(cam, tile, min_t)
    t, r, vs, cc = O.project(ocam, idx, V.astype(np.float32))
    keys, vals, ranges = O.bin_tiles(ocam, t, r, vs)
    H, W = cam.height, cam.width
    rng = np.random.default_rng(seed)
    jit = rng.uniform(0, 1, (H, W, 2)).astype(np.float32)
    img, dist, nc, margin = O.composite_fwd(ocam, vals, ranges, idx, V, F, jit)
    wi = rng.standard_normal((H, W, 4))
    wd = rng.standard_normal((H, W, 5)) * 0.1
    gv, gf = O.composite_bwd(ocam, vals, ranges, idx, V, F, jit, img, dist, nc, wi, wd)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), vertices=V.astype(np.float32), indices=idx,
                        features=F.astype(np.float32), w2v=cam.w2v, K=cam.K, cam_pos=cam.cam_pos,
                        hw=np.array([H, W]), tile=tile, min_t=min_t, jitter=jit, tiles=t, rect=r,
                        depth=vs[:, 2], circumcenter=cc, sorted_idx=vals, ranges=ranges, img=img, dist=dist,
                        n_contrib=nc, margin=margin, w_img=wi, w_dist=wd, grad_vertices=gv, grad_features=gf)
    print(name, "tets", len(idx), "pairs", len(vals), "max|img|", np.abs(img).max())


def main():
    rng = np.random.default_rng(SEED)
    # single tet, camera inside
    r = 0.2
    V = ((rng.uniform(0, 1, (4, 3)) * 2 - 1) * r).astype(np.float32).astype(np.float64)
    idx = np.array([[0, 1, 2, 3]], np.int32)
    b = rng.uniform(0, 1, 4)
    b /= b.sum()
    origin = V.T @ b
    cam = camera(view_from(origin, random_rotation(rng)), 32, 32, 60.0)
    F = np.array([[10.0, 0.25, 0.5, 0.1, -0.15, 0.25, 0.1]])
    render("single_tet", V, idx, F, cam, 16, 0.0, SEED)
    # multi tet: 20 points in a ball of radius 10, camera at the centre
    pts = (rng.uniform(0, 1, (200, 3)) * 2 - 1) * 10
    pts = pts[np.linalg.norm(pts, axis=1) <= 10][:20].astype(np.float32).astype(np.float64)
    tets = scenes.tetrahedralize(pts)
    base = rng.uniform(0, 1, (len(tets), 3))
    grd = (2 * rng.uniform(0, 1, (len(tets), 3)) - 1) / np.sqrt(3) * 0.1
    F = np.concatenate([np.ones((len(tets), 1)), base, grd], 1)
    cam = camera(view_from(np.zeros(3), random_rotation(rng)), 16, 16, 70.0)
    render("multi_tet", pts, tets, F, cam, 8, 0.1, SEED + 1)


if __name__ == "__main__":
    main()

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
  face_view / face_view_in / frustum / multi_views: the parameter grids of test_face_view (and the
              same with the camera looking into the tet), test_frustum_point
              (single_tet_test.py:143-217), test_center_view and test_outside_view
              (multi_tet_test.py:80-122), several cases per file (keys '<case>/<field>')
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


def case(V, idx, F, cam, tile, min_t, seed, jitter=None):
    """Inputs and the oracle's outputs of one scene: projection, binning, FP64 image and adjoint of
    sum(img * w_img) + sum(dist * w_dist). jitter None: random per-pixel jitter (seeded)."""
    V = V.astype(np.float32).astype(np.float64)   # the stored (float32) inputs are the rendered ones
    F = F.astype(np.float32).astype(np.float64)
    ocam = O.make_camera(cam, tile, min_t)
    t, r, vs, cc = O.project(ocam, idx, V.astype(np.float32))
    keys, vals, ranges = O.bin_tiles(ocam, t, r, vs)
    H, W = cam.height, cam.width
    rng = np.random.default_rng(seed)
    jit = rng.uniform(0, 1, (H, W, 2)).astype(np.float32) if jitter is None else \
        np.full((H, W, 2), jitter, np.float32)
    img, dist, nc, margin = O.composite_fwd(ocam, vals, ranges, idx, V, F, jit)
    wi = rng.standard_normal((H, W, 4)).astype(np.float32)          # stored as float32
    wd = (rng.standard_normal((H, W, 5)) * 0.1).astype(np.float32)
    gv, gf = O.composite_bwd(ocam, vals, ranges, idx, V, F, jit, img, dist, nc, wi.astype(np.float64),
                             wd.astype(np.float64))
    return dict(vertices=V.astype(np.float32), indices=np.asarray(idx, np.int32), features=F.astype(np.float32),
                w2v=cam.w2v, K=cam.K, cam_pos=cam.cam_pos, hw=np.array([H, W]), tile=tile, min_t=min_t,
                jitter=jit, tiles=t, rect=r, depth=vs[:, 2], circumcenter=cc, sorted_idx=vals, ranges=ranges,
                img=img, dist=dist, n_contrib=nc, margin=margin, w_img=wi, w_dist=wd, grad_vertices=gv,
                grad_features=gf)


def render(name, V, idx, F, cam, tile, min_t, seed):
    c = case(V, idx, F, cam, tile, min_t, seed)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **c)
    print(name, "tets", len(idx), "pairs", len(c["sorted_idx"]), "max|img|", np.abs(c["img"]).max())


def write_family(name, cases):
    """Several cases in one file, keys '<case>/<field>'."""
    out = {}
    for cname, c in cases:
        for k, v in c.items():
            out[f"{cname}/{k}"] = v
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    print(name, len(cases), "cases:", ", ".join(f"{n} ({int(c['tiles'].astype(bool).sum())} visible, "
                                                 f"max|img| {np.abs(c['img']).max():.3f})" for n, c in cases))


def harness_camera(w2v, H, W, fov_deg=90.0):
    """The camera of the reference's test harness (utils/compare_quad.py:36-57): focal length
    H / tan(fov / 2) for both axes, principal point at the image centre."""
    f = H / np.tan(np.radians(fov_deg) / 2)
    K = np.array([[f, 0, W / 2], [0, f, H / 2], [0, 0, 1]], np.float32)
    cam_pos = np.linalg.inv(w2v.astype(np.float64))[:3, 3].astype(np.float32)
    return scenes.Camera(w2v.astype(np.float32), K, cam_pos, H, W, 2 * np.arctan(W / (2 * f)),
                         2 * np.arctan(H / (2 * f)))


def circumcenters(tets):
    """Circumcenters of [T,4,3] tets (the formula of utils/topo_utils.py:120-160, in FP64)."""
    a, b, c = tets[:, 1] - tets[:, 0], tets[:, 2] - tets[:, 0], tets[:, 3] - tets[:, 0]
    num = ((a * a).sum(1, keepdims=True) * np.cross(b, c) + (b * b).sum(1, keepdims=True) * np.cross(c, a)
           + (c * c).sum(1, keepdims=True) * np.cross(a, b))
    return tets[:, 0] + num / (2 * (a * np.cross(b, c)).sum(1, keepdims=True))


def offset_normalize(rgb, grd, cc, tets):
    """utils/model_util.py:81-91 in FP64: gradient scaled by the smallest colour channel over the
    circumradius, colour moved from the circumcenter to vertex 0."""
    g = grd * rgb.min(1, keepdims=True)
    ng = g / np.linalg.norm(tets[:, 0] - cc, axis=1, keepdims=True)
    return rgb + (ng * (tets[:, 0] - cc)).sum(1, keepdims=True), ng


def rainbow(n):
    """plt.cm.rainbow(np.linspace(0, 1, n))[:, :3] (utils/test_util.py:23-28): the 256-entry table
    of r = |2x - 1/2|, g = sin(pi x), b = cos(pi x / 2), clipped to [0, 1]."""
    lut_x = np.linspace(0, 1, 256)
    lut = np.clip(np.stack([np.abs(2 * lut_x - 0.5), np.sin(np.pi * lut_x), np.cos(np.pi * lut_x / 2)], 1), 0, 1)
    i = np.minimum((np.linspace(0, 1, n) * 256).astype(int), 255)
    return lut[i]


def single_tet_features(V):
    """[density, colour, gradient] of single_tet_test.py:43-56: the inline raw values moved to
    vertex 0 by offset_normalize around the circumcenter."""
    raw = np.array([[10.0, 0.25, 0.5, 0.1, -0.15, 0.25, 0.1]])
    tets = V[None]
    col, ng = offset_normalize(raw[:, 1:4], raw[:, 4:7], circumcenters(tets), tets)
    return np.concatenate([raw[:, :1], col, ng], 1)


def face_view_cases():
    """tests/single_tet_test.py:143-186 (test_face_view): a tet of vertices U(-r, r), the camera on
    a random face at a random barycentric point, moved along the face normal by U(-1, 1) x offset,
    looking along the normal; 32x32, 16x16 tiles, the harness camera (fov 90, tmin 0.01). One
    case per (offset, radius) of the test's grid."""
    rng = np.random.default_rng(SEED + 10)
    out = []
    idx = np.array([[0, 1, 2, 3]], np.int32)
    for off in (0, 0.1, 1, 5, 10):
        for radius in (0.05, 0.1, 0.2, 0.4):
            V = ((rng.uniform(0, 1, (4, 3)) * 2 - 1) * radius).astype(np.float32).astype(np.float64)
            fv = [[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]][int(rng.integers(0, 4))]
            bary = rng.uniform(0, 1, 3)
            bary /= bary.sum()
            fvs = V[fv]
            point = fvs.T @ bary
            n = np.cross(fvs[1] - fvs[0], fvs[2] - fvs[0])
            n /= np.linalg.norm(n)
            pos = point + n * (rng.uniform() * 2 - 1) * off
            right = np.cross(n, rng.standard_normal(3))
            right /= np.linalg.norm(right)
            up = np.cross(right, n)
            c2w = np.eye(4)
            c2w[:3, 0], c2w[:3, 1], c2w[:3, 2], c2w[:3, 3] = right, up, n, pos
            cam = harness_camera(np.linalg.inv(c2w), 32, 32)
            out.append((f"off{off}_r{radius}", case(V, idx, single_tet_features(V), cam, 16, 0.01,
                                                    SEED + 11, jitter=0.5)))
    return out


def face_view_in_cases():
    """The face-view scene with the camera looking into the tet: on the face plane or outside it
    by the offset (0, 0.1, 1, 5, 10, along the face's outward normal), looking along the inward
    normal, so the face is seen head-on; at offset 0 the rays start on the face plane itself,
    which drives the crossing's near-parallel / grazing arithmetic (intersect.slang:393-624) and
    the near-plane clip of the projection (camera.slang:443-508)."""
    rng = np.random.default_rng(SEED + 15)
    out = []
    idx = np.array([[0, 1, 2, 3]], np.int32)
    for off in (0, 0.1, 1, 5, 10):
        for radius in (0.05, 0.1, 0.2, 0.4):
            V = ((rng.uniform(0, 1, (4, 3)) * 2 - 1) * radius).astype(np.float32).astype(np.float64)
            fi = int(rng.integers(0, 4))
            fv = [[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]][fi]
            bary = rng.uniform(0, 1, 3)
            bary /= bary.sum()
            fvs = V[fv]
            n = np.cross(fvs[1] - fvs[0], fvs[2] - fvs[0])
            n /= np.linalg.norm(n)
            if n @ (V[fi] - fvs[0]) > 0:       # outward: away from the opposite vertex
                n = -n
            pos = fvs.T @ bary + n * off
            fwd = -n
            right = np.cross(fwd, rng.standard_normal(3))
            right /= np.linalg.norm(right)
            up = np.cross(right, fwd)
            c2w = np.eye(4)
            c2w[:3, 0], c2w[:3, 1], c2w[:3, 2], c2w[:3, 3] = right, up, fwd, pos
            cam = harness_camera(np.linalg.inv(c2w), 32, 32)
            out.append((f"off{off}_r{radius}", case(V, idx, single_tet_features(V), cam, 16, 0.01,
                                                    SEED + 16, jitter=0.5)))
    return out


def frustum_cases():
    """tests/single_tet_test.py:188-217 (test_frustum_point): a tet of vertices U(-1, 1) moved to a
    point of the view frustum at depth U(d/2, d), viewmat = [random rotation | N(0,1) origin]
    (used as given, not inverted), for d in {0, 1, 5, 10}; two cases per depth."""
    rng = np.random.default_rng(SEED + 20)
    out = []
    idx = np.array([[0, 1, 2, 3]], np.int32)
    for depth in (0, 1, 5, 10):
        for k in range(2):
            V = (rng.uniform(0, 1, (4, 3)) * 2 - 1)
            z = rng.uniform() * depth / 2 + depth / 2
            x, y = (rng.uniform() * 2 - 1) * z, (rng.uniform() * 2 - 1) * z
            vm = np.eye(4)
            vm[:3, :3] = random_rotation(rng)
            vm[:3, 3] = rng.standard_normal(3)
            wp = vm @ np.array([x, y, z, 1.0])
            V = (V + wp[:3] / wp[3]).astype(np.float32).astype(np.float64)
            cam = harness_camera(vm, 32, 32)
            out.append((f"d{depth}_{k}", case(V, idx, single_tet_features(V), cam, 16, 0.01, SEED + 21,
                                              jitter=0.5)))
    return out


def point_cloud(rng, n, radius):
    p = (rng.uniform(0, 1, (n, 3)) * 2 - 1) * radius
    return p[np.linalg.norm(p, axis=1) <= radius]


def multi_tet_features(pts, tets):
    """tests/multi_tet_test.py:54-69: rainbow base colours and gradients (2 c - 1) / sqrt(3) of the
    second half of the palette, moved to vertex 0 around the circumcenter; density 1."""
    N = len(tets)
    allc = rainbow(2 * N)
    base, grd = allc[:N], (2 * allc[N:] - 1) / np.sqrt(3)
    tv = pts[tets]
    col, ng = offset_normalize(base, grd, circumcenters(tv), tv)
    return np.concatenate([np.ones((N, 1)), col, ng], 1)


def multi_view_cases():
    """tests/multi_tet_test.py:80-122: points in a ball (n in {10, 20} drawn, radius 10 and 100),
    scipy Delaunay; the camera at one of the points with a random rotation (test_center_view), or
    outside at radius + offset (offset 1, 10, 100) looking at the mean (test_outside_view);
    16x16, 8x8 tiles, the harness camera (fov 90, tmin 0.01)."""
    rng = np.random.default_rng(SEED + 30)
    out = []
    for n in (10, 20):
        for radius in (10, 100):
            pts = point_cloud(rng, n, radius)
            tets = scenes.tetrahedralize(pts)
            pts = pts.astype(np.float32).astype(np.float64)
            c2w = np.eye(4)
            c2w[:3, :3] = random_rotation(rng)
            c2w[:3, 3] = pts[int(rng.integers(0, len(pts)))]
            cam = harness_camera(np.linalg.inv(c2w), 16, 16)
            out.append((f"center_n{n}_r{radius}", case(pts, tets, multi_tet_features(pts, tets), cam, 8, 0.01,
                                                        SEED + 31, jitter=0.5)))
    for radius in (10, 100):
        for off in (1, 10, 100):
            pts = point_cloud(rng, 20, radius)
            tets = scenes.tetrahedralize(pts)
            pts = pts.astype(np.float32).astype(np.float64)
            d = rng.standard_normal(3)
            d /= np.linalg.norm(d)
            pos = d * (radius + off)
            fwd = pts.mean(0) - pos
            fwd /= np.linalg.norm(fwd)
            right = np.cross(fwd, [0.0, 1.0, 0.0])
            right /= np.linalg.norm(right)
            up = np.cross(right, fwd)
            c2w = np.eye(4)
            c2w[:3, 0], c2w[:3, 1], c2w[:3, 2], c2w[:3, 3] = right, up, fwd, pos
            cam = harness_camera(np.linalg.inv(c2w), 16, 16)
            out.append((f"outside_r{radius}_off{off}", case(pts, tets, multi_tet_features(pts, tets), cam, 8,
                                                             0.01, SEED + 32, jitter=0.5)))
    return out


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
    write_family("face_view", face_view_cases())
    write_family("face_view_in", face_view_in_cases())
    write_family("frustum", frustum_cases())
    write_family("multi_views", multi_view_cases())


if __name__ == "__main__":
    main()

"""Seeded synthetic workloads for the rasterizer: point sets, their Delaunay tetrahedralisation,
per-tet features and cameras.

The reference trains on COLMAP captures; these generators reproduce the synthetic scenes the
benchmarks and parity tests are quoted on (CPU Delaunay via scipy, tets re-oriented to positive
volume as models/ingp_color.py:538-542 does after every triangulation).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Camera:
    """Pinhole camera in the reference's tcam convention (data/camera.py:72-91):
    world_view_transform is the row-major 4x4 world->view matrix for column vectors, view space
    has +z forward, +x right, +y down; K = [[fx,0,W/2],[0,fy,H/2],[0,0,1]]."""
    w2v: np.ndarray
    K: np.ndarray
    cam_pos: np.ndarray
    height: int
    width: int
    fovx: float
    fovy: float
    distortion: np.ndarray = field(default_factory=lambda: np.zeros(4, np.float32))
    camera_type: int = 0

    def tcam(self, tile_size: int, min_t: float):
        """The reference's tcam dict (utils/train_util.py:31-38) with numpy/CPU values."""
        import torch
        gh = (self.height + tile_size - 1) // tile_size
        gw = (self.width + tile_size - 1) // tile_size
        return dict(tile_height=tile_size, tile_width=tile_size, grid_height=gh, grid_width=gw,
                    min_t=float(min_t),
                    world_view_transform=torch.from_numpy(self.w2v.astype(np.float32)),
                    cam_pos=torch.from_numpy(self.cam_pos.astype(np.float32)),
                    K=torch.from_numpy(self.K.astype(np.float32)),
                    image_height=self.height, image_width=self.width,
                    fovy=float(self.fovy), fovx=float(self.fovx),
                    distortion_params=torch.from_numpy(self.distortion.astype(np.float32)),
                    camera_type=int(self.camera_type))


def look_at_camera(eye, target, up, width, height, fovy_deg, fovx_deg=None) -> Camera:
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, np.float64))
    r /= np.linalg.norm(r)
    d = np.cross(f, r)
    R = np.stack([r, d, f])
    w2v = np.eye(4)
    w2v[:3, :3] = R
    w2v[:3, 3] = -R @ eye
    fovy = math.radians(fovy_deg)
    fy = height / (2 * math.tan(fovy / 2))
    if fovx_deg is None:
        fx = fy
        fovx = 2 * math.atan(width / (2 * fx))
    else:
        fovx = math.radians(fovx_deg)
        fx = width / (2 * math.tan(fovx / 2))
    K = np.array([[fx, 0, width / 2], [0, fy, height / 2], [0, 0, 1]], np.float64)
    w2v32 = w2v.astype(np.float32)
    # camera centre from the float32 matrix, as Camera.update() inverts the stored transform
    cam_pos = np.linalg.inv(w2v32.astype(np.float64))[:3, 3]
    return Camera(w2v32, K.astype(np.float32), cam_pos.astype(np.float32), height, width, fovx, fovy)


def orbit_camera(radius, azimuth_deg, elevation_deg, width, height, fovy_deg, target=(0, 0, 0)):
    az, el = math.radians(azimuth_deg), math.radians(elevation_deg)
    eye = np.array([radius * math.cos(el) * math.sin(az), radius * math.sin(el),
                    radius * math.cos(el) * math.cos(az)]) + np.asarray(target)
    return look_at_camera(eye, target, (0, 1, 0), width, height, fovy_deg)


def tetrahedralize(points: np.ndarray) -> np.ndarray:
    """scipy Delaunay; every tet re-oriented to positive signed volume."""
    from scipy.spatial import Delaunay
    tets = Delaunay(points).simplices.astype(np.int32)
    p = points.astype(np.float64)
    v = p[tets]
    vol = np.einsum("ij,ij->i", np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), v[:, 3] - v[:, 0])
    neg = vol < 0
    tets[neg, 2], tets[neg, 3] = tets[neg, 3].copy(), tets[neg, 2].copy()
    return np.ascontiguousarray(tets)


CUBE_CORNERS = np.array([[i, j, k] for i in (-1, 1) for j in (-1, 1) for k in (-1, 1)], np.float64)


def tet_features(n_tets: int, seed: int, n_features: int = 7, dtype=np.float32) -> np.ndarray:
    """density = exp(N(0,1))*2, base rgb ~ U(0,1), colour gradient rows ~ N(0, 0.1).
    n_features 7: one gradient row shared by rgb (the reference's layout); 13: a 3x3 gradient."""
    rng = np.random.default_rng(seed)
    dens = np.exp(rng.standard_normal(n_tets)) * 2.0
    base = rng.uniform(0, 1, (n_tets, 3))
    grd = rng.normal(0, 0.1, (n_tets, 3, 3))
    if n_features == 7:
        f = np.concatenate([dens[:, None], base, grd[:, 0]], 1)
    elif n_features == 13:
        f = np.concatenate([dens[:, None], base, grd.reshape(n_tets, 9)], 1)
    else:
        raise ValueError("n_features must be 7 or 13")
    return np.ascontiguousarray(f.astype(dtype))


@dataclass
class Scene:
    vertices: np.ndarray   # [V,3] f32 (or f64)
    indices: np.ndarray    # [T,4] i32
    features: np.ndarray   # [T,F]
    camera: Camera
    tile_size: int
    min_t: float
    name: str


def plain_small(n_points=10000, seed=0, tile_size=16, n_features=7, width=256, height=256) -> Scene:
    """10k points U[-1,1]^3 + 8 cube corners, 256x256 pinhole, 60 deg FOV, radius 3 on +z."""
    rng = np.random.default_rng(seed)
    pts = np.concatenate([rng.uniform(-1, 1, (n_points, 3)), CUBE_CORNERS])
    tets = tetrahedralize(pts)
    feats = tet_features(len(tets), seed + 1, n_features)
    cam = look_at_camera((0, 0, 3), (0, 0, 0), (0, 1, 0), width, height, 60.0)
    return Scene(pts.astype(np.float32), tets, feats, cam, tile_size, 0.1, "plain_10k_256")


def gmm_points(n_points: int, seed: int, n_comp: int = 32) -> np.ndarray:
    rng = np.random.default_rng(seed)
    means = rng.uniform(-2, 2, (n_comp, 3))
    stds = rng.uniform(0.05, 0.4, n_comp)
    comp = rng.integers(0, n_comp, n_points)
    pts = means[comp] + rng.standard_normal((n_points, 3)) * stds[comp, None]
    return np.concatenate([pts, CUBE_CORNERS * 20.0])


def gmm_scene(n_points=1_000_000, seed=0, tile_size=16, n_features=7, width=1280, height=840,
              azimuth=30.0, elevation=20.0, radius=6.0) -> Scene:
    """n points from a seeded 32-component Gaussian mixture (means U[-2,2]^3, std U(0.05,0.4))
    plus 8 far corners at +-20; orbit camera at radius 6, 50 deg vertical FOV."""
    pts = gmm_points(n_points, seed)
    tets = tetrahedralize(pts)
    feats = tet_features(len(tets), seed + 1, n_features)
    cam = orbit_camera(radius, azimuth, elevation, width, height, 50.0)
    return Scene(pts.astype(np.float32), tets, feats, cam, tile_size, 0.1, f"gmm_{n_points}_{width}x{height}")


def ball_scene(n_points=2000, seed=0, tile_size=16, n_features=7, size=64, dtype=np.float64) -> Scene:
    """n points uniform in the unit ball, 64x64 camera at radius 3 (differentiation checks)."""
    rng = np.random.default_rng(seed)
    pts = []
    while sum(len(p) for p in pts) < n_points:
        c = rng.uniform(-1, 1, (2 * n_points, 3))
        pts.append(c[np.linalg.norm(c, axis=1) <= 1.0])
    pts = np.concatenate(pts)[:n_points]
    tets = tetrahedralize(pts)
    feats = tet_features(len(tets), seed + 1, n_features, dtype)
    cam = look_at_camera((0.3, 0.4, 2.6), (0, 0, 0), (0, 1, 0), size, size, 50.0)
    return Scene(pts.astype(dtype), tets, feats, cam, tile_size, 0.1, f"ball_{n_points}_{size}")

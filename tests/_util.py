"""Shared helpers for the parity tests: scenes on the device, tcam dicts, oracle cameras."""
from __future__ import annotations

import functools

import numpy as np

from radiance_meshes_amd import scenes


@functools.lru_cache(maxsize=None)
def scene(kind: str, tile: int = 16, nf: int = 7, **kw):
    if kind == "plain":
        return scenes.plain_small(tile_size=tile, n_features=nf, **kw)
    if kind == "ball":
        return scenes.ball_scene(tile_size=tile, n_features=nf, dtype=np.float32, **kw)
    if kind == "gmm":
        return scenes.gmm_scene(tile_size=tile, n_features=nf, **kw)
    raise ValueError(kind)


def device_inputs(sc, jitter_seed=None):
    import torch
    dev = "cuda"
    verts = torch.from_numpy(np.ascontiguousarray(sc.vertices, np.float32)).to(dev)
    idx = torch.from_numpy(np.ascontiguousarray(sc.indices, np.int32)).to(dev)
    feats = torch.from_numpy(np.ascontiguousarray(sc.features, np.float32)).to(dev)
    H, W = sc.camera.height, sc.camera.width
    jit = jitter(sc, jitter_seed)
    return verts, idx, feats, torch.from_numpy(jit).to(dev)


def jitter(sc, seed=None):
    H, W = sc.camera.height, sc.camera.width
    if seed is None:
        return np.full((H, W, 2), 0.5, np.float32)
    return np.random.default_rng(seed).uniform(0, 1, (H, W, 2)).astype(np.float32)


def tcam_of(sc, device="cuda"):
    t = sc.camera.tcam(sc.tile_size, sc.min_t)
    return {k: (v.to(device) if hasattr(v, "to") else v) for k, v in t.items()}


def grid_of(sc):
    from radiance_meshes_amd.rasterizer import RenderGrid
    return RenderGrid(sc.camera.height, sc.camera.width, sc.tile_size, sc.tile_size)


def f32_bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.int32)


def with_camera(sc, kind: str):
    """A copy of scene sc seen through another camera model: 'distorted' (pinhole with radial and
    tangential distortion k1, k2, p1, p2), 'fisheye' (equidistant, camera_type 1), 'inside'
    (pinhole at the centre of the mesh) or 'away' (pinhole looking away from the mesh)."""
    import dataclasses
    cam = sc.camera
    if kind == "distorted":
        cam = dataclasses.replace(cam, distortion=np.array([0.08, -0.02, 0.004, -0.003], np.float32))
    elif kind == "fisheye":
        cam = dataclasses.replace(cam, camera_type=1)
    elif kind == "inside":
        cam = scenes.look_at_camera((0.05, -0.03, 0.02), (1.0, 0.2, -0.3), (0, 1, 0), cam.width, cam.height, 70.0)
    elif kind == "away":
        p = np.asarray(cam.cam_pos, np.float64)
        cam = scenes.look_at_camera(tuple(p), tuple(2 * p), (0, 1, 0), cam.width, cam.height, 40.0)
    else:
        raise ValueError(kind)
    return dataclasses.replace(sc, camera=cam, name=f"{sc.name}_{kind}")

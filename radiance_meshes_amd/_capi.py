"""ctypes binding of librmesh.so (the C ABI declared in include/rmesh.h).

Tensors cross the boundary as raw device pointers plus sizes; every call is enqueued on torch's
current HIP stream. A missing library is an error: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librmesh.so")

RM_ERRORS = {1: "invalid argument", 2: "unsupported tile size", 3: "workspace too small",
             4: "HIP error", 5: "capacity exceeded"}


class RmCamera(C.Structure):
    _fields_ = [("w2v", C.c_float * 16), ("K", C.c_float * 9), ("cam_pos", C.c_float * 3),
                ("dist", C.c_float * 4), ("min_t", C.c_float), ("fovy", C.c_float),
                ("fovx", C.c_float), ("image_height", C.c_int32), ("image_width", C.c_int32),
                ("grid_height", C.c_int32), ("grid_width", C.c_int32),
                ("tile_height", C.c_int32), ("tile_width", C.c_int32),
                ("camera_type", C.c_int32), ("_pad", C.c_int32)]


class RmScene(C.Structure):
    _fields_ = [("indices_dev", C.c_void_p), ("vertices_dev", C.c_void_p), ("features_dev", C.c_void_p),
                ("n_tets", C.c_int64), ("n_verts", C.c_int64), ("n_features", C.c_int32), ("_pad", C.c_int32)]


class RmBatchInfo(C.Structure):
    _fields_ = [("n_views", C.c_int64), ("total_tets", C.c_int64), ("total_tiles", C.c_int64),
                ("total_pixels", C.c_int64), ("total_workgroups", C.c_int64), ("plan_bytes", C.c_int64),
                ("tile", C.c_int32), ("n_features", C.c_int32), ("total_bin_blocks", C.c_int64)]


_P = C.c_void_p
_I64 = C.c_int64
_I32 = C.c_int32
_SZ = C.c_size_t

# name -> (restype, argtypes)
SIGNATURES = {
    "rm_last_error": (C.c_char_p, []),
    "rm_version": (C.c_int, []),
    "rm_project_fwd": (C.c_int, [C.POINTER(RmCamera), _P, _P, _I64, _I64, _P, _P, _P, _P, _P]),
    "rm_project_bwd": (C.c_int, [C.POINTER(RmCamera), _P, _P, _I64, _I64, _P, _P, _P, _P, _P]),
    "rm_scan_workspace_size": (_SZ, [_I64]),
    "rm_inclusive_scan_i32": (C.c_int, [_P, _P, _I64, _P, _SZ, _P]),
    "rm_generate_keys": (C.c_int, [_P, _P, _P, _I64, _I32, _I32, _P, _P, _P]),
    "rm_sort_pairs_workspace_size": (_SZ, [_I64]),
    "rm_sort_pairs": (C.c_int, [_P, _P, _P, _P, _I64, _I32, _P, _SZ, _P]),
    "rm_tile_ranges": (C.c_int, [_P, _I64, _P, _I64, _P]),
    "rm_bin_workspace_size": (_SZ, [_I64, _I64, _I64]),
    "rm_bin_count": (C.c_int, [_P, _I64, _P, _P, _SZ, _P]),
    "rm_bin_sort": (C.c_int, [_P, _P, _P, _I64, _I64, _I64, _I32, _I32, _P, _P, _P, _SZ, _P]),
    "rm_composite_fwd": (C.c_int, [C.POINTER(RmCamera), _P, _I64, _P, _P, _P, _P, _I32, _P, _P, _P, _P, _P]),
    "rm_composite_fwd_flags": (C.c_int, [C.POINTER(RmCamera), _P, _I64, _P, _P, _P, _P, _I32, _P, _P, _P, _P, _P,
                                         _P, _P]),
    "rm_set_intersect_precision": (C.c_int, [C.c_int]),
    "rm_composite_bwd": (C.c_int, [C.POINTER(RmCamera), _P, _I64, _P, _P, _P, _P, _I32, _P, _P, _P, _P,
                                   _P, _P, _P, _P, _P]),
    "rm_composite_bwd_workspace_size": (_SZ, [_I64]),
    "rm_composite_bwd_ws": (C.c_int, [C.POINTER(RmCamera), _P, _I64, _P, _P, _I64, _P, _P, _I32, _P, _P, _P, _P,
                                      _P, _P, _P, _P, _P, _SZ, _P]),
    "rm_composite_bwd_flags": (C.c_int, [C.POINTER(RmCamera), _P, _I64, _P, _P, _I64, _P, _P, _I32, _P, _P, _P, _P,
                                         _P, _P, _P, _P, _P, _P, _P, _SZ, _P]),
    "rm_tile_order_workspace_size": (_SZ, [_I64]),
    "rm_tile_order": (C.c_int, [_P, _I64, _P, _P, _SZ, _P]),
    "rm_composite_bwd_noalloc_max_tets": (_I64, []),
    "rm_composite_bwd_noalloc": (C.c_int, [C.POINTER(RmCamera), _P, _I64, _P, _P, _I64, _P, _P, _I32, _P, _P, _P,
                                           _P, _P, _P, _P, _P, _P]),
    "rm_activate_fwd": (C.c_int, [_P, _P, _P, _P, _P, _I32, _I32, _P, _P, _P, _I64, _P, _P]),
    "rm_activate_bwd": (C.c_int, [_P, _P, _P, _P, _P, _I32, _I32, _P, _P, _P, _I64, _P, _P, _P, _P, _P, _P, _P, _P]),
    "rm_tet_err": (C.c_int, [C.POINTER(RmCamera), _P, _I64, _P, _P, _P, _P, _I32, _P, _P, _P, _P, _I64, _P, _P]),
    "rm_batch_plan": (C.c_int, [C.POINTER(RmScene), _I32, C.POINTER(RmCamera), _P, _I32, _P, _SZ,
                                C.POINTER(RmBatchInfo)]),
    "rm_batch_workspace_size": (_SZ, [C.POINTER(RmBatchInfo), _I64]),
    "rm_batch_render": (C.c_int, [_P, C.POINTER(RmBatchInfo), _I64, _P, _P, _P, _P, _P, _SZ, _P]),
    "rm_hashgrid_fwd": (C.c_int, [_P, _I64, _P, _P, _P, _I32, _I32, _I32, _P, _P]),
    "rm_hashgrid_bwd": (C.c_int, [_P, _I64, _P, _P, _P, _I32, _I32, _I32, _P, _P]),
    "rm_hashgrid_bwd_workspace_size": (_SZ, [_I64, _P, _P, _I32, _I32, _I32]),
    "rm_hashgrid_bwd_ws": (C.c_int, [_P, _I64, _P, _P, _P, _I32, _I32, _I32, _P, _P, _SZ, _P]),
    "rm_mlp_fwd": (C.c_int, [_P, _I64, _P, _P, _P, _P, _P, _P, _I32, _I32, _P, _P]),
    "rm_mlp_bwd": (C.c_int, [_P, _I64, _P, _P, _P, _P, _P, _P, _I32, _I32, _P, _P, _P, _P, _P, _P, _P,
                             _P, _P]),
    "rm_field_head_fwd": (C.c_int, [_P, _I64, _I32, _P, _P, _P, _P, _P, _P, _I32, _I32, _P, _P]),
    "rm_field_head_bwd": (C.c_int, [_P, _I64, _I32, _P, _P, _P, _P, _P, _P, _I32, _I32, _P, _P, _P, _P, _P, _P,
                                    _P, _P, _P]),
    "rm_field_query": (C.c_int, [_P, _P, _I64, _P, C.c_float, _I32, _P, _P, _P, _P]),
    "rm_field_downweight": (C.c_int, [_P, _P, _I64, _I32, _I32, C.c_double, C.c_double, _P, _P]),
    "rm_field_outputs_fwd": (C.c_int, [_P, _P, _P, _P, _I64, _I32, C.c_float, _P, _P, _P, _P, _P]),
    "rm_field_outputs_bwd": (C.c_int, [_P, _P, _P, _P, _P, _P, _I64, _I32, _P, _P, _P, _P, _P]),
    "rm_field_points": (C.c_int, [_P, _P, _P, _I64, C.c_float, _P, _P]),
    "rm_field_scatter_fwd": (C.c_int, [_P, _P, _I64, _I32, _P, _P]),
    "rm_field_scatter_bwd": (C.c_int, [_P, _P, _I64, _I32, _P, _P, _P]),
    "rm_delaunay_workspace_size": (_SZ, [_I64, _I64]),
    "rm_delaunay": (C.c_int, [_P, _I64, _P, _I64, _P, _P, _SZ, _P]),
    "rm_delaunay_insphere_host": (C.c_int, [_P, _P]),
}

_lib = None
_lock = threading.Lock()


def lib():
    """Loads librmesh.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise RuntimeError(
                        f"{LIB_PATH} is missing: build it with `python -m radiance_meshes_amd.build` "
                        "(or __graft_entry__.build()); there is no non-HIP fallback")
                L = C.CDLL(LIB_PATH)
                for name, (res, args) in SIGNATURES.items():
                    fn = getattr(L, name)
                    fn.restype = res
                    fn.argtypes = args
                _lib = L
    return _lib


def check(rc: int, what: str):
    if rc != 0:
        msg = lib().rm_last_error().decode(errors="replace")
        kind = RM_ERRORS.get(rc, f"error {rc}")
        if rc == 2:
            raise AssertionError(f"{what}: {kind}: {msg}")
        raise RuntimeError(f"{what}: {kind}: {msg}")


def ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def stream_of(device=None):
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _host_floats(x, n):
    if isinstance(x, torch.Tensor):
        x = x.detach().reshape(-1)
    return [float(v) for v in list(x)[:n]]


def camera_from_tcam(tcam: dict) -> RmCamera:
    """Packs the reference's tcam dict (utils/train_util.py:31-38 + data/camera.py:72-91).
    The small camera tensors are brought to the host with one copy."""
    parts = []
    for key, n in (("world_view_transform", 16), ("K", 9), ("cam_pos", 3), ("distortion_params", 4)):
        v = tcam.get(key)
        if v is None:
            if key == "distortion_params":
                v = torch.zeros(4)
            else:
                raise KeyError(f"tcam is missing '{key}'")
        if not isinstance(v, torch.Tensor):
            v = torch.as_tensor(v)
        parts.append(v.detach().reshape(-1).float().to("cpu", non_blocking=False)[:n])
    flat = torch.cat(parts).tolist()
    c = RmCamera()
    c.w2v[:] = flat[0:16]
    c.K[:] = flat[16:25]
    c.cam_pos[:] = flat[25:28]
    c.dist[:] = flat[28:32]
    c.min_t = float(tcam.get("min_t", 0.1))
    c.fovy = float(tcam.get("fovy", 0.0))
    c.fovx = float(tcam.get("fovx", 0.0))
    c.image_height = int(tcam["image_height"])
    c.image_width = int(tcam["image_width"])
    c.tile_height = int(tcam["tile_height"])
    c.tile_width = int(tcam["tile_width"])
    c.grid_height = int(tcam.get("grid_height", (c.image_height + c.tile_height - 1) // c.tile_height))
    c.grid_width = int(tcam.get("grid_width", (c.image_width + c.tile_width - 1) // c.tile_width))
    c.camera_type = int(tcam.get("camera_type", 0))
    return c

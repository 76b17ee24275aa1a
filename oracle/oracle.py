"""ctypes wrapper of the CPU oracle (librmesh_oracle.so) plus the numpy field oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline leg. The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librmesh_oracle.so")


class OrcCamera(C.Structure):
    _fields_ = [("w2v", C.c_float * 16), ("K", C.c_float * 9), ("cam_pos", C.c_float * 3),
                ("dist", C.c_float * 4), ("min_t", C.c_float), ("fovy", C.c_float),
                ("fovx", C.c_float), ("H", C.c_int32), ("W", C.c_int32), ("gh", C.c_int32),
                ("gw", C.c_int32), ("th", C.c_int32), ("tw", C.c_int32),
                ("camera_type", C.c_int32), ("pad_", C.c_int32)]


def make_camera(cam, tile_size: int, min_t: float) -> OrcCamera:
    c = OrcCamera()
    c.w2v[:] = [float(v) for v in np.asarray(cam.w2v, np.float32).reshape(-1)]
    c.K[:] = [float(v) for v in np.asarray(cam.K, np.float32).reshape(-1)]
    c.cam_pos[:] = [float(v) for v in np.asarray(cam.cam_pos, np.float32).reshape(-1)]
    c.dist[:] = [float(v) for v in np.asarray(cam.distortion, np.float32).reshape(-1)]
    c.min_t = min_t
    c.fovy = cam.fovy
    c.fovx = cam.fovx
    c.H, c.W = cam.height, cam.width
    c.th = c.tw = tile_size
    c.gh = (cam.height + tile_size - 1) // tile_size
    c.gw = (cam.width + tile_size - 1) // tile_size
    c.camera_type = cam.camera_type
    return c


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_project.argtypes = [C.POINTER(OrcCamera), P, P, C.c_int64, P, P, P, P]
        L.orc_bin.argtypes = [P, P, P, C.c_int64, C.c_int32, C.c_int32, P, P, P]
        L.orc_bin.restype = C.c_int64
        L.orc_sortable.argtypes = [C.c_float]
        L.orc_sortable.restype = C.c_uint32
        L.orc_composite_fwd.argtypes = [C.POINTER(OrcCamera), P, P, P, P, P, C.c_int, P, P, P, P, P]
        L.orc_composite_bwd.argtypes = [C.POINTER(OrcCamera), P, P, P, P, P, C.c_int, P, P, P, P,
                                        P, P, P, P]
        L.orc_composite_fwd_tiles.argtypes = [C.POINTER(OrcCamera), P, P, P, P, P, C.c_int, P, P, C.c_int64,
                                              P, P, P, P]
        L.orc_composite_bwd_tiles.argtypes = [C.POINTER(OrcCamera), P, P, P, P, P, C.c_int, P, P, C.c_int64,
                                              P, P, P, P, P, P, P]
        L.orc_quadrature.argtypes = [C.POINTER(OrcCamera), P, P, C.c_int64, C.c_double, C.c_int,
                                     P, P, P, C.c_int, C.c_int64, P]
        L.orc_get_ray.argtypes = [C.POINTER(OrcCamera), C.c_double, C.c_double, P, P]
        L.orc_intersect.argtypes = [P, P, P, P, P]
        L.orc_intersect.restype = C.c_int
        L.orc_power_of_circumsphere.argtypes = [P, P, P, P, P, P]
        L.orc_power_of_circumsphere.restype = C.c_double
        L.orc_power_bwd.argtypes = [P, P, C.c_double, P, P]
        L.orc_tet_err.argtypes = [C.POINTER(OrcCamera), P, P, P, P, P, C.c_int, P, P, P, P, P, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def project(cam: OrcCamera, indices, vertices):
    idx = np.ascontiguousarray(indices, np.int32)
    v = np.ascontiguousarray(vertices, np.float32)
    T = idx.shape[0]
    tiles = np.zeros(T, np.int32)
    rect = np.zeros((T, 4), np.int32)
    vs = np.zeros((T, 3), np.float32)
    cc = np.zeros((T, 3), np.float32)
    lib().orc_project(C.byref(cam), _p(idx), _p(v), T, _p(tiles), _p(rect), _p(vs), _p(cc))
    return tiles, rect, vs, cc


def bin_tiles(cam: OrcCamera, tiles, rect, vs):
    P = int(tiles.astype(np.int64).sum())
    keys = np.zeros(max(P, 1), np.int64)
    vals = np.zeros(max(P, 1), np.int32)
    ranges = np.zeros((cam.gh * cam.gw, 2), np.int32)
    lib().orc_bin(_p(np.ascontiguousarray(tiles)), _p(np.ascontiguousarray(rect)),
                  _p(np.ascontiguousarray(vs)), tiles.shape[0], cam.gh, cam.gw, _p(keys),
                  _p(vals), _p(ranges))
    return keys[:P], vals[:P], ranges


def composite_fwd(cam: OrcCamera, sorted_idx, ranges, indices, vertices, feats, jitter, tiles=None):
    """FP64 forward (img [H,W,4], dist [H,W,5], n_contrib [H,W], margin [H,W]); with `tiles` (tile ids)
    only those tiles' pixels are computed and every other pixel stays zero (margin 0)."""
    H, W = cam.H, cam.W
    F = feats.shape[1]
    img = np.zeros((H, W, 4))
    dist = np.zeros((H, W, 5))
    nc = np.zeros((H, W), np.int32)
    margin = np.zeros((H, W))
    v = np.ascontiguousarray(vertices, np.float64)
    f = np.ascontiguousarray(feats, np.float64)
    j = np.ascontiguousarray(jitter, np.float32)
    args = [_p(np.ascontiguousarray(sorted_idx, np.int32)), _p(np.ascontiguousarray(ranges, np.int32)),
            _p(np.ascontiguousarray(indices, np.int32)), _p(v), _p(f), F, _p(j)]
    if tiles is None:
        lib().orc_composite_fwd(C.byref(cam), *args, _p(img), _p(dist), _p(nc), _p(margin))
    else:
        tl = np.ascontiguousarray(tiles, np.int32)
        lib().orc_composite_fwd_tiles(C.byref(cam), *args, _p(tl), len(tl), _p(img), _p(dist), _p(nc), _p(margin))
    return img, dist, nc, margin


def tet_err(cam: OrcCamera, sorted_idx, ranges, indices, vertices, feats, pixel_err, ssim_err, graze=False):
    """FP64 per-tet error votes of calc_tet_err: (err [T,16], count [T,2] int32, margin [H,W]);
    with graze=True also grazes [T] int32, the pixels hitting each tet with a chord below 1e-12."""
    T = indices.shape[0]
    F = feats.shape[1]
    err = np.zeros((T, 16))
    count = np.zeros((T, 2), np.int32)
    margin = np.zeros((cam.H, cam.W))
    gz = np.zeros((T,), np.int32)
    lib().orc_tet_err(C.byref(cam), _p(np.ascontiguousarray(sorted_idx, np.int32)),
                      _p(np.ascontiguousarray(ranges, np.int32)), _p(np.ascontiguousarray(indices, np.int32)),
                      _p(np.ascontiguousarray(vertices, np.float64)), _p(np.ascontiguousarray(feats, np.float64)),
                      F, _p(np.ascontiguousarray(pixel_err, np.float64)),
                      _p(np.ascontiguousarray(ssim_err, np.float64)), _p(err), _p(count), _p(margin), _p(gz))
    return (err, count, margin, gz) if graze else (err, count, margin)


def composite_bwd(cam: OrcCamera, sorted_idx, ranges, indices, vertices, feats, jitter, img, dist,
                  nc, gimg, gdist, tiles=None):
    """FP64 adjoint (grad_vertices [V,3], grad_features [T,F]); with `tiles` only those tiles' pixels
    are walked (the other pixels' incoming gradients are ignored)."""
    F = feats.shape[1]
    v = np.ascontiguousarray(vertices, np.float64)
    f = np.ascontiguousarray(feats, np.float64)
    gv = np.zeros_like(v)
    gf = np.zeros_like(f)
    if tiles is not None:
        tl = np.ascontiguousarray(tiles, np.int32)
        lib().orc_composite_bwd_tiles(C.byref(cam), _p(np.ascontiguousarray(sorted_idx, np.int32)),
                                      _p(np.ascontiguousarray(ranges, np.int32)),
                                      _p(np.ascontiguousarray(indices, np.int32)), _p(v), _p(f), F,
                                      _p(np.ascontiguousarray(jitter, np.float32)), _p(tl), len(tl),
                                      _p(np.ascontiguousarray(img, np.float64)),
                                      _p(np.ascontiguousarray(dist, np.float64)),
                                      _p(np.ascontiguousarray(nc, np.int32)),
                                      _p(np.ascontiguousarray(gimg, np.float64)),
                                      _p(np.ascontiguousarray(gdist, np.float64)), _p(gv), _p(gf))
        return gv, gf
    lib().orc_composite_bwd(C.byref(cam), _p(np.ascontiguousarray(sorted_idx, np.int32)),
                            _p(np.ascontiguousarray(ranges, np.int32)),
                            _p(np.ascontiguousarray(indices, np.int32)), _p(v), _p(f), F,
                            _p(np.ascontiguousarray(jitter, np.float32)),
                            _p(np.ascontiguousarray(img, np.float64)),
                            _p(np.ascontiguousarray(dist, np.float64)),
                            _p(np.ascontiguousarray(nc, np.int32)),
                            _p(np.ascontiguousarray(gimg, np.float64)),
                            _p(np.ascontiguousarray(gdist, np.float64)), _p(gv), _p(gf))
    return gv, gf


def quadrature(cam: OrcCamera, jitter, pix, t_far, n_samples, indices, vertices, feats):
    pix = np.ascontiguousarray(pix, np.int64)
    out = np.zeros((len(pix), 4))
    lib().orc_quadrature(C.byref(cam), _p(np.ascontiguousarray(jitter, np.float32)), _p(pix),
                         len(pix), float(t_far), int(n_samples),
                         _p(np.ascontiguousarray(indices, np.int32)),
                         _p(np.ascontiguousarray(vertices, np.float64)),
                         _p(np.ascontiguousarray(feats, np.float64)), feats.shape[1],
                         indices.shape[0], _p(out))
    return out


def get_ray(cam: OrcCamera, px, py):
    o = np.zeros(3)
    d = np.zeros(3)
    lib().orc_get_ray(C.byref(cam), float(px), float(py), _p(o), _p(d))
    return o, d


def intersect(v12, o, d):
    v = np.ascontiguousarray(v12, np.float64).reshape(12)
    e = np.zeros(1)
    x = np.zeros(1)
    hit = lib().orc_intersect(_p(v), _p(np.ascontiguousarray(o, np.float64)),
                              _p(np.ascontiguousarray(d, np.float64)), _p(e), _p(x))
    return bool(hit), float(e[0]), float(x[0])


def power_of_circumsphere(A, B, C_, D, P):
    ctr = np.zeros(3)
    arrs = [np.ascontiguousarray(a, np.float64) for a in (A, B, C_, D, P)]
    pw = lib().orc_power_of_circumsphere(*[_p(a) for a in arrs], _p(ctr))
    return pw, ctr


def power_bwd(V12, P, g_power, g_center):
    g = np.zeros(12)
    lib().orc_power_bwd(_p(np.ascontiguousarray(V12, np.float64).reshape(12)),
                        _p(np.ascontiguousarray(P, np.float64)), float(g_power),
                        _p(np.ascontiguousarray(g_center, np.float64)), _p(g))
    return g.reshape(4, 3)


def project_bwd(cam: OrcCamera, indices, vertices, tiles, grad_vs, grad_cc):
    """FP64 adjoint of the projection outputs (vs[:,2] and circumcenter) w.r.t. vertices."""
    v = np.asarray(vertices, np.float64)
    gv = np.zeros_like(v)
    P = np.array(cam.cam_pos[:], np.float64)
    for t in np.nonzero(tiles > 0)[0]:
        q = indices[t]
        g = power_bwd(v[q].reshape(12), P, float(grad_vs[t, 2]), np.asarray(grad_cc[t], np.float64))
        np.add.at(gv, q, g)
    return gv


# --------------------------------------------------------------------------- field oracle

PRIMES = (1, 2654435761, 805459861)


def hashgrid_levels(n_levels, base_res, growth, log2_size):
    """Per-level resolution floor(base * growth**l) and table sizes min(2^log2, res^3)
    (utils/hashgrid.py:35-48)."""
    res = [int(np.floor(base_res * growth ** l)) for l in range(n_levels)]
    sizes = [min(2 ** log2_size, r ** 3) for r in res]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    return np.array(res, np.int32), offs


def _grid_cells(x, r):
    """Cell corner and fractions at resolution r, with the reference's FP32 position arithmetic
    (utils/hashgrid.py:78-84: clamp, x * r, + 0.5, floor, frac = pos - floor, all in float32)."""
    xf = np.clip(np.asarray(x, np.float32), np.float32(0), np.float32(1))
    pos = (xf * np.float32(r)).astype(np.float32) + np.float32(0.5)
    fl = np.floor(pos)
    return fl.astype(np.int64), (pos - fl).astype(np.float64)


def _corner_hash(vi, r, log2_size):
    if r ** 3 <= (1 << log2_size):
        w = vi % r
        return w[:, 0] + w[:, 1] * r + w[:, 2] * r * r
    u = vi.astype(np.uint64)
    h = (u[:, 0] * np.uint64(PRIMES[0])) ^ (u[:, 1] * np.uint64(PRIMES[1])) ^ (u[:, 2] * np.uint64(PRIMES[2]))
    return (h & np.uint64((1 << log2_size) - 1)).astype(np.int64)


def hashgrid_fwd(x, table, res, offs, n_feat, log2_size):
    """Restatement of HashEmbedderOptimized.forward (utils/hashgrid.py:74-124) on the [0,1]^3 box:
    FP32 cell position as the reference, then 8 corners with the dense modulo index or the XOR-prime
    hash, trilinear interpolation in FP64. Returns [N, L*n_feat] level-major."""
    N = np.asarray(x).shape[0]
    L = len(res)
    out = np.zeros((N, L * n_feat))
    for l in range(L):
        r = int(res[l])
        fl, fr = _grid_cells(x, r)
        acc = np.zeros((N, n_feat))
        for c in range(8):
            off = np.array([(c >> 2) & 1, (c >> 1) & 1, c & 1])
            h = _corner_hash(fl + off, r, log2_size)
            wt = np.prod(np.where(off[None, :] == 1, fr, 1.0 - fr), axis=1)
            acc += wt[:, None] * np.asarray(table, np.float64)[offs[l] + h]
        out[:, l * n_feat:(l + 1) * n_feat] = acc
    return out


def hashgrid_bwd(x, grad_out, res, offs, n_feat, log2_size, table_rows):
    """Adjoint of hashgrid_fwd w.r.t. the table (the encode is linear in the table): scatter-add of
    corner weight * output gradient, in FP64."""
    g = np.zeros((table_rows, n_feat))
    go = np.asarray(grad_out, np.float64)
    for l in range(len(res)):
        r = int(res[l])
        fl, fr = _grid_cells(x, r)
        for c in range(8):
            off = np.array([(c >> 2) & 1, (c >> 1) & 1, c & 1])
            h = _corner_hash(fl + off, r, log2_size)
            wt = np.prod(np.where(off[None, :] == 1, fr, 1.0 - fr), axis=1)
            np.add.at(g, offs[l] + h, wt[:, None] * go[:, l * n_feat:(l + 1) * n_feat])
    return g


SELU_ALPHA = 1.6732632423543772
SELU_SCALE = 1.0507009873554805


def selu(x):
    return SELU_SCALE * np.where(x > 0, x, SELU_ALPHA * np.expm1(np.minimum(x, 0)))


def mlp_fwd(x, w1, b1, w2, b2, w3, b3):
    """FP64 Linear->SELU->Linear->SELU->Linear (models/ingp_color.py:141-151)."""
    h1 = x @ w1.T + b1
    h2 = selu(h1) @ w2.T + b2
    return selu(h2) @ w3.T + b3


# --------------------------------------------------------------------- the reference model's field

def circumspheres(tets):
    """FP64 circumcenters [T,3] and radii [T] of tets [T,4,3], by solving
    2 (v_i - v_0) . c' = |v_i - v_0|^2 (i = 1..3) for c' = c - v_0 (not the reference's
    cross-product form, utils/topo_utils.py:120-164, which is the same point)."""
    t = np.asarray(tets, np.float64)
    E = t[:, 1:] - t[:, :1]
    rhs = 0.5 * (E * E).sum(-1)
    rel = np.linalg.solve(E, rhs[..., None])[..., 0]
    return t[:, 0] + rel, np.linalg.norm(rel, axis=1)


def contract_mean_std(x, std, eps=1.1920928955078125e-07):
    """FP64 contract_mean_std (utils/contraction.py:6-20): z = x inside the unit ball, else
    (2 - 1/|x|) x/|x|; std scaled by det_13 = ((2|x| - 1)^(1/3) / |x|)^2 (the cube root of the
    contraction's Jacobian determinant) outside; 2|x| - 1 clamped to [1e-6, 1e6] as safe_pow does."""
    x = np.asarray(x, np.float64)
    msq = np.maximum((x * x).sum(-1, keepdims=True), eps)
    mag = np.sqrt(msq)
    inside = msq <= 1
    z = np.where(inside, x, ((2 * mag - 1) / msq) * x)
    det13 = (np.cbrt(np.clip(2 * mag - 1, 1e-6, 1e6)) / mag) ** 2
    return z, np.where(inside[..., 0], std, det13[..., 0] * std)


def field_query(vertices, indices, center, scene_scaling, anchor="centroid"):
    """FP64 query of Model.compute_batch_features (models/ingp_color.py:314-349, k_samples 1,
    alpha 0): anchor (centroid or circumcenter), radius = circumradius, (x, cr) with
    x = (cv/2 + 1)/2 of the contracted normalised anchor. Returns (x [T,3], cr [T], anchor [T,3])."""
    tets = np.asarray(vertices, np.float64)[np.asarray(indices, np.int64)]
    cc, radius = circumspheres(tets)
    anc = tets.mean(1) if anchor == "centroid" else cc
    c = np.asarray(center, np.float64).reshape(1, 3)
    z, cr = contract_mean_std((anc - c) / scene_scaling, radius / scene_scaling)
    return (z / 2 + 1) / 2, cr, anc


def approx_erf(x):
    """sign(x) sqrt(1 - exp(-(4/pi) x^2)) (models/ingp_color.py:84-86)."""
    return np.sign(x) * np.sqrt(1 - np.exp(-(4 / np.pi) * x * x))


def downweight_scale(cr, n_levels, scale_multi, per_level_scale):
    """[N, L] scales of iNGPDW._encode (models/ingp_color.py:174-183): approx_erf(safe_div(1,
    safe_sqrt(per_level_scale * 4 * n * cr * scale_multi))) for n = 0..L-1; safe_sqrt clamps its
    input to at least float32 tiny, so n = 0 gives 1/sqrt(tiny) and a scale of exactly 1."""
    tiny = float(np.finfo(np.float32).tiny)
    n = np.arange(n_levels, dtype=np.float64)[None, :]
    d = np.maximum(per_level_scale * 4 * n * (np.asarray(cr, np.float64).reshape(-1, 1) * scale_multi), tiny)
    return approx_erf(1.0 / np.sqrt(d))


def downweight(enc, cr, n_levels, n_feat, scale_multi, per_level_scale):
    """enc [N, L*F] level-major, viewed as [N, F, L] (the reference's reshape(-1, dim, L)) and
    scaled along the last axis."""
    e = np.asarray(enc, np.float64).reshape(-1, n_feat, n_levels)
    s = downweight_scale(cr, n_levels, scale_multi, per_level_scale)
    return (e * s[:, None, :]).reshape(-1, n_levels * n_feat)


def field_outputs(sigma, rgb, grd, sh, density_offset):
    """iNGPDW.forward's activations (models/ingp_color.py:196-205), FP64: exp(sigma + offset)
    (input clamped above at log(float32 max)), rgb + 0.5, grd / sqrt(|grd|^2 + 1); sh unrounded."""
    s = np.minimum(np.asarray(sigma, np.float64) + density_offset, np.log(np.finfo(np.float32).max))
    g = np.asarray(grd, np.float64)
    return np.exp(s), np.asarray(rgb, np.float64) + 0.5, g / np.sqrt((g * g).sum(-1, keepdims=True) + 1), \
        np.asarray(sh, np.float64)


def ingp_forward(x, cr, table, res, offs, n_feat, log2_size, heads, scale_multi, per_level_scale,
                 density_offset, ablate_downweighing=False):
    """FP64 iNGPDW.forward: hash grid -> down-weighting -> the four heads (density, colour,
    gradient, SH; each a list [w1, b1, w2, b2, w3, b3]) -> activations."""
    enc = hashgrid_fwd(x, table, res, offs, n_feat, log2_size)
    if not ablate_downweighing:
        enc = downweight(enc, cr, len(res), n_feat, scale_multi, per_level_scale)
    raw = [mlp_fwd(enc, *[np.asarray(a, np.float64) for a in h]) for h in heads]
    return field_outputs(*raw, density_offset)


# --------------------------------------------------------------------- per-tet feature activation

def _sh_real_scipy(dirs, deg):
    """Real SH basis for degrees 1..deg at unit directions [N,3], from scipy's complex harmonics
    (Condon-Shortley phase): sqrt(2) Im (m < 0), Re (m = 0), sqrt(2) Re (m > 0)."""
    from scipy import special
    polar = np.arccos(np.clip(dirs[:, 2], -1, 1))
    azim = np.arctan2(dirs[:, 1], dirs[:, 0])
    cols = []
    for l in range(1, deg + 1):
        for m in range(-l, l + 1):
            Y = special.sph_harm_y(l, abs(m), polar, azim)
            cols.append(np.sqrt(2) * Y.imag if m < 0 else (Y.real if m == 0 else np.sqrt(2) * Y.real))
    return np.stack(cols, 1) if cols else np.zeros((len(dirs), 0))


def activate_output(camera_center, density, rgb, grd, sh, indices, circumcenters, vertices, deg, offset=None):
    """FP64 per-tet features [T,7] of the reference's activate_output (utils/model_util.py:93-105,
    offset_normalize :81-91, compute_vertex_colors_from_field :62-78, eval_sh
    utils/eval_sh_py.py:137-195, softplus beta 10 threshold 20; safe_div utils/safe_math.py:143-171).
    offset [T,3], when given, replaces v0 - c in the vertex-0 colour: the reference detaches the
    vertices and circumcenters there, so central differences of its gradient hold it fixed."""
    V = np.asarray(vertices, np.float64)
    tets = V[np.asarray(indices, np.int64)]                       # [T,4,3]
    rgb = np.asarray(rgb, np.float64)
    grd = np.asarray(grd, np.float64).reshape(-1, 3)
    cc = np.asarray(circumcenters, np.float64)
    g2 = grd * rgb.min(axis=1, keepdims=True)                     # grd * min channel
    radius = np.linalg.norm(tets[:, 0] - cc, axis=1, keepdims=True)
    tiny = np.finfo(np.float32).tiny
    ng = np.where(np.abs(radius) < tiny, 0.0, g2 / np.where(np.abs(radius) < tiny, tiny, radius))
    off = tets[:, 0] - cc if offset is None else np.asarray(offset, np.float64)
    base = rgb + (ng * off).sum(1, keepdims=True)                 # vertex-0 colour of the linear field
    md = tets.mean(1) - np.asarray(camera_center, np.float64).reshape(1, 3)
    dirs = md / np.maximum(np.linalg.norm(md, axis=1, keepdims=True), 1e-8)
    raw = 0.28209479177387814 * ((base - 0.5) / 0.28209479177387814) + 0.5
    if deg > 0:
        B = _sh_real_scipy(dirs, deg)
        raw = raw + np.einsum("tk,tkc->tc", B, np.asarray(sh, np.float64)[:, :B.shape[1]])
    col = np.where(10 * raw > 20, raw, np.log1p(np.exp(np.minimum(10 * raw, 20))) / 10)
    return np.concatenate([np.asarray(density, np.float64).reshape(-1, 1), col, ng], 1)


# ---- Delaunay tetrahedralisation (models/ingp_color.py:523-544) ---------------------------------
def delaunay(points):
    """The reference's high-precision triangulation branch: scipy.spatial.Delaunay (Qhull) simplices
    of the points (models/ingp_color.py:526-528), then the orientation fix of :536-540 (tets with
    topo_utils.tet_volumes < 0, utils/topo_utils.py:9-24, get columns 0 and 1 swapped). int64 [T,4]."""
    from scipy.spatial import Delaunay
    p = np.asarray(points, np.float64)
    t = Delaunay(p).simplices.astype(np.int64)
    a, b, c, d = (p[t[:, k]] for k in range(4))
    neg = np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a) < 0
    t[neg] = t[neg][:, [1, 0, 2, 3]]
    return t


def canonical_tets(t):
    """Tets as a set: vertex ids sorted within each tet, rows sorted lexicographically
    (the normalisation of tests/delaunay_update_test.py:12-13 and tests/test_del.py:45-46)."""
    s = np.sort(np.asarray(t, np.int64).reshape(-1, 4), axis=1)
    return s[np.lexsort(s.T[::-1])] if len(s) else s


def _fracs(points):
    from fractions import Fraction
    return [[Fraction(float(v)) for v in row] for row in np.asarray(points, np.float32)]


def _det(m):
    """Exact determinant (Fraction entries) by fraction-free elimination."""
    m = [list(r) for r in m]
    n, sign, det = len(m), 1, 1
    for c in range(n):
        piv = next((r for r in range(c, n) if m[r][c] != 0), None)
        if piv is None:
            return 0
        if piv != c:
            m[c], m[piv] = m[piv], m[c]
            sign = -sign
        det *= m[c][c]
        for r in range(c + 1, n):
            f = m[r][c] / m[c][c]
            if f:
                for k in range(c, n):
                    m[r][k] -= f * m[c][k]
    return sign * det


def orient_exact(P, i, j, k, l):
    """Sign of det(P_j - P_i, P_k - P_i, P_l - P_i) in exact arithmetic (P from _fracs)."""
    a, b, c, d = P[i], P[j], P[k], P[l]
    d3 = _det([[b[x] - a[x] for x in range(3)], [c[x] - a[x] for x in range(3)], [d[x] - a[x] for x in range(3)]])
    return (d3 > 0) - (d3 < 0)


def inside_sphere_exact(P, tet, q):
    """True when point q is strictly inside the circumsphere of tet (exact, no perturbation):
    |q - c|^2 < |a - c|^2 with the circumcentre c solved exactly."""
    a = P[tet[0]]
    rows, rhs = [], []
    for v in tet[1:]:
        b = P[v]
        rows.append([b[x] - a[x] for x in range(3)])
        rhs.append(sum(b[x] * b[x] - a[x] * a[x] for x in range(3)) / 2)
    D = _det(rows)
    c = []
    for x in range(3):
        m = [list(r) for r in rows]
        for r in range(3):
            m[r][x] = rhs[r]
        c.append(_det(m) / D)
    r2 = sum((a[x] - c[x]) ** 2 for x in range(3))
    return sum((P[q][x] - c[x]) ** 2 for x in range(3)) < r2


def delaunay_brute(points):
    """Every 4-subset of non-coplanar points whose circumsphere holds no other point inside or on
    it, positively oriented; exact. For points in general position this is the Delaunay
    tetrahedralisation itself (small inputs only: O(N^5))."""
    import itertools
    P = _fracs(points)
    n = len(P)
    out = []
    for t in itertools.combinations(range(n), 4):
        o = orient_exact(P, *t)
        if o == 0:
            continue
        tt = t if o > 0 else (t[1], t[0], t[2], t[3])
        a = P[tt[0]]
        rows, rhs = [], []
        for v in tt[1:]:
            b = P[v]
            rows.append([b[x] - a[x] for x in range(3)])
            rhs.append(sum(b[x] * b[x] - a[x] * a[x] for x in range(3)) / 2)
        D = _det(rows)
        c = []
        for x in range(3):
            m = [list(r) for r in rows]
            for r in range(3):
                m[r][x] = rhs[r]
            c.append(_det(m) / D)
        r2 = sum((a[x] - c[x]) ** 2 for x in range(3))
        if all(sum((P[q][x] - c[x]) ** 2 for x in range(3)) > r2 for q in range(n) if q not in t):
            out.append(tt)
    return np.array(out, np.int64).reshape(-1, 4)


def perturbed_inside(points5, ids5):
    """Whether point 4 is inside the circumsphere of the positively oriented tet of points 0-3
    under the lift perturbation lift_k -= eps^id_k (eps -> 0+), from the 5x5 lifted determinant and
    its lift-column cofactors, each computed as an exact determinant (rows in the given order)."""
    P = _fracs(points5)
    M = [[P[r][0], P[r][1], P[r][2], sum(v * v for v in P[r]), 1] for r in range(5)]
    D = _det(M)
    if D != 0:
        s = (D > 0) - (D < 0)
    else:
        s = 0
        for r in sorted(range(5), key=lambda k: ids5[k]):
            Mc = [list(row) for row in M]
            for k in range(5):
                Mc[k][3] = 1 if k == r else 0
            C = _det(Mc)
            if C != 0:
                s = -((C > 0) - (C < 0))   # d/d(lift_r) = C_r and lift_r decreases
                break
    return s < 0

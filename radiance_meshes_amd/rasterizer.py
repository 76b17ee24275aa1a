"""Host side of the tetrahedral tile rasterizer, with the reference's Python interface:

* ``RenderGrid``                     delaunay_rasterization/internal/render_grid.py:20-27
* ``vertex_and_tile_shader``         delaunay_rasterization/internal/tile_shader_slang.py:24-72
* ``VertexShader``                   delaunay_rasterization/internal/tile_shader_slang.py:75-150
* ``AlphaBlendTiledRender``          delaunay_rasterization/internal/alphablend_tiled_slang_interp.py:9-112
* ``sort_by_keys``                   delaunay_rasterization/internal/sort_by_keys/sort_by_keys.cu:9-53
* ``split_tensors/recombine_tensors`` delaunay_rasterization/internal/util.py:3-23

Every op runs the HIP kernels of librmesh.so on torch's current stream. Names, argument order,
return tuples and error behaviour (AssertionError for an unsupported tile size or NaN vertices,
RuntimeError for dtype/shape misuse) follow the reference.
"""
from __future__ import annotations

import math

import torch

from . import _capi, timing
from ._capi import check, lib, ptr, stream_of

SUPPORTED_TILES = (4, 8, 16)


def ceil_div(x, y):
    return (x + y - 1) // y


class RenderGrid:
    def __init__(self, image_height, image_width, tile_width, tile_height):
        self.image_height = image_height
        self.image_width = image_width
        self.tile_height = tile_height
        self.tile_width = tile_width
        self.grid_height = ceil_div(image_height, tile_height)
        self.grid_width = ceil_div(image_width, tile_width)


def recombine_tensors(non_tensor_data, tensor_data):
    tensor_keys = non_tensor_data["_tensor_keys"]
    out = {k: v for k, v in non_tensor_data.items() if k != "_tensor_keys"}
    out.update({key: t for key, t in zip(tensor_keys, tensor_data)})
    return out


def split_tensors(tcam):
    tensors, tensor_keys, non_tensor = [], [], {}
    for key, value in tcam.items():
        if isinstance(value, torch.Tensor):
            tensors.append(value)
            tensor_keys.append(key)
        else:
            non_tensor[key] = value
    non_tensor["_tensor_keys"] = tensor_keys
    return non_tensor, tensors


def _require(t, dtype, name, cols=None):
    if not isinstance(t, torch.Tensor):
        raise RuntimeError(f"{name} must be a tensor")
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a device tensor (got {t.device})")
    if t.dtype != dtype:
        raise RuntimeError(f"{name} must be {dtype}, got {t.dtype}")
    if cols is not None and (t.dim() != 2 or t.shape[1] not in (cols if isinstance(cols, tuple) else (cols,))):
        raise RuntimeError(f"{name} must have shape [N, {cols}], got {tuple(t.shape)}")
    return t.contiguous()


def _cam(tcam, render_grid=None):
    c = _capi.camera_from_tcam(tcam)
    if render_grid is not None:
        c.tile_height, c.tile_width = render_grid.tile_height, render_grid.tile_width
        c.grid_height, c.grid_width = render_grid.grid_height, render_grid.grid_width
        c.image_height, c.image_width = render_grid.image_height, render_grid.image_width
    return c


class VertexShader(torch.autograd.Function):
    """Per-tet projection + tile rectangles (forward) and the FP64 adjoint of the circumsphere
    power / circumcenter w.r.t. the vertices (backward)."""

    @staticmethod
    def forward(ctx, indices, vertices, tcam, render_grid, device="cuda"):
        assert not torch.isnan(vertices).any(), "Tensor contains NaN values!"
        indices = _require(indices, torch.int32, "indices", 4)
        vertices = _require(vertices, torch.float32, "vertices", 3)
        T = indices.shape[0]
        dev = vertices.device
        tiles = torch.empty((T,), device=dev, dtype=torch.int32)
        rect = torch.empty((T, 4), device=dev, dtype=torch.int32)
        vs = torch.empty((T, 3), device=dev, dtype=torch.float32)
        cc = torch.empty((T, 3), device=dev, dtype=torch.float32)
        cam = _cam(tcam, render_grid)
        with timing.span("project_fwd"):
            check(lib().rm_project_fwd(cam, ptr(indices), ptr(vertices), T, vertices.shape[0], ptr(tiles),
                                       ptr(rect), ptr(vs), ptr(cc), stream_of(dev)), "rm_project_fwd")
        non_tensor, tensor_data = split_tensors(tcam)
        ctx.non_tensor_data = non_tensor
        ctx.render_grid = render_grid
        ctx.cam = cam
        ctx.save_for_backward(indices, vertices, tiles, *tensor_data)
        ctx.mark_non_differentiable(tiles, rect)
        return tiles, rect, vs, cc

    @staticmethod
    def backward(ctx, grad_tiles, grad_rect, grad_vs, grad_cc):
        indices, vertices, tiles = ctx.saved_tensors[:3]
        grad_vertices = torch.zeros_like(vertices)
        if grad_vs is None and grad_cc is None:
            return None, grad_vertices, None, None
        gvs = grad_vs.contiguous() if grad_vs is not None else None
        gcc = grad_cc.contiguous() if grad_cc is not None else None
        with timing.span("project_bwd"):
            check(lib().rm_project_bwd(ctx.cam, ptr(indices), ptr(vertices), indices.shape[0], vertices.shape[0],
                                   ptr(tiles), ptr(gvs), ptr(gcc), ptr(grad_vertices),
                                   stream_of(vertices.device)), "rm_project_bwd")
        return None, grad_vertices, None, None


# Binning workspace kept between calls, one per (device, stream): the count phase leaves its results
# at the start of the workspace and the sort phase needs room for the pairs, so a workspace that is
# already large enough saves allocating a larger one and copying those results into it. Calls on one
# stream run in order, so reusing the buffer there is safe.
_bin_ws = {}


def bin_and_sort(tiles_touched, rect_tile_space, vs_tetra, render_grid):
    """Fused binning: (tile, depth)-ordered tet list and per-tile [start, end) ranges, bit-identical
    to the reference's keys + stable CUB sort + compute_tile_ranges (tile_shader_slang.py:36-68)."""
    dev = tiles_touched.device
    T = tiles_touched.shape[0]
    G = render_grid.grid_height * render_grid.grid_width
    L = lib()
    st = stream_of(dev)
    ws_bytes = L.rm_bin_workspace_size(T, 0, G)
    key = (str(dev), torch.cuda.current_stream(dev).cuda_stream)
    ws = _bin_ws.get(key)
    if ws is None or ws.numel() < ws_bytes:
        ws = torch.empty((max(ws_bytes, 1),), device=dev, dtype=torch.uint8)
    counts = torch.empty((2,), device=dev, dtype=torch.int64)
    with timing.span("bin_count"):
        check(L.rm_bin_count(ptr(tiles_touched), T, ptr(counts), ptr(ws), ws_bytes, st), "rm_bin_count")
    n_vis, P = (int(v) for v in counts.tolist())  # host sync, as the reference's index_buffer_offset[-1]
    full = L.rm_bin_workspace_size(T, P, G)
    if full > ws.numel():
        # room for a quarter more pairs, so that small changes of the view reuse it
        ws2 = torch.empty((max(full, L.rm_bin_workspace_size(T, P + P // 4, G)),), device=dev, dtype=torch.uint8)
        ws2[:ws_bytes].copy_(ws[:ws_bytes])
        ws = ws2
    _bin_ws[key] = ws
    sorted_idx = torch.empty((P,), device=dev, dtype=torch.int32)
    ranges = torch.empty((G, 2), device=dev, dtype=torch.int32)
    with timing.span("bin_sort"):
        check(L.rm_bin_sort(ptr(tiles_touched), ptr(rect_tile_space), ptr(vs_tetra), T, n_vis, P,
                        render_grid.grid_height, render_grid.grid_width, ptr(sorted_idx), ptr(ranges),
                        ptr(ws), ws.numel(), st), "rm_bin_sort")
    return sorted_idx, ranges


def inclusive_scan(x):
    x = _require(x, torch.int32, "x")
    out = torch.empty_like(x)
    n = x.numel()
    wsb = lib().rm_scan_workspace_size(n)
    ws = torch.empty((max(wsb, 1),), device=x.device, dtype=torch.uint8)
    check(lib().rm_inclusive_scan_i32(ptr(x), ptr(out), n, ptr(ws), wsb, stream_of(x.device)),
          "rm_inclusive_scan_i32")
    return out


def generate_keys(vs_tetra, rect_tile_space, index_buffer_offset, grid_height, grid_width):
    P = int(index_buffer_offset[-1].item()) if index_buffer_offset.numel() else 0
    dev = vs_tetra.device
    keys = torch.zeros((P,), device=dev, dtype=torch.int64)
    vals = torch.zeros((P,), device=dev, dtype=torch.int32)
    check(lib().rm_generate_keys(ptr(vs_tetra.contiguous()), ptr(rect_tile_space.contiguous()),
                                 ptr(index_buffer_offset.contiguous()), vs_tetra.shape[0], grid_height,
                                 grid_width, ptr(keys), ptr(vals), stream_of(dev)), "rm_generate_keys")
    return keys, vals


def sort_by_keys(keys, values, highest_tile_id_msb):
    """Stable radix sort of (int64 key, int32 value) over key bits [0, 32 + msb)."""
    if keys.shape != values.shape:
        raise RuntimeError("keys and values must have the same shape")
    if keys.dtype != torch.int64 or values.dtype != torch.int32:
        raise RuntimeError("keys must be int64 and values int32")
    keys = keys.contiguous()
    values = values.contiguous()
    n = keys.numel()
    ko = torch.empty_like(keys)
    vo = torch.empty_like(values)
    if n == 0:
        return ko, vo
    wsb = lib().rm_sort_pairs_workspace_size(n)
    ws = torch.empty((wsb,), device=keys.device, dtype=torch.uint8)
    check(lib().rm_sort_pairs(ptr(keys), ptr(values), ptr(ko), ptr(vo), n, 32 + int(highest_tile_id_msb),
                              ptr(ws), wsb, stream_of(keys.device)), "rm_sort_pairs")
    return ko, vo


def tile_ranges(sorted_keys, n_tiles):
    ranges = torch.empty((n_tiles, 2), device=sorted_keys.device, dtype=torch.int32)
    check(lib().rm_tile_ranges(ptr(sorted_keys.contiguous()), sorted_keys.numel(), ptr(ranges), n_tiles,
                               stream_of(sorted_keys.device)), "rm_tile_ranges")
    return ranges


class _SortByKeysModule:
    """Stands in for the JIT-built `sort_by_keys_cub` extension object."""
    sort_by_keys = staticmethod(sort_by_keys)


sort_by_keys_cub = _SortByKeysModule()


def vertex_and_tile_shader(indices, vertices, cam, render_grid, fused=True):
    """Projection + binning. Returns (sorted_tetra_idx, tile_ranges, vs_tetra, circumcenter, mask,
    rect_tile_space). ``fused=False`` runs the reference's own decomposition (inclusive scan,
    generate_keys, 64-bit key sort, compute_tile_ranges) on the same kernels' building blocks."""
    tiles, rect, vs, cc = VertexShader.apply(indices, vertices, cam, render_grid)
    with torch.no_grad():
        if fused:
            sorted_idx, ranges = bin_and_sort(tiles, rect, vs, render_grid)
        else:
            offs = inclusive_scan(tiles)
            keys, vals = generate_keys(vs, rect, offs, render_grid.grid_height, render_grid.grid_width)
            msb = (render_grid.grid_width * render_grid.grid_height).bit_length()
            skeys, sorted_idx = sort_by_keys(keys, vals, msb)
            ranges = tile_ranges(skeys, render_grid.grid_height * render_grid.grid_width)
    mask = tiles > 0
    return sorted_idx, ranges, vs, cc, mask, rect


# rm_tile_order workspace kept between calls, one per (device, stream) (calls on one stream run in
# order, so reusing it there is safe).
_order_ws = {}


def tile_order(tile_ranges):
    """int32 [G]: the tiles by decreasing list length (rm_tile_order), the order in which the 4x4-tile
    compositing kernels take them so that the longest walks start first."""
    dev = tile_ranges.device
    G = tile_ranges.shape[0]
    L = lib()
    wsb = L.rm_tile_order_workspace_size(G)
    key = (str(dev), torch.cuda.current_stream(dev).cuda_stream)
    ws = _order_ws.get(key)
    if ws is None or ws.numel() < wsb:
        ws = _order_ws[key] = torch.empty((max(wsb, 1),), device=dev, dtype=torch.uint8)
    order = torch.empty((G,), device=dev, dtype=torch.int32)
    check(L.rm_tile_order(ptr(tile_ranges), G, ptr(order), ptr(ws), ws.numel(), stream_of(dev)), "rm_tile_order")
    return order


class AlphaBlendTiledRender(torch.autograd.Function):
    """Per-tile compositing. forward -> (output_img [H,W,4] = rgb + log-transmittance,
    distortion_img [H,W,5], tet_alive [T] bool (all False, as the reference never writes it))."""

    @staticmethod
    def forward(ctx, sorted_tetra_idx, tile_ranges, indices, vertices, tet_density, render_grid, tcam,
                ray_jitter, device="cuda"):
        ts = (render_grid.tile_height, render_grid.tile_width)
        assert ts[0] == ts[1] and ts[0] in SUPPORTED_TILES, (
            "Alpha Blend Shader was not compiled for this tile"
            f" {ts[0]}x{ts[1]} configuration, available configurations:"
            f" {[(t, t) for t in SUPPORTED_TILES]}")
        sorted_tetra_idx = _require(sorted_tetra_idx, torch.int32, "sorted_tetra_idx")
        tile_ranges = _require(tile_ranges, torch.int32, "tile_ranges", 2)
        indices = _require(indices, torch.int32, "indices", 4)
        vertices = _require(vertices, torch.float32, "vertices", 3)
        tet_density = _require(tet_density, torch.float32, "tet_density", (7, 13))
        ray_jitter = _require(ray_jitter, torch.float32, "ray_jitter")
        H, W = render_grid.image_height, render_grid.image_width
        if tuple(ray_jitter.shape) != (H, W, 2):
            raise RuntimeError(f"ray_jitter must be [{H},{W},2], got {tuple(ray_jitter.shape)}")
        dev = vertices.device
        output_img = torch.empty((H, W, 4), device=dev, dtype=torch.float32)
        distortion_img = torch.empty((H, W, 5), device=dev, dtype=torch.float32)
        n_contributors = torch.empty((H, W, 1), device=dev, dtype=torch.int32)
        tet_alive = torch.zeros((indices.shape[0],), dtype=torch.bool, device=dev)
        cam = _cam(tcam, render_grid)
        F = tet_density.shape[1]
        # per list entry: did some pixel of its tile blend the tet (4x4 and 8x8 tiles); the backward
        # then skips the entries no pixel blended
        P = sorted_tetra_idx.numel()
        flags = torch.empty((P if ts[0] <= 8 else 0,), device=dev, dtype=torch.uint8)
        order = None
        with timing.span("composite_fwd"):
            if flags.numel():
                # 4x4 tiles: the longest lists first (the kernels take any order; results do not depend on it)
                order = tile_order(tile_ranges) if ts[0] == 4 else None
                check(lib().rm_composite_fwd_flags(cam, ptr(sorted_tetra_idx), P, ptr(tile_ranges), ptr(indices),
                                                   ptr(vertices), ptr(tet_density), F, ptr(ray_jitter), ptr(output_img),
                                                   ptr(distortion_img), ptr(n_contributors), ptr(flags), ptr(order),
                                                   stream_of(dev)), "rm_composite_fwd_flags")
            else:
                check(lib().rm_composite_fwd(cam, ptr(sorted_tetra_idx), P, ptr(tile_ranges), ptr(indices),
                                             ptr(vertices), ptr(tet_density), F, ptr(ray_jitter), ptr(output_img),
                                             ptr(distortion_img), ptr(n_contributors), stream_of(dev)),
                      "rm_composite_fwd")
        non_tensor, tensor_data = split_tensors(tcam)
        ctx.save_for_backward(sorted_tetra_idx, tile_ranges, indices, vertices, tet_density, output_img,
                              distortion_img, n_contributors, ray_jitter, flags, *tensor_data)
        ctx.non_tensor_data = non_tensor
        ctx.render_grid = render_grid
        ctx.cam = cam
        ctx.tile_order = order
        ctx.mark_non_differentiable(tet_alive)
        return output_img, distortion_img, tet_alive

    @staticmethod
    def backward(ctx, grad_output_img, grad_distortion_img, grad_tet_alive):
        (sorted_tetra_idx, tile_ranges, indices, vertices, tet_density, output_img, distortion_img,
         n_contributors, ray_jitter, flags) = ctx.saved_tensors[:10]
        if grad_output_img is None:
            grad_output_img = torch.zeros_like(output_img)
        if grad_distortion_img is None:
            grad_distortion_img = torch.zeros_like(distortion_img)
        grad_output_img = grad_output_img.contiguous().float()
        grad_distortion_img = grad_distortion_img.contiguous().float()
        vertices_grad = torch.zeros_like(vertices)
        tet_density_grad = torch.zeros_like(tet_density)
        L = lib()
        n_tets = indices.shape[0]
        wsb = L.rm_composite_bwd_workspace_size(n_tets)
        ws = torch.empty((max(wsb, 1),), device=vertices.device, dtype=torch.uint8)
        with timing.span("composite_bwd"):
            if flags.numel():
                check(L.rm_composite_bwd_flags(ctx.cam, ptr(sorted_tetra_idx), sorted_tetra_idx.numel(),
                                               ptr(tile_ranges), ptr(indices), n_tets, ptr(vertices), ptr(tet_density),
                                               tet_density.shape[1], ptr(ray_jitter), ptr(output_img),
                                               ptr(distortion_img), ptr(n_contributors), ptr(flags),
                                               ptr(ctx.tile_order), ptr(grad_output_img), ptr(grad_distortion_img),
                                               ptr(vertices_grad),
                                               ptr(tet_density_grad), ptr(ws), wsb, stream_of(vertices.device)),
                      "rm_composite_bwd_flags")
            else:
                check(L.rm_composite_bwd_ws(ctx.cam, ptr(sorted_tetra_idx), sorted_tetra_idx.numel(),
                                            ptr(tile_ranges), ptr(indices), n_tets, ptr(vertices), ptr(tet_density),
                                            tet_density.shape[1], ptr(ray_jitter), ptr(output_img),
                                            ptr(distortion_img), ptr(n_contributors), ptr(grad_output_img),
                                            ptr(grad_distortion_img), ptr(vertices_grad), ptr(tet_density_grad),
                                            ptr(ws), wsb, stream_of(vertices.device)), "rm_composite_bwd_ws")
        return None, None, None, vertices_grad, tet_density_grad, None, None, None


def set_intersect_precision(fp64: bool):
    """FP64 (default) or FP32 ray/plane intersection in the compositing kernels."""
    check(lib().rm_set_intersect_precision(1 if fp64 else 0), "rm_set_intersect_precision")


def distortion_loss(distortion_img):
    """utils/train_util.py:66-67."""
    total_density = (distortion_img[:, :, 2] ** 2).clip(min=1e-6)
    return (((distortion_img[:, :, 0] - distortion_img[:, :, 1]) + distortion_img[:, :, 4])
            / total_density).clip(min=0)

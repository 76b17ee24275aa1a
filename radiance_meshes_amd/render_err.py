"""Per-tet error votes for densification, with the reference's interface
(delaunay_rasterization/internal/render_err.py:33-159, utils/ssim.py:29-66):

* ``render_err(gt_image, camera, model, tile_size=16, min_t=0.1)`` renders the view (projection,
  binning and compositing on the HIP kernels), forms the per-pixel signed L1 error and SSIM error,
  masks both with the camera's alpha mask, and gathers per-tet votes with ``calc_tet_err``;
  returns ``(tet_err [T,16], {tet_count [T,2], pixel_err, ssim_err, render_img, cell_values})``;
* ``calc_tet_err`` is the vote kernel alone (``rm_tet_err``): channel layout in ``include/rmesh.h``;
* ``ssim`` is the per-pixel SSIM map (Gaussian window 11, sigma 1.5, zero padding) and
  ``gaussian_blur`` the depth-wise Gaussian blur the reference module also exports.

The reference allocates the forward's distortion buffer with 4 channels while its kernel writes 5
(render_err.py:70-72); here it has the 5 channels the compositing kernel writes.
"""
from __future__ import annotations

import torch
import torch.nn.functional as F

from . import timing
from ._capi import check, lib, ptr, stream_of
from .rasterizer import RenderGrid, _cam, _require, tile_order, vertex_and_tile_shader


def _gauss1d(size: int, sigma: float, device=None, dtype=torch.float32) -> torch.Tensor:
    x = torch.arange(size, dtype=torch.float64) - size // 2
    g = torch.exp(-(x * x) / (2.0 * sigma * sigma))
    return (g / g.sum()).to(device=device, dtype=dtype)


def gaussian_blur(img: torch.Tensor, kernel_size: int = 5, sigma: float = 1.5) -> torch.Tensor:
    """Depth-wise Gaussian blur of a [C,H,W] image (render_err.py:12-31), zero padded."""
    g = _gauss1d(kernel_size, sigma, img.device, img.dtype)
    k = torch.outer(g, g).expand(img.shape[0], 1, kernel_size, kernel_size).contiguous()
    return F.conv2d(img.unsqueeze(0), k, padding=kernel_size // 2, groups=img.shape[0]).squeeze(0)


def ssim(img1: torch.Tensor, img2: torch.Tensor, window_size: int = 11) -> torch.Tensor:
    """Per-pixel SSIM map of two [C,H,W] (or [N,C,H,W]) images (utils/ssim.py:39-66): local
    means, variances and covariance under a normalised Gaussian window (sigma 1.5, zero padding),
    stabilisers C1 = 0.01^2 and C2 = 0.03^2."""
    ch = img1.size(-3)
    g = _gauss1d(window_size, 1.5, img1.device, img1.dtype)
    win = torch.outer(g, g).expand(ch, 1, window_size, window_size).contiguous()
    pad = window_size // 2

    def blur(x):
        return F.conv2d(x, win, padding=pad, groups=ch)

    mu1, mu2 = blur(img1), blur(img2)
    mu11, mu22, mu12 = mu1 * mu1, mu2 * mu2, mu1 * mu2
    s11 = blur(img1 * img1) - mu11
    s22 = blur(img2 * img2) - mu22
    s12 = blur(img1 * img2) - mu12
    c1, c2 = 0.01 ** 2, 0.03 ** 2
    return ((2 * mu12 + c1) * (2 * s12 + c2)) / ((mu11 + mu22 + c1) * (s11 + s22 + c2))


def calc_tet_err(sorted_tetra_idx, tile_ranges, indices, vertices, tet_density, render_grid, tcam, pixel_err,
                 ssim_err, tet_err=None, tet_count=None):
    """Adds the per-tet votes of one view into tet_err [T,16] f32 / tet_count [T,2] i32 (fresh zeros
    when not given) and returns both."""
    indices = _require(indices, torch.int32, "indices", 4)
    vertices = _require(vertices, torch.float32, "vertices", 3)
    tet_density = _require(tet_density.detach(), torch.float32, "tet_density", (7, 13))
    sorted_tetra_idx = _require(sorted_tetra_idx, torch.int32, "sorted_tetra_idx")
    tile_ranges = _require(tile_ranges, torch.int32, "tile_ranges")
    H, W = render_grid.image_height, render_grid.image_width
    pixel_err = _require(pixel_err.detach(), torch.float32, "pixel_err")
    ssim_err = _require(ssim_err.detach(), torch.float32, "ssim_err")
    if tuple(pixel_err.shape) != (H, W) or tuple(ssim_err.shape) != (H, W):
        raise RuntimeError(f"pixel_err and ssim_err must have shape [{H}, {W}]")
    T = indices.shape[0]
    dev = vertices.device
    if tet_err is None:
        tet_err = torch.zeros((T, 16), device=dev, dtype=torch.float32)
    if tet_count is None:
        tet_count = torch.zeros((T, 2), device=dev, dtype=torch.int32)
    if tet_err.shape != (T, 16) or tet_err.dtype != torch.float32 or not tet_err.is_contiguous():
        raise RuntimeError("tet_err must be a contiguous float32 [T, 16] tensor")
    if tet_count.shape != (T, 2) or tet_count.dtype != torch.int32 or not tet_count.is_contiguous():
        raise RuntimeError("tet_count must be a contiguous int32 [T, 2] tensor")
    with timing.span("tet_err"):
        # 4x4 tiles: the longest lists first (the votes do not depend on the order)
        order = tile_order(tile_ranges.reshape(-1, 2)) if render_grid.tile_height == 4 else None
        check(lib().rm_tet_err(_cam(tcam, render_grid), ptr(sorted_tetra_idx), sorted_tetra_idx.numel(),
                               ptr(tile_ranges), ptr(indices), ptr(vertices), ptr(tet_density),
                               tet_density.shape[1], ptr(pixel_err), ptr(ssim_err), ptr(tet_err), ptr(tet_count),
                               T, ptr(order), stream_of(dev)), "rm_tet_err")
    return tet_err, tet_count


def render_err(gt_image, camera, model, tile_size=16, min_t=0.1, **kwargs):
    """Votes of one training view for densification (render_err.py:33-159). `model` needs
    `indices`, `vertices`, `feature_dim` and `get_cell_values(camera, mask)`; `camera` needs
    `image_height`, `image_width`, `to_dict(device)` and `gt_alpha_mask` [1|3,H,W]."""
    indices, vertices = model.indices, model.vertices
    device = indices.device
    render_grid = RenderGrid(camera.image_height, camera.image_width, tile_height=tile_size, tile_width=tile_size)
    tcam = dict(tile_height=tile_size, tile_width=tile_size, grid_height=render_grid.grid_height,
                grid_width=render_grid.grid_width, min_t=min_t, **camera.to_dict(device))
    with torch.no_grad():
        sorted_tetra_idx, tile_ranges, vs_tetra, circumcenter, mask, _ = vertex_and_tile_shader(
            indices, vertices, tcam, render_grid)
        cell_values = torch.zeros((mask.shape[0], model.feature_dim), device=device)
        _, cell_values[mask] = model.get_cell_values(camera, mask)
        H, W = render_grid.image_height, render_grid.image_width
        output_img = torch.empty((H, W, 4), device=device)
        distortion_img = torch.empty((H, W, 5), device=device)
        n_contributors = torch.empty((H, W), device=device, dtype=torch.int32)
        ray_jitter = torch.full((H, W, 2), 0.5, device=device)
        cv = cell_values.contiguous()
        with timing.span("composite_fwd"):
            check(lib().rm_composite_fwd(_cam(tcam, render_grid), ptr(sorted_tetra_idx), sorted_tetra_idx.numel(),
                                         ptr(tile_ranges), ptr(indices.contiguous()), ptr(vertices.contiguous()),
                                         ptr(cv), cv.shape[1], ptr(ray_jitter), ptr(output_img),
                                         ptr(distortion_img), ptr(n_contributors), stream_of(device)),
                  "rm_composite_fwd")
        render_img = output_img.permute(2, 0, 1)[:3, ...].clip(min=0, max=1)
        pixel_err = (render_img - gt_image).mean(dim=0).contiguous()
        ssim_err = (1 - ssim(render_img, gt_image).mean(dim=0)).clip(min=0, max=1)
        amask = camera.gt_alpha_mask.to(device)
        pixel_err = pixel_err * amask[0]
        ssim_err = ssim_err * amask[0]
        render_img = render_img * amask
        tet_err, tet_count = calc_tet_err(sorted_tetra_idx, tile_ranges, indices, vertices, cv, render_grid, tcam,
                                          pixel_err, ssim_err)
    return tet_err, dict(tet_count=tet_count, pixel_err=pixel_err, ssim_err=ssim_err, render_img=render_img,
                         cell_values=cell_values)

"""`render()` with the signature and return dict of the reference's entry point
(utils/train_util.py:15-77), running projection, binning and compositing on the HIP kernels.

`model` needs the attributes the reference's render() reads: `vertices` [V,3] f32, `indices`
[T,4] i32, `device`, `feature_dim`, `mask_values`, and `get_cell_values(camera, mask, circumcenters)`
returning (shs [T',C,3] or None, features [T',feature_dim]); `camera` needs `to_dict(device)`,
`image_height`, `image_width` and `gt_alpha_mask` (data/camera.py:72-91).
"""
from __future__ import annotations

import torch

from .rasterizer import AlphaBlendTiledRender, RenderGrid, distortion_loss, vertex_and_tile_shader


def weigh_degree(sh_coeffs: torch.Tensor, values) -> torch.Tensor:
    """Per-coefficient weight by SH degree floor(sqrt(i)) (utils/eval_sh_py.py:28-40)."""
    C = sh_coeffs.shape[1]
    idx = torch.arange(C, device=sh_coeffs.device, dtype=torch.float32)
    deg = torch.floor(torch.sqrt(idx)).long()
    return torch.tensor(values, device=sh_coeffs.device, dtype=sh_coeffs.dtype)[deg].view(1, C, 1)


def render(camera, model, cell_values=None, tile_size=16, min_t=0.1, scene_scaling=1, clip_multi=0,
           ray_jitter=None, **kwargs):
    device = model.device
    if ray_jitter is None:
        ray_jitter = 0.5 * torch.ones((camera.image_height, camera.image_width, 2), device=device)
    else:
        assert ray_jitter.shape[0] == camera.image_height
        assert ray_jitter.shape[1] == camera.image_width
        assert ray_jitter.shape[2] == 2
    vertices = model.vertices
    render_grid = RenderGrid(camera.image_height, camera.image_width, tile_height=tile_size,
                             tile_width=tile_size)
    tcam = dict(tile_height=tile_size, tile_width=tile_size, grid_height=render_grid.grid_height,
                grid_width=render_grid.grid_width, min_t=min_t, **camera.to_dict(device))
    sorted_tetra_idx, tile_ranges, vs_tetra, circumcenter, mask, _ = vertex_and_tile_shader(
        model.indices, vertices, tcam, render_grid)
    sh_reg = 0
    if cell_values is None:
        cell_values = torch.zeros((mask.shape[0], model.feature_dim), device=circumcenter.device)
        if mask.sum() > 0 and model.mask_values:
            shs, values = model.get_cell_values(camera, mask, circumcenter[mask])
            cell_values[mask] = values
        else:
            shs, cell_values = model.get_cell_values(camera, all_circumcenters=circumcenter)
        if shs is not None and shs.numel() > 0:
            sh_reg = ((weigh_degree(shs, [0, 0.1, 0.5, 1]) * shs) ** 2).mean()
    image_rgb, distortion_img, tet_alive = AlphaBlendTiledRender.apply(
        sorted_tetra_idx, tile_ranges, model.indices, vertices, cell_values, render_grid, tcam, ray_jitter)
    alpha = image_rgb.permute(2, 0, 1)[3, ...]
    return {
        "render": image_rgb.permute(2, 0, 1)[:3, ...] * camera.gt_alpha_mask.to(device),
        "alpha": alpha,
        "distortion_loss": distortion_loss(distortion_img).mean(),
        "mask": mask,
        "sh_reg": sh_reg,
    }

"""Per-tet feature activation of the reference's field model on one HIP kernel per direction.

`activate_output(camera_center, density, rgb, grd, sh, indices, circumcenters, vertices,
current_sh_deg, max_sh_deg)` has the signature and result of the reference's
(utils/model_util.py:93-105): features [T,7] = [density, softplus_10(eval_sh(...)), normed_grd],
with offset_normalize (:81-91) and the vertex-0 colour of the linear field (:62-78) inside. The
backward returns gradients for density, rgb, grd, sh, circumcenters and vertices with the
reference's detach pattern (the vertex-0 offset used for the colour is a constant; the radius and
the view direction are not).
"""
from __future__ import annotations

import torch

from . import timing
from ._capi import check, lib, ptr, stream_of


def _f32(t, shape_tail, name):
    t = t.reshape(-1, *shape_tail) if shape_tail else t.reshape(-1)
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a device tensor")
    return t.to(torch.float32).contiguous()


class ActivateOutput(torch.autograd.Function):
    @staticmethod
    def forward(ctx, camera_center, density, rgb, grd, sh, indices, circumcenters, vertices, sh_degree):
        T = indices.shape[0]
        cam = _f32(camera_center.detach(), None, "camera_center")
        dens = _f32(density.detach(), None, "density")
        rgb_ = _f32(rgb.detach(), (3,), "rgb")
        grd_ = _f32(grd.detach(), (3,), "grd")
        sh_ = _f32(sh.detach(), (sh.shape[-2], 3), "sh") if sh is not None and sh.numel() else None
        K = sh_.shape[1] if sh_ is not None else 0
        idx = indices.to(torch.int32).contiguous()
        cc = _f32(circumcenters.detach(), (3,), "circumcenters")
        verts = _f32(vertices.detach(), (3,), "vertices")
        if dens.shape[0] != T or rgb_.shape[0] != T or grd_.shape[0] != T or cc.shape[0] != T or \
                (sh_ is not None and sh_.shape[0] != T):
            raise RuntimeError("per-tet inputs must have one row per tet")
        out = torch.empty((T, 7), device=verts.device, dtype=torch.float32)
        with timing.span("activate_fwd"):
            check(lib().rm_activate_fwd(ptr(cam), ptr(dens), ptr(rgb_), ptr(grd_), ptr(sh_), K, int(sh_degree),
                                        ptr(idx), ptr(cc), ptr(verts), T, ptr(out), stream_of(verts.device)),
                  "rm_activate_fwd")
        ctx.save_for_backward(cam, dens, rgb_, grd_, sh_ if sh_ is not None else torch.empty(0), idx, cc, verts)
        ctx.meta = (K, int(sh_degree), density.shape, rgb.shape, grd.shape, None if sh is None else sh.shape,
                    circumcenters.shape, vertices.shape)
        return out

    @staticmethod
    def backward(ctx, gout):
        cam, dens, rgb_, grd_, sh_, idx, cc, verts = ctx.saved_tensors
        K, deg, ds, rs, gs, ss, cs, vs = ctx.meta
        need = ctx.needs_input_grad
        T = idx.shape[0]
        g = gout.to(torch.float32).contiguous()
        gd = torch.empty_like(dens) if need[1] else None
        gr = torch.empty_like(rgb_) if need[2] else None
        gg = torch.empty_like(grd_) if need[3] else None
        gs_ = torch.empty_like(sh_) if (need[4] and K > 0) else None
        gc = torch.empty_like(cc) if need[6] else None
        gv = torch.zeros_like(verts) if need[7] else None
        with timing.span("activate_bwd"):
            check(lib().rm_activate_bwd(ptr(cam), ptr(dens), ptr(rgb_), ptr(grd_), ptr(sh_) if K else None, K, deg,
                                        ptr(idx), ptr(cc), ptr(verts), T, ptr(g), ptr(gd), ptr(gr), ptr(gg), ptr(gs_),
                                        ptr(gv), ptr(gc), stream_of(verts.device)), "rm_activate_bwd")
        shape = lambda t, s: None if t is None else t.reshape(s)  # noqa: E731
        return (None, shape(gd, ds), shape(gr, rs), shape(gg, gs), shape(gs_, ss) if gs_ is not None else None,
                None, shape(gc, cs), shape(gv, vs), None)


def activate_output(camera_center, density, rgb, grd, sh, indices, circumcenters, vertices,
                    current_sh_deg: int, max_sh_deg: int):
    """utils/model_util.py:93-105 on the HIP kernels: features [T,7]. sh is [T, K, 3] with
    K >= (current_sh_deg + 1)^2 - 1 (the first coefficients are used); degrees up to 3."""
    if current_sh_deg > 3:
        raise ValueError("SH degree above 3 is not supported")
    return ActivateOutput.apply(camera_center, density, rgb, grd, sh, indices, circumcenters, vertices,
                                int(current_sh_deg))

"""The reference model's field on the HIP kernels: iNGPDW (models/ingp_color.py:100-205) and the
field half of Model (compute_batch_features, get_cell_values; models/ingp_color.py:314-384).

* ``iNGPDW``                 same constructor arguments and forward(x, cr) -> (density [N,1],
                             rgb [N,3], grd [N,1,3], sh [N,sh_dim] float16): the hash grid
                             (HashEmbedderOptimized's level table, computed with the same float32
                             torch ops), the erf down-weighting (rm_field_downweight), the four SELU
                             heads on MFMA (rm_field_head_fwd/bwd) and the output activations
                             (rm_field_outputs_fwd/bwd). Built for hidden_dim 64, L * hashmap_dim
                             in {32, 64} and sh_dim <= 48 (train.py's field is L = 8, F = 8,
                             log2 23, base 64, per-level scale 2, 45 SH outputs); k_samples 1.
* ``compute_batch_features`` query anchor, FP64 circumradius and contraction (rm_field_query),
                             then the backbone; returns (anchor, density, rgb, grd, sh).
* ``Model``                  the parts of the reference's Model that render() and evaluate.py
                             use (vertices, indices, get_cell_values through activate_output,
                             compute_batch_features) and load_ckpt for its ckpt.pth + config.json.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import timing
from ._capi import check, lib, ptr, stream_of
from .activate import activate_output
from .field import HashGrid, _MLP


def reference_level_table(n_levels: int, base_resolution: int, finest_resolution: float, log2_size: int):
    """HashEmbedderOptimized's per-level resolutions (utils/hashgrid.py:31-48): growth
    b = exp((log(finest) - log(base)) / (L - 1)) and res_l = floor(base * b^l), all in float32 torch
    ops as the reference computes them; level sizes min(2^log2, res^3), concatenated."""
    base = torch.tensor(base_resolution, dtype=torch.float32)
    finest = torch.tensor(finest_resolution, dtype=torch.float32)
    b = torch.exp((finest.log() - base.log()) / (n_levels - 1))
    res = [int(torch.floor(base * (b ** i)).item()) for i in range(n_levels)]
    sizes = [min(2 ** log2_size, r ** 3) for r in res]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    return np.array(res, np.int32), offs


class _Downweight(torch.autograd.Function):
    """enc * s(cr) per element (rm_field_downweight); cr is detached, so the adjoint is the same map."""

    @staticmethod
    def forward(ctx, enc, cr, n_levels, n_feat, scale_multi, per_level_scale):
        enc = enc.contiguous()
        cr = cr.detach().reshape(-1).to(torch.float64).contiguous()
        out = torch.empty_like(enc)
        with timing.span("field_glue"):
            check(lib().rm_field_downweight(ptr(enc), ptr(cr), enc.shape[0], n_levels, n_feat, float(scale_multi),
                                            float(per_level_scale), ptr(out), stream_of(enc.device)),
                  "rm_field_downweight")
        ctx.save_for_backward(cr)
        ctx.meta = (n_levels, n_feat, float(scale_multi), float(per_level_scale))
        return out

    @staticmethod
    def backward(ctx, gout):
        (cr,) = ctx.saved_tensors
        L, F, sm, pls = ctx.meta
        g = gout.contiguous()
        gin = torch.empty_like(g)
        with timing.span("field_glue"):
            check(lib().rm_field_downweight(ptr(g), ptr(cr), g.shape[0], L, F, sm, pls, ptr(gin), stream_of(g.device)),
                  "rm_field_downweight")
        return gin, None, None, None, None, None


class _Outputs(torch.autograd.Function):
    """Raw head outputs -> (density, rgb, grd, sh half) (rm_field_outputs_fwd/bwd)."""

    @staticmethod
    def forward(ctx, sigma, rgb, grd, sh, density_offset):
        n = sigma.shape[0]
        n_sh = sh.shape[1]
        dev = sigma.device
        sigma, rgb, grd, sh = (t.contiguous() for t in (sigma, rgb, grd, sh))
        density = torch.empty((n, 1), device=dev, dtype=torch.float32)
        rgb_out = torch.empty((n, 3), device=dev, dtype=torch.float32)
        grd_out = torch.empty((n, 3), device=dev, dtype=torch.float32)
        sh_out = torch.empty((n, n_sh), device=dev, dtype=torch.float16)
        with timing.span("field_glue"):
            check(lib().rm_field_outputs_fwd(ptr(sigma), ptr(rgb), ptr(grd), ptr(sh) if n_sh else None, n, n_sh,
                                             float(density_offset), ptr(density), ptr(rgb_out), ptr(grd_out),
                                             ptr(sh_out) if n_sh else None, stream_of(dev)), "rm_field_outputs_fwd")
        ctx.save_for_backward(density, grd)
        ctx.n_sh = n_sh
        return density, rgb_out, grd_out, sh_out

    @staticmethod
    def backward(ctx, g_density, g_rgb, g_grd, g_sh):
        density, grd = ctx.saved_tensors
        n, n_sh = density.shape[0], ctx.n_sh
        dev = density.device
        c = lambda t, dt: None if t is None else t.to(dt).contiguous()  # noqa: E731
        g_density, g_rgb, g_grd = c(g_density, torch.float32), c(g_rgb, torch.float32), c(g_grd, torch.float32)
        g_sh = c(g_sh, torch.float16)
        gs = torch.empty((n, 1), device=dev, dtype=torch.float32)
        gr = torch.empty((n, 3), device=dev, dtype=torch.float32)
        gg = torch.empty((n, 3), device=dev, dtype=torch.float32)
        gsh = torch.empty((n, n_sh), device=dev, dtype=torch.float32)
        with timing.span("field_glue"):
            check(lib().rm_field_outputs_bwd(ptr(density), ptr(grd), ptr(g_density), ptr(g_rgb), ptr(g_grd),
                                             ptr(g_sh) if n_sh else None, n, n_sh, ptr(gs), ptr(gr), ptr(gg),
                                             ptr(gsh) if n_sh else None, stream_of(dev)), "rm_field_outputs_bwd")
        return gs, gr, gg, gsh, None


class _Head(torch.nn.Module):
    """mk_head(n) (models/ingp_color.py:141-151): Linear(n_in, 64) SELU Linear(64, 64) SELU
    Linear(64, n), weights [out, in] as nn.Linear; evaluated by the fused MFMA kernels."""

    def __init__(self, n_in, hidden, n_out, last_eps, precision, device, gen):
        super().__init__()
        gain = math.sqrt(2.0)      # nn.init.calculate_gain('relu')

        def xavier(o, i):
            a = gain * math.sqrt(6.0 / (i + o))
            return torch.nn.Parameter(((torch.rand((o, i), generator=gen) * 2 - 1) * a).to(device))

        self.w1, self.w2 = xavier(hidden, n_in), xavier(hidden, hidden)
        self.w3 = torch.nn.Parameter(((torch.rand((n_out, hidden), generator=gen) * 2 - 1) * last_eps).to(device))
        self.b1 = torch.nn.Parameter(torch.zeros(hidden, device=device))
        self.b2 = torch.nn.Parameter(torch.zeros(hidden, device=device))
        self.b3 = torch.nn.Parameter(torch.zeros(n_out, device=device))
        self.precision = precision
        self.n_out = n_out

    def forward(self, h):
        if self.n_out == 0:
            return torch.zeros((h.shape[0], 0), device=h.device, dtype=torch.float32)
        return _MLP.apply(h, self.w1, self.b1, self.w2, self.b2, self.w3, self.b3, self.precision)


class iNGPDW(torch.nn.Module):  # noqa: N801  (the reference's class name)
    """models/ingp_color.py:100-205 with the same constructor arguments; see the module docstring
    for the supported shapes. `precision` "fp32" or "bf16" picks the heads' MFMA operand type."""

    def __init__(self, sh_dim=0, scale_multi=0.5, log2_hashmap_size=16, base_resolution=16, per_level_scale=2,
                 L=10, hashmap_dim=4, hidden_dim=64, g_init=1, s_init=1e-4, d_init=0.1, c_init=0.6,
                 density_offset=-4, ablate_downweighing=False, k_samples=1, trunc_sigma=1, use_tcnn=False,
                 precision="fp32", device="cuda", seed=0, **kwargs):
        super().__init__()
        if use_tcnn:
            raise ValueError("use_tcnn: the tiny-cuda-nn encoding is not part of this build")
        if k_samples != 1:
            raise ValueError("k_samples > 1 (sampled-sphere queries) is not built; use k_samples=1 as train.py does")
        if hidden_dim != 64 or L * hashmap_dim not in (32, 64) or not 0 <= sh_dim <= 48:
            raise ValueError("built for hidden_dim 64, L * hashmap_dim in {32, 64} and sh_dim <= 48")
        self.scale_multi = scale_multi
        self.L = L
        self.dim = hashmap_dim
        self.per_level_scale = per_level_scale
        self.base_resolution = base_resolution
        self.density_offset = density_offset
        self.ablate_downweighing = ablate_downweighing
        self.k_samples = k_samples
        self.trunc_sigma = trunc_sigma
        self.compile = False
        self.config = dict(per_level_scale=per_level_scale, n_levels=L, otype="HashGrid",
                           n_features_per_level=hashmap_dim, base_resolution=base_resolution,
                           log2_hashmap_size=log2_hashmap_size)
        res, offs = reference_level_table(L, base_resolution, base_resolution * per_level_scale ** L,
                                          log2_hashmap_size)
        self.encoding = HashGrid(n_levels=L, n_features_per_level=hashmap_dim, log2_hashmap_size=log2_hashmap_size,
                                 device=device, seed=seed, level_table=(res, offs))
        self.encoding.n_output_dims = L * hashmap_dim
        prec = {"fp32": 0, "bf16": 1}[precision]
        gen = torch.Generator().manual_seed(seed + 1)
        n_in = L * hashmap_dim
        self.density_net = _Head(n_in, hidden_dim, 1, d_init, prec, device, gen)
        self.color_net = _Head(n_in, hidden_dim, 3, c_init, prec, device, gen)
        self.gradient_net = _Head(n_in, hidden_dim, 3, g_init, prec, device, gen)
        self.sh_net = _Head(n_in, hidden_dim, sh_dim, s_init, prec, device, gen)

    def _encode(self, x, cr):
        out = self.encoding(x)
        if not self.ablate_downweighing:
            out = _Downweight.apply(out, cr, self.L, self.dim, self.scale_multi, self.per_level_scale)
        return out

    def forward(self, x, cr):
        h = self._encode(x, cr)
        density, rgb, grd, sh = _Outputs.apply(self.density_net(h), self.color_net(h), self.gradient_net(h),
                                               self.sh_net(h), self.density_offset)
        return density, rgb, grd.reshape(-1, 1, 3), sh


def field_query(vertices, indices, center, scene_scaling, anchor="centroid"):
    """(x [T,3] float32 grid coordinates, cr [T] float64, anchor [T,3] float32) for every tet of
    indices (rm_field_query). No gradient: the reference detaches the anchor and the radius."""
    vertices = vertices.detach().to(torch.float32).contiguous()
    idx = indices.to(torch.int32).contiguous()
    T = idx.shape[0]
    dev = vertices.device
    x = torch.empty((T, 3), device=dev, dtype=torch.float32)
    cr = torch.empty((T,), device=dev, dtype=torch.float64)
    anc = torch.empty((T, 3), device=dev, dtype=torch.float32)
    c = np.ascontiguousarray(torch.as_tensor(center).detach().reshape(3).cpu().numpy(), np.float32)
    with timing.span("field_glue"):
        check(lib().rm_field_query(ptr(idx), ptr(vertices), T, c.ctypes.data, float(scene_scaling),
                                   {"centroid": 0, "circumcenter": 1}[anchor], ptr(x), ptr(cr), ptr(anc),
                                   stream_of(dev)), "rm_field_query")
    return x, cr, anc


def compute_batch_features(backbone, vertices, indices, start, end, center, scene_scaling,
                           ablate_circumsphere=True):
    """Model.compute_batch_features (models/ingp_color.py:314-357) for k_samples = 1 and alpha = 0:
    (anchor, density, rgb, grd, sh) of tets [start, end). The circumradius is always recomputed
    in FP64 from the tet (the reference's precomputed-circumcenter path gives the same radius up
    to FP64 rounding)."""
    x, cr, anc = field_query(vertices, indices[start:end], center, scene_scaling,
                             "centroid" if ablate_circumsphere else "circumcenter")
    return (anc, *backbone(x, cr))


_HEADS = ("density_net", "color_net", "gradient_net", "sh_net")


class Model(torch.nn.Module):
    """The parts of models/ingp_color.Model (:208-474) that render() and evaluate.py use: the
    vertices (contracted_vertices + ext_vertices), indices, the backbone iNGPDW and the scene
    normalisation, get_cell_values, compute_batch_features and load_ckpt. Constructor arguments as
    the reference's; `density_offset` is consumed here and not passed on, so the backbone keeps
    its own default (-4), as in the reference."""

    def __init__(self, vertices, ext_vertices, center, scene_scaling, density_offset=-1, current_sh_deg=2,
                 max_sh_deg=2, ablate_circumsphere=False, ablate_gradient=False, device="cuda", **kwargs):
        super().__init__()
        self.device = torch.device(device)
        self.density_offset = density_offset
        self.max_sh_deg = max_sh_deg
        self.current_sh_deg = current_sh_deg
        self.sh_dim = (max_sh_deg + 1) ** 2 - 1
        self.backbone = iNGPDW(self.sh_dim * 3, device=device, **kwargs)
        self.register_buffer("ext_vertices", torch.as_tensor(ext_vertices, dtype=torch.float32).to(device))
        self.register_buffer("center", torch.as_tensor(center, dtype=torch.float32).reshape(1, 3).to(device))
        self.register_buffer("scene_scaling", torch.as_tensor(float(scene_scaling), dtype=torch.float32).to(device))
        self.contracted_vertices = torch.nn.Parameter(torch.as_tensor(vertices, dtype=torch.float32).to(device))
        self.indices = torch.empty((0, 4), dtype=torch.int32, device=device)
        self.empty_indices = torch.empty((0, 4), dtype=torch.int32, device=device)
        self.chunk_size = 308576
        self.mask_values = True
        self.feature_dim = 7
        self.alpha = 0
        self.ablate_circumsphere = ablate_circumsphere
        self.ablate_gradient = ablate_gradient
        self.min_t = 0.1

    @property
    def vertices(self):
        return torch.cat([self.contracted_vertices, self.ext_vertices])

    @property
    def num_int_verts(self):
        return self.contracted_vertices.shape[0]

    def compute_batch_features(self, vertices, indices, start, end, circumcenters=None):
        """(anchor, density, rgb, grd, sh) of tets [start, end) (models/ingp_color.py:314-357).
        The radius is recomputed in FP64 from the tet when circumcenters are given too (the
        reference then takes |circumcenter - v0| from them; both agree to float32 rounding)."""
        return compute_batch_features(self.backbone, vertices, indices, start, end, self.center,
                                      float(self.scene_scaling), self.ablate_circumsphere)

    def get_cell_values(self, camera, mask=None, all_circumcenters=None, radii=None):
        """models/ingp_color.py:361-384: (shs [T, sh_dim, 3] float32, features [T, 7]) over the
        (masked) tets, in the reference's chunks, through activate_output at camera.camera_center."""
        indices = self.indices[mask] if mask is not None else self.indices
        vertices = self.vertices
        T = indices.shape[0]
        features = torch.empty((T, self.feature_dim), device=self.device)
        shs = torch.empty((T, self.sh_dim, 3), device=self.device)
        cam_center = torch.as_tensor(camera.camera_center).to(self.device)
        for start in range(0, T, self.chunk_size):
            end = min(start + self.chunk_size, T)
            _, density, rgb, grd, sh = self.compute_batch_features(vertices, indices, start, end, all_circumcenters)
            if self.ablate_gradient:
                grd = torch.zeros_like(grd)
            centroids = vertices[indices[start:end].long()].mean(dim=1)
            sh3 = sh.reshape(-1, self.sh_dim, 3)
            shs[start:end] = sh3
            features[start:end] = activate_output(cam_center, density, rgb, grd, sh3, indices[start:end], centroids,
                                                  vertices.detach(), self.current_sh_deg, self.max_sh_deg)
        return shs, features

    # ------------------------------------------------------------------ checkpoints
    def load_reference_state_dict(self, sd):
        """Loads a state dict in the reference's naming (Model.state_dict(), with or without the
        '_orig_mod.' level torch.compile adds under backbone): geometry buffers, the per-level
        embedding tables (concatenated into the grid's table) and the four heads' Linear layers
        (indices 0, 2, 4 of each nn.Sequential). Other entries (the unused 'network' head, the
        encoding's constant buffers, sphere_pattern) are checked where they constrain the shape
        and otherwise ignored."""
        sd = {k.replace("backbone._orig_mod.", "backbone."): v for k, v in sd.items()}
        dev = self.device
        with torch.no_grad():
            self.contracted_vertices.data = sd["contracted_vertices"].to(dev, torch.float32).contiguous()
            self.ext_vertices = sd["ext_vertices"].to(dev, torch.float32).contiguous()
            self.center.copy_(sd["center"].reshape(1, 3))
            self.scene_scaling.copy_(sd["scene_scaling"].reshape(()))
            enc = self.backbone.encoding
            if "backbone.encoding.level_resolutions" in sd:
                lr = sd["backbone.encoding.level_resolutions"].cpu().numpy().astype(np.int64)
                if not np.array_equal(lr, enc.res.astype(np.int64)):
                    raise ValueError(f"level resolutions {lr.tolist()} differ from this grid's {enc.res.tolist()}")
            for l in range(enc.n_levels):
                w = sd[f"backbone.encoding.embeddings.{l}.weight"]
                rows = int(enc.offs[l + 1] - enc.offs[l])
                if tuple(w.shape) != (rows, enc.n_feat):
                    raise ValueError(f"embedding {l}: shape {tuple(w.shape)}, expected {(rows, enc.n_feat)}")
                enc.table.data[int(enc.offs[l]):int(enc.offs[l + 1])] = w.to(dev, torch.float32)
            for h in _HEADS:
                head = getattr(self.backbone, h)
                if head.n_out == 0:
                    continue
                for j, (wn, bn) in enumerate((("w1", "b1"), ("w2", "b2"), ("w3", "b3"))):
                    getattr(head, wn).data.copy_(sd[f"backbone.{h}.{2 * j}.weight"])
                    getattr(head, bn).data.copy_(sd[f"backbone.{h}.{2 * j}.bias"])

    def reference_state_dict(self):
        """The inverse of load_reference_state_dict: CPU tensors under the reference's key names
        (with torch.compile's '_orig_mod.' level, as its train.py saves them), plus indices and
        empty_indices as train.py:365-369 adds them."""
        enc = self.backbone.encoding
        p = "backbone._orig_mod."
        sd = {"contracted_vertices": self.contracted_vertices.detach().cpu(), "ext_vertices": self.ext_vertices.cpu(),
              "center": self.center.cpu(), "scene_scaling": self.scene_scaling.cpu(),
              p + "encoding.level_resolutions": torch.tensor(enc.res.astype(np.float32))}
        for l in range(enc.n_levels):
            sd[p + f"encoding.embeddings.{l}.weight"] = enc.table.detach()[int(enc.offs[l]):int(enc.offs[l + 1])].cpu()
        for h in _HEADS:
            head = getattr(self.backbone, h)
            for j, (wn, bn) in enumerate((("w1", "b1"), ("w2", "b2"), ("w3", "b3"))):
                sd[p + f"{h}.{2 * j}.weight"] = getattr(head, wn).detach().cpu()
                sd[p + f"{h}.{2 * j}.bias"] = getattr(head, bn).detach().cpu()
        sd["indices"] = self.indices.cpu()
        sd["empty_indices"] = self.empty_indices.cpu()
        return sd

    @staticmethod
    def load_ckpt(path, device="cuda"):
        """models/ingp_color.py:452-474: config.json (the training arguments) and ckpt.pth from
        the directory `path`. The checkpoint is read with torch.load(weights_only=True): tensors
        only, nothing in the file is executed."""
        import json
        import os
        with open(os.path.join(path, "config.json")) as f:
            config = json.load(f)
        ckpt = torch.load(os.path.join(path, "ckpt.pth"), map_location="cpu", weights_only=True)
        indices = ckpt.pop("indices")
        empty = ckpt.pop("empty_indices", torch.empty((0, 4), dtype=indices.dtype))
        kw = {k: v for k, v in config.items() if k != "device"}
        model = Model(ckpt["contracted_vertices"], ckpt["ext_vertices"], ckpt["center"],
                      float(ckpt["scene_scaling"]), device=device, **kw)
        model.load_reference_state_dict(ckpt)
        model.min_t = float(model.scene_scaling) * float(config.get("base_min_t", 0.1))
        model.indices = torch.as_tensor(indices).to(device, torch.int32)
        model.empty_indices = torch.as_tensor(empty).to(device, torch.int32)
        return model

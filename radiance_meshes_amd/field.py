"""Hash-grid field and its MLP on the HIP kernels, as torch autograd functions.

* ``HashGrid``     utils/hashgrid.py:5-124 (HashEmbedderOptimized with conditional hashing on the
                   [0,1]^3 box, level-major output)
* ``FieldMLP``     a field head {32, 64} -> 64 -> 64 -> n_out (n_out <= 48) with SELU
                   (models/ingp_color.py:141-151), FP32 or BF16 operands on MFMA, FP32 accumulation
* ``contract``     utils/contraction.py:6-20 (contract_mean_std, position part), mapped to [0,1]^3
                   as models/ingp_color.py:347-349
* ``activate``     raw 13 outputs -> [density, base rgb, 3x3 colour gradient] features
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import timing
from ._capi import check, lib, ptr, stream_of


def level_table(n_levels: int, base_res: int, growth: float, log2_size: int):
    """Per-level resolution floor(base * growth^l) and level offsets into the concatenated table
    (sizes min(2^log2, res^3), utils/hashgrid.py:35-48)."""
    res = [int(math.floor(base_res * growth ** l)) for l in range(n_levels)]
    sizes = [min(2 ** log2_size, r ** 3) for r in res]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    return np.array(res, np.int32), offs


class _HashEncode(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, table, res, offs, n_levels, n_feat, log2_size):
        x = x.contiguous()
        N = x.shape[0]
        out = torch.empty((N, n_levels * n_feat), device=x.device, dtype=torch.float32)
        with timing.span("hashgrid_fwd"):
            check(lib().rm_hashgrid_fwd(ptr(x), N, ptr(table), res.ctypes.data, offs.ctypes.data, n_levels, n_feat,
                                    log2_size, ptr(out), stream_of(x.device)), "rm_hashgrid_fwd")
        ctx.save_for_backward(x)
        ctx.meta = (res, offs, n_levels, n_feat, log2_size, table.shape)
        return out

    @staticmethod
    def backward(ctx, gout):
        (x,) = ctx.saved_tensors
        res, offs, n_levels, n_feat, log2_size, tshape = ctx.meta
        gtable = torch.zeros(tshape, device=x.device, dtype=torch.float32)
        L = lib()
        n = x.shape[0]
        wsb = L.rm_hashgrid_bwd_workspace_size(n, res.ctypes.data, offs.ctypes.data, n_levels, n_feat, log2_size)
        ws = torch.empty((max(wsb, 1),), device=x.device, dtype=torch.uint8)
        with timing.span("hashgrid_bwd"):
            check(L.rm_hashgrid_bwd_ws(ptr(x), n, ptr(gout.contiguous()), res.ctypes.data, offs.ctypes.data,
                                       n_levels, n_feat, log2_size, ptr(gtable), ptr(ws), wsb, stream_of(x.device)),
                  "rm_hashgrid_bwd_ws")
        return None, gtable, None, None, None, None, None


class HashGrid(torch.nn.Module):
    def __init__(self, n_levels=16, n_features_per_level=2, log2_hashmap_size=19, base_resolution=16,
                 growth=1.38, device="cuda", seed=0, level_table=None):
        """level_table: (res [L] int32, offs [L+1] int64) to use instead of floor(base * growth^l)."""
        super().__init__()
        self.n_levels, self.n_feat, self.log2_size = n_levels, n_features_per_level, log2_hashmap_size
        if level_table is None:
            self.res, self.offs = globals()["level_table"](n_levels, base_resolution, growth, log2_hashmap_size)
        else:
            self.res = np.ascontiguousarray(level_table[0], np.int32)
            self.offs = np.ascontiguousarray(level_table[1], np.int64)
        g = torch.Generator(device=device).manual_seed(seed)
        init = (torch.rand((int(self.offs[-1]), n_features_per_level), generator=g, device=device) * 2 - 1) * 1e-4
        self.table = torch.nn.Parameter(init)
        self.out_dim = n_levels * n_features_per_level

    def forward(self, x):
        return _HashEncode.apply(x, self.table, self.res, self.offs, self.n_levels, self.n_feat, self.log2_size)


class _MLP(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, w1, b1, w2, b2, w3, b3, precision):
        x = x.contiguous()
        N = x.shape[0]
        n_out = w3.shape[0]
        y = torch.empty((N, n_out), device=x.device, dtype=torch.float32)
        ws = [t.contiguous() for t in (w1, b1, w2, b2, w3, b3)]
        with timing.span("mlp_fwd"):
            check(lib().rm_field_head_fwd(ptr(x), N, x.shape[1], *[ptr(t) for t in ws], n_out, precision, ptr(y),
                                          stream_of(x.device)), "rm_field_head_fwd")
        ctx.save_for_backward(x, *ws)
        ctx.precision = precision
        return y

    @staticmethod
    def backward(ctx, gy):
        x, w1, b1, w2, b2, w3, b3 = ctx.saved_tensors
        grads = [torch.zeros_like(t) for t in (w1, b1, w2, b2, w3, b3)]
        gx = torch.empty_like(x) if ctx.needs_input_grad[0] else None
        with timing.span("mlp_bwd"):
            check(lib().rm_field_head_bwd(ptr(x), x.shape[0], x.shape[1], ptr(w1), ptr(b1), ptr(w2), ptr(b2), ptr(w3),
                                          ptr(b3), w3.shape[0], ctx.precision, ptr(gy.contiguous()), ptr(gx),
                                          ptr(grads[0]), ptr(grads[1]), ptr(grads[2]), ptr(grads[3]), ptr(grads[4]),
                                          ptr(grads[5]), stream_of(x.device)), "rm_field_head_bwd")
        return (gx, *grads, None)


class FieldMLP(torch.nn.Module):
    """n_in -> 64 -> 64 -> n_out, SELU (n_in 32 or 64, n_out <= 48); weights ~ N(0, 1/fan_in), zero
    biases."""

    def __init__(self, n_in=32, hidden=64, n_out=13, precision="fp32", device="cuda", seed=0):
        super().__init__()
        if n_in not in (32, 64) or hidden != 64 or not 1 <= n_out <= 48:
            raise ValueError("the fused MLP kernels are built for {32, 64} -> 64 -> 64 -> n_out (n_out <= 48)")
        g = torch.Generator().manual_seed(seed)

        def w(o, i):
            return torch.nn.Parameter((torch.randn((o, i), generator=g) / math.sqrt(i)).to(device))

        self.w1, self.w2, self.w3 = w(hidden, n_in), w(hidden, hidden), w(n_out, hidden)
        self.b1 = torch.nn.Parameter(torch.zeros(hidden, device=device))
        self.b2 = torch.nn.Parameter(torch.zeros(hidden, device=device))
        self.b3 = torch.nn.Parameter(torch.zeros(n_out, device=device))
        self.precision = {"fp32": 0, "bf16": 1}[precision]

    def forward(self, x):
        return _MLP.apply(x, self.w1, self.b1, self.w2, self.b2, self.w3, self.b3, self.precision)


def contract(x, center=0.0, scale=1.0, eps=1.1920928955078125e-07):
    """contract_mean_std position part: |z| <= 1 kept, else (2 - 1/|z|) z/|z|; mapped to [0,1]^3."""
    z = (x - center) / scale
    mag_sq = (z * z).sum(-1, keepdim=True).clamp_min(eps)
    mag = mag_sq.sqrt()
    zc = torch.where(mag_sq <= 1, z, ((2 * mag - 1) / mag_sq) * z)
    return (zc / 2 + 1) / 2


def activate(raw, n_features=13):
    """raw MLP outputs -> per-tet features [density, base rgb, colour gradient rows]."""
    dens = torch.exp(raw[:, :1].clamp(max=math.log(1e20)))
    rgb = raw[:, 1:4] + 0.5
    if n_features == 13:
        return torch.cat([dens, rgb, raw[:, 4:13]], 1)
    return torch.cat([dens, rgb, raw[:, 4:7]], 1)


def field_points(vertices, indices, tet_ids, scale):
    """Query points of the listed tets: centroid / scale contracted into [0,1]^3 (rm_field_points;
    the same arithmetic as contract(vertices[indices[tet_ids]].mean(1), 0, scale)). No gradient."""
    vertices = vertices.detach().contiguous().float()
    out = torch.empty((tet_ids.shape[0], 3), device=vertices.device, dtype=torch.float32)
    with timing.span("field_glue"):
        check(lib().rm_field_points(ptr(indices.contiguous()), ptr(vertices), ptr(tet_ids), tet_ids.shape[0],
                                    float(scale), ptr(out), stream_of(vertices.device)), "rm_field_points")
    return out


class _FieldScatter(torch.autograd.Function):
    """activate(raw) written into rows tet_ids of a zero [n_tets, F] feature tensor
    (rm_field_scatter_fwd) and its adjoint (rm_field_scatter_bwd)."""

    @staticmethod
    def forward(ctx, raw, tet_ids, n_tets):
        raw = raw.contiguous()
        F = raw.shape[1]
        feats = torch.zeros((n_tets, F), device=raw.device, dtype=torch.float32)
        with timing.span("field_glue"):
            check(lib().rm_field_scatter_fwd(ptr(raw), ptr(tet_ids), raw.shape[0], F, ptr(feats),
                                             stream_of(raw.device)), "rm_field_scatter_fwd")
        ctx.save_for_backward(raw, tet_ids)
        return feats

    @staticmethod
    def backward(ctx, gfeats):
        raw, tet_ids = ctx.saved_tensors
        graw = torch.empty_like(raw)
        with timing.span("field_glue"):
            check(lib().rm_field_scatter_bwd(ptr(raw), ptr(tet_ids), raw.shape[0], raw.shape[1],
                                             ptr(gfeats.contiguous()), ptr(graw), stream_of(raw.device)),
                  "rm_field_scatter_bwd")
        return graw, None, None


class Field(torch.nn.Module):
    """Tet-centroid field: centroid -> contraction -> hash grid -> MLP -> features."""

    def __init__(self, n_features=13, precision="fp32", scale=2.0, device="cuda", **grid_kw):
        super().__init__()
        self.grid = HashGrid(device=device, **grid_kw)
        self.mlp = FieldMLP(n_in=self.grid.out_dim, n_out=n_features, precision=precision, device=device)
        self.scale = scale
        self.n_features = n_features

    def forward(self, vertices, indices, mask=None):
        idx = indices if mask is None else indices[mask]
        centroid = vertices[idx.long()].mean(dim=1)
        h = self.grid(contract(centroid, 0.0, self.scale).detach())
        return activate(self.mlp(h), self.n_features)

    def features(self, vertices, indices, mask):
        """Per-tet features [T, n_features], zero rows outside mask: the same values as
        forward() scattered into a [T, F] tensor, with the glue fused into three kernels."""
        tet_ids = torch.nonzero(mask).reshape(-1).to(torch.int32)
        x = field_points(vertices, indices, tet_ids, self.scale)
        raw = self.mlp(self.grid(x))
        return _FieldScatter.apply(raw, tet_ids, indices.shape[0])

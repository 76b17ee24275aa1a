"""Mixed-precision workload: the 10k-point plain mesh (uniform in [-1,1]^3 plus the cube corners,
about 65k tets, 256x256, 16x16 tiles) with the training step's field (hash grid L16 F2 2^19 base
16 growth 1.38, table U(-1,1) seed 3, MLP 32-64-64-13 with weights N(0, 1/fan_in)) evaluated at
every visible tet's centroid. The MLP runs with BF16 operands and FP32 accumulation on MFMA, and
with FP32 operands; both against an FP64 evaluation of the same weights on the same encodings.
Images: the GPU composites (FP32) the BF16 and the FP32 feature sets; the reference image is the
FP64 oracle's composite of the FP64 features. Reported: per-pixel max/mean error and the
density (relative) / colour-parameter (absolute) errors."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_bf16_field_to_image(oracle_lib):
    import torch
    import torch.nn.functional as Fn
    from radiance_meshes_amd import field as fieldmod, scenes
    from radiance_meshes_amd._capi import check, lib, ptr, stream_of
    from radiance_meshes_amd.rasterizer import RenderGrid, _cam, vertex_and_tile_shader
    O = oracle_lib
    sc = scenes.plain_small(tile_size=16, n_features=13)
    H, W = sc.camera.height, sc.camera.width
    print(f"mixed precision: plain mesh T={len(sc.indices)} {W}x{H}", flush=True)
    verts = torch.from_numpy(sc.vertices).cuda()
    idx = torch.from_numpy(sc.indices).cuda()
    tcam = {k: (v.cuda() if hasattr(v, "cuda") else v) for k, v in sc.camera.tcam(16, sc.min_t).items()}
    rg = RenderGrid(H, W, 16, 16)
    fld = fieldmod.Field(n_features=13, precision="fp32", scale=2.0, device="cuda")
    with torch.no_grad():
        fld.grid.table.copy_(torch.rand(fld.grid.table.shape, generator=torch.Generator().manual_seed(3)) * 2 - 1)
        sidx, ranges, vs, cc, mask, rect = vertex_and_tile_shader(idx, verts, tcam, rg)
        tet_ids = torch.nonzero(mask).reshape(-1).to(torch.int32)
        h = fld.grid(fieldmod.field_points(verts, idx, tet_ids, fld.scale))
        m = fld.mlp
        raw32 = m(h)
        m.precision = 1
        raw16 = m(h)
        m.precision = 0
        w = [t.detach().cpu().double() for t in (m.w1, m.b1, m.w2, m.b2, m.w3, m.b3)]
        z = Fn.selu(h.cpu().double() @ w[0].T + w[1])
        z = Fn.selu(z @ w[2].T + w[3])
        raw64 = (z @ w[4].T + w[5]).numpy()
    T = len(sc.indices)
    ids = tet_ids.cpu().numpy()

    def feats64(raw):
        f = np.zeros((T, 13))
        f[ids, 0] = np.exp(np.minimum(raw[:, 0], np.log(1e20)))
        f[ids, 1:4] = raw[:, 1:4] + 0.5
        f[ids, 4:] = raw[:, 4:13]
        return f

    cam = O.make_camera(sc.camera, 16, sc.min_t)
    jit = np.full((H, W, 2), 0.5, np.float32)
    ref, _, _, margin = O.composite_fwd(cam, sidx.cpu().numpy(), ranges.cpu().numpy(), sc.indices, sc.vertices,
                                        feats64(raw64), jit)
    out = {}
    for name, raw in (("bf16", raw16), ("fp32", raw32)):
        feats = torch.zeros((T, 13), device="cuda")
        feats[tet_ids.long()] = fieldmod.activate(raw, 13)
        img = torch.empty((H, W, 4), device="cuda")
        dist = torch.empty((H, W, 5), device="cuda")
        nc = torch.empty((H, W), device="cuda", dtype=torch.int32)
        check(lib().rm_composite_fwd(_cam(tcam, rg), ptr(sidx), sidx.numel(), ptr(ranges), ptr(idx), ptr(verts),
                                     ptr(feats), 13, ptr(torch.from_numpy(jit).cuda()), ptr(img), ptr(dist),
                                     ptr(nc), stream_of()), "rm_composite_fwd")
        torch.cuda.synchronize()
        r = raw.cpu().double().numpy()
        d_rel = np.abs(np.exp(r[:, 0]) - np.exp(raw64[:, 0])) / np.exp(raw64[:, 0])
        c_abs = np.abs(r[:, 1:] - raw64[:, 1:])
        e = np.abs(img.cpu().numpy()[..., :3] - ref[..., :3])
        keep = margin >= 1e-4
        out[name] = dict(density_rel_max=d_rel.max(), density_rel_mean=d_rel.mean(), colour_abs_max=c_abs.max(),
                         colour_abs_mean=c_abs.mean(), pixel_abs_max=e.max(), pixel_abs_mean=e.mean(),
                         pixel_abs_max_off_threshold=e[keep].max())
        print(f"{name}: " + ", ".join(f"{k} {v:.2e}" for k, v in out[name].items()), flush=True)
    assert out["fp32"]["density_rel_max"] <= 1e-5 and out["fp32"]["colour_abs_max"] <= 1e-5
    assert out["fp32"]["pixel_abs_max_off_threshold"] <= 1e-5
    assert out["bf16"]["density_rel_max"] <= 5e-2 and out["bf16"]["colour_abs_max"] <= 5e-2
    assert out["bf16"]["pixel_abs_max"] <= 5e-2 and out["bf16"]["pixel_abs_mean"] <= 5e-3

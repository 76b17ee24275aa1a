"""GPU parity of the per-tet error votes (rm_tet_err, render_err) against the FP64 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from tests._util import device_inputs, grid_of, scene, tcam_of

CASES = [("plain", 16, 7), ("plain", 4, 7), ("plain", 8, 13), ("ball", 16, 7), ("ball", 4, 13), ("ball", 8, 7)]


def _compare(O, sc, vals, ranges, feats64, pe, se, g_err, g_cnt, label):
    cam = O.make_camera(sc.camera, sc.tile_size, sc.min_t)
    o_err, o_cnt, margin, grazes = O.tet_err(cam, vals, ranges, sc.indices, sc.vertices, feats64, pe, se, graze=True)
    # pixels whose termination decision sits within FP32 noise of -5.5 may vote for a different
    # set of tets: every tet listed in such a pixel's tile is left out of the comparison
    amb = np.argwhere(margin < 1e-4)
    skip = np.zeros(len(sc.indices), bool)
    gw = (sc.camera.width + sc.tile_size - 1) // sc.tile_size
    for y, x in amb:
        t = (y // sc.tile_size) * gw + x // sc.tile_size
        skip[vals[ranges[t, 0]:ranges[t, 1]]] = True
    keep = ~skip
    assert o_cnt[:, 0].sum() > 100
    # a hit whose chord is within rounding of 0 (a ray through an edge) votes with weight ~0 in one
    # precision and not at all in another: such hits bound the count difference per tet
    dc = np.abs(g_cnt[keep, 0].astype(np.int64) - o_cnt[keep, 0])
    assert np.all(dc <= grazes[keep]), (int(dc.max()), int((dc > grazes[keep]).sum()))
    dmax = np.abs(g_cnt[keep, 1].astype(np.int64) - o_cnt[keep, 1])
    assert np.all(dmax <= 2 + 1e-4 * o_cnt[keep, 1]), dmax.max()
    worst = 0.0
    for ch in [0, 1, 2, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15]:
        a, b = g_err[keep, ch].astype(np.float64), o_err[keep, ch]
        scale = np.abs(b).max()
        e = np.abs(a - b).max() / max(scale, 1e-30)
        worst = max(worst, e)
        assert e <= 2e-5, (ch, e)
    assert np.all(g_err[:, 3:5] == 0)
    print(f"{label}: max channel err / channel scale = {worst:.2e}, voted tets {int((o_cnt[:, 0] > 0).sum())}, "
          f"near-threshold pixels {len(amb)}, tets left out {int(skip.sum())}, "
          f"count differences {int(dc.sum())} (grazing hits {int(grazes.sum())})")


@pytest.mark.parametrize("kind,tile,nf", CASES)
def test_tet_err_matches_fp64_oracle(oracle_lib, kind, tile, nf):
    import torch
    from radiance_meshes_amd.rasterizer import vertex_and_tile_shader
    from radiance_meshes_amd.render_err import calc_tet_err
    O = oracle_lib
    sc = scene(kind, tile, nf)
    print(f"case {kind} tile={tile} nf={nf} T={len(sc.indices)}", flush=True)
    verts, idx, feats, _ = device_inputs(sc)
    tcam, rg = tcam_of(sc), grid_of(sc)
    H, W = sc.camera.height, sc.camera.width
    rng = np.random.default_rng(3)
    pe = rng.uniform(-0.5, 0.5, (H, W)).astype(np.float32)
    se = rng.uniform(0, 1, (H, W)).astype(np.float32)
    sidx, ranges, vs, cc, mask, rect = vertex_and_tile_shader(idx, verts, tcam, rg)
    err, cnt = calc_tet_err(sidx, ranges, idx, verts, feats, rg, tcam, torch.from_numpy(pe).cuda(),
                            torch.from_numpy(se).cuda())
    torch.cuda.synchronize()
    _compare(O, sc, sidx.cpu().numpy(), ranges.cpu().numpy(), sc.features.astype(np.float64), pe, se,
             err.cpu().numpy(), cnt.cpu().numpy(), f"{kind} {tile} {nf}")


def test_tet_err_accumulates_over_views(oracle_lib):
    """Votes of two calls add into the same buffers (counts add, maxima combine)."""
    import torch
    from radiance_meshes_amd.rasterizer import vertex_and_tile_shader
    from radiance_meshes_amd.render_err import calc_tet_err
    sc = scene("ball", 4, 7)
    verts, idx, feats, _ = device_inputs(sc)
    tcam, rg = tcam_of(sc), grid_of(sc)
    H, W = sc.camera.height, sc.camera.width
    pe = torch.full((H, W), 0.3, device="cuda")
    se = torch.full((H, W), 0.1, device="cuda")
    sidx, ranges, *_ = vertex_and_tile_shader(idx, verts, tcam, rg)
    e1, c1 = calc_tet_err(sidx, ranges, idx, verts, feats, rg, tcam, pe, se)
    e2, c2 = calc_tet_err(sidx, ranges, idx, verts, feats, rg, tcam, pe, se, e1.clone(), c1.clone())
    torch.cuda.synchronize()
    np.testing.assert_allclose(e2.cpu().numpy(), 2 * e1.cpu().numpy(), rtol=1e-5, atol=1e-6)
    np.testing.assert_array_equal(c2[:, 0].cpu().numpy(), 2 * c1[:, 0].cpu().numpy())
    np.testing.assert_array_equal(c2[:, 1].cpu().numpy(), c1[:, 1].cpu().numpy())


class _Camera:
    def __init__(self, cam, tile):
        import torch
        self.c, self.tile = cam, tile
        self.image_height, self.image_width = cam.height, cam.width
        m = torch.ones((1, cam.height, cam.width))
        m[:, :, : cam.width // 5] = 0     # part of the view masked out
        self.gt_alpha_mask = m

    def to_dict(self, device):
        t = self.c.tcam(self.tile, 0.1)
        keep = ("world_view_transform", "cam_pos", "K", "image_height", "image_width", "fovy", "fovx",
                "distortion_params", "camera_type")
        return {k: (v.to(device) if hasattr(v, "to") else v) for k, v in t.items() if k in keep}


class _Model:
    def __init__(self, sc):
        import torch
        self.device = "cuda"
        self.vertices = torch.from_numpy(sc.vertices).cuda()
        self.indices = torch.from_numpy(sc.indices).cuda()
        self.feats = torch.from_numpy(sc.features).cuda()
        self.feature_dim = sc.features.shape[1]

    def get_cell_values(self, camera, mask=None, all_circumcenters=None):
        return None, self.feats[mask]


def test_render_err_through_reference_path(oracle_lib):
    """render_err() as densification calls it (utils/densification.py:136-160), through the
    reference's module path: images, per-pixel errors and votes against the FP64 oracle."""
    import torch
    import radiance_meshes_amd
    radiance_meshes_amd.install_reference_aliases()
    from delaunay_rasterization.internal.render_err import render_err
    O = oracle_lib
    sc = scene("ball", 4, 7)
    print(f"render_err T={len(sc.indices)}", flush=True)
    H, W = sc.camera.height, sc.camera.width
    gt = torch.rand((3, H, W), generator=torch.Generator().manual_seed(4)).cuda()
    cam, model = _Camera(sc.camera, 4), _Model(sc)
    tet_err, extras = render_err(gt, cam, model, tile_size=4)
    torch.cuda.synchronize()
    assert set(extras) == {"tet_count", "pixel_err", "ssim_err", "render_img", "cell_values"}
    ocam = O.make_camera(sc.camera, 4, 0.1)
    t, r, vs, cc = O.project(ocam, sc.indices, sc.vertices)
    keys, vals, ranges = O.bin_tiles(ocam, t, r, vs)
    f = sc.features.astype(np.float64) * (t > 0)[:, None]
    img, dist, nc, _ = O.composite_fwd(ocam, vals, ranges, sc.indices, sc.vertices, f, np.full((H, W, 2), 0.5, np.float32))
    mask = cam.gt_alpha_mask.numpy()
    ref_img = np.clip(img[..., :3].transpose(2, 0, 1), 0, 1) * mask
    assert np.abs(extras["render_img"].cpu().numpy() - ref_img).max() <= 1e-5
    ref_pe = (np.clip(img[..., :3].transpose(2, 0, 1), 0, 1) - gt.cpu().numpy()).mean(0) * mask[0]
    assert np.abs(extras["pixel_err"].cpu().numpy() - ref_pe).max() <= 1e-5
    se = extras["ssim_err"].cpu().numpy()
    assert se.min() >= 0 and se.max() <= 1 and np.all(se[:, : W // 5] == 0)
    _compare(O, sc, vals, ranges, f, extras["pixel_err"].cpu().numpy().astype(np.float64), se.astype(np.float64),
             tet_err.cpu().numpy(), extras["tet_count"].cpu().numpy(), "render_err")

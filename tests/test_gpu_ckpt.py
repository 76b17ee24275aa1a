"""A checkpoint in the reference's on-disk layout (ckpt.pth state dict + indices, config.json;
train.py:365-369, models/ingp_color.py:452-474) loaded with Model.load_ckpt and rendered through
render() on the GPU, against the FP64 oracle chain (query, field, activate_output, compositing)."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CONFIG = dict(L=8, hashmap_dim=8, log2_hashmap_size=14, base_resolution=16, per_level_scale=2, scale_multi=0.35,
              hidden_dim=64, max_sh_deg=3, current_sh_deg=3, ablate_circumsphere=True, density_offset=-4,
              base_min_t=0.05, iterations=30000, lr=0.01)


class _Camera:
    """data/camera.Camera's surface as render() and get_cell_values use it."""

    def __init__(self, cam, tile):
        import torch
        self.c, self.tile = cam, tile
        self.image_height, self.image_width = cam.height, cam.width
        self.gt_alpha_mask = torch.ones((1, cam.height, cam.width))
        self.camera_center = torch.as_tensor(np.asarray(cam.cam_pos, np.float32))

    def to_dict(self, device):
        t = self.c.tcam(self.tile, 0.1)
        keep = ("world_view_transform", "cam_pos", "K", "image_height", "image_width", "fovy", "fovx",
                "distortion_params", "camera_type")
        return {k: (v.to(device) if hasattr(v, "to") else v) for k, v in t.items() if k in keep}


def _write_checkpoint(tmp_path):
    """A trained-looking model in the reference's layout: Delaunay mesh of 1,500 interior points plus
    a ring of 60 exterior vertices, table U(-1, 1), head weights N(0, 1/fan_in)."""
    import torch
    from radiance_meshes_amd import scenes
    from radiance_meshes_amd.ingp import Model
    rng = np.random.default_rng(21)
    pts = rng.uniform(-1, 1, (4000, 3))
    pts = pts[np.linalg.norm(pts, axis=1) <= 1][:1500].astype(np.float32)
    phi = np.arange(60) * np.pi * (3 - np.sqrt(5))
    z = 1 - 2 * (np.arange(60) + 0.5) / 60
    ext = (1.6 * np.stack([np.sqrt(1 - z * z) * np.cos(phi), np.sqrt(1 - z * z) * np.sin(phi), z], 1)).astype(np.float32)
    tets = scenes.tetrahedralize(np.concatenate([pts, ext]).astype(np.float64))
    m = Model(pts, ext, np.array([0.05, -0.02, 0.01]), 1.3, device="cuda", **CONFIG)
    m.indices = torch.from_numpy(tets.astype(np.int32)).cuda()
    g = torch.Generator(device="cuda").manual_seed(4)
    with torch.no_grad():
        t = m.backbone.encoding.table
        t.copy_(torch.rand(t.shape, generator=g, device="cuda") * 2 - 1)
        for h in (m.backbone.density_net, m.backbone.color_net, m.backbone.gradient_net, m.backbone.sh_net):
            for w in (h.w1, h.w2, h.w3):
                w.copy_(torch.randn(w.shape, generator=g, device="cuda") / np.sqrt(w.shape[1]))
            for b in (h.b1, h.b2, h.b3):
                b.copy_(torch.randn(b.shape, generator=g, device="cuda") * 0.1)
    torch.save(m.reference_state_dict(), tmp_path / "ckpt.pth")
    (tmp_path / "config.json").write_text(json.dumps(CONFIG))
    return m


def test_checkpoint_round_trip_and_render(oracle_lib, tmp_path):
    import torch
    from radiance_meshes_amd.ingp import Model
    from radiance_meshes_amd.render import render
    from radiance_meshes_amd import scenes
    O = oracle_lib
    print("writing a reference-layout checkpoint", flush=True)
    src = _write_checkpoint(tmp_path)
    sd = torch.load(tmp_path / "ckpt.pth", map_location="cpu", weights_only=True)
    assert "backbone._orig_mod.encoding.embeddings.7.weight" in sd and "indices" in sd
    m = Model.load_ckpt(tmp_path, "cuda")
    assert torch.equal(m.backbone.encoding.table, src.backbone.encoding.table)
    assert torch.equal(m.backbone.sh_net.w3, src.backbone.sh_net.w3)
    assert torch.equal(m.vertices, src.vertices) and torch.equal(m.indices, src.indices)
    assert m.backbone.density_offset == -4 and abs(m.min_t - 1.3 * 0.05) < 1e-7
    H = W = 48
    cam = scenes.look_at_camera((0.4, 0.5, 2.4), (0, 0, 0), (0, 1, 0), H, W, 50.0)
    camera = _Camera(cam, 4)
    print(f"render of the loaded checkpoint: {m.indices.shape[0]} tets, {H}x{W}", flush=True)
    with torch.no_grad():
        pkg = render(camera, m, tile_size=4, min_t=0.1)
        shs, feats = m.get_cell_values(camera)
    # FP64 chain from the GPU's float32 query coordinates (rm_field_query, checked against the FP64
    # query in test_gpu_ingp.py; a 1-ulp change of a coordinate moves a fine level of a random table
    # by ~res * 6e-8, so the chain starts from the same coordinates): field in FP64 (SH rounded to
    # FP16 by the GPU, as the reference rounds it, and taken from there), activate_output, then the
    # compositor on the projection's mask
    from radiance_meshes_amd.ingp import field_query
    V = m.vertices.detach().cpu().numpy().astype(np.float64)
    idx = m.indices.cpu().numpy()
    xg, crg, _ = field_query(m.vertices, m.indices, m.center.cpu().numpy()[0], float(m.scene_scaling))
    x, cr = xg.cpu().numpy(), crg.cpu().numpy()
    rx, rcr, _ = O.field_query(V, idx, m.center.cpu().numpy()[0], float(m.scene_scaling), "centroid")
    assert np.abs(x - rx).max() <= 1e-6 and (np.abs(cr - rcr) / rcr).max() <= 2e-6
    enc = m.backbone.encoding
    heads = [[getattr(h, k).detach().cpu().numpy() for k in ("w1", "b1", "w2", "b2", "w3", "b3")]
             for h in (m.backbone.density_net, m.backbone.color_net, m.backbone.gradient_net, m.backbone.sh_net)]
    table = enc.table.detach().cpu().numpy().astype(np.float64)
    dens, rgb, grd, _ = O.ingp_forward(x, cr, table, enc.res, enc.offs, 8, 14, heads, 0.35, 2, -4)
    sh16 = shs.cpu().numpy().astype(np.float64)
    of = O.activate_output(cam.cam_pos, dens, rgb, grd, sh16, idx, V[idx].mean(1), V, 3)
    fe = np.abs(feats.cpu().numpy() - of).max(0) / np.maximum(np.abs(of).max(0), 1e-6)
    print("features max err / max |ref| per channel", np.array2string(fe, precision=2))
    assert fe.max() <= 2e-5
    ocam = O.make_camera(cam, 4, 0.1)
    t, r, vs, cc = O.project(ocam, idx, m.vertices.detach().cpu().numpy())
    keys, vals, ranges = O.bin_tiles(ocam, t, r, vs)
    img, dist, nc, margin = O.composite_fwd(ocam, vals, ranges, idx, V, of * (t > 0)[:, None],
                                            np.full((H, W, 2), 0.5, np.float32))
    far = margin >= 1e-4
    err = np.abs(pkg["render"].cpu().numpy() - img[..., :3].transpose(2, 0, 1))[:, far].max()
    print(f"rendered checkpoint: image max-abs err vs FP64 {err:.3e} ({int((~far).sum())} near-threshold pixels)")
    assert err <= 2e-5
    assert (img[..., 3] < -1e-3).mean() > 0.2     # the view sees the mesh

"""render() as train.py calls it (utils/train_util.py:15-77), through the reference's module
paths, with a camera/model shaped like data/camera.Camera and models/ingp_color.Model."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Camera:
    def __init__(self, cam):
        import torch
        self.c = cam
        self.image_height, self.image_width = cam.height, cam.width
        self.gt_alpha_mask = torch.ones((1, cam.height, cam.width))

    def to_dict(self, device):
        t = self.c.tcam(4, 0.1)
        keep = ("world_view_transform", "cam_pos", "K", "image_height", "image_width", "fovy", "fovx",
                "distortion_params", "camera_type")
        return {k: (v.to(device) if hasattr(v, "to") else v) for k, v in t.items() if k in keep}


class _Model:
    def __init__(self, sc):
        import torch
        self.device = "cuda"
        self.vertices = torch.from_numpy(sc.vertices).cuda().requires_grad_(True)
        self.indices = torch.from_numpy(sc.indices).cuda()
        self.feats = torch.from_numpy(sc.features).cuda().requires_grad_(True)
        self.feature_dim = 7
        self.mask_values = True

    def get_cell_values(self, camera, mask=None, all_circumcenters=None):
        import torch
        return torch.zeros((int(mask.sum()), 0, 3), device="cuda"), self.feats[mask]


def test_render_through_reference_paths(oracle_lib):
    import torch
    import radiance_meshes_amd
    radiance_meshes_amd.install_reference_aliases()
    from radiance_meshes_amd.render import render
    from radiance_meshes_amd import scenes
    O = oracle_lib
    sc = scenes.ball_scene(n_points=400, seed=2, size=48, tile_size=4, n_features=7, dtype=np.float32)
    cam, model = _Camera(sc.camera), _Model(sc)
    print("render() through aliases", flush=True)
    pkg = render(cam, model, tile_size=4, min_t=0.1)
    loss = pkg["render"].sum() + pkg["distortion_loss"] + pkg["alpha"].mean()
    loss.backward()
    torch.cuda.synchronize()
    assert set(pkg) == {"render", "alpha", "distortion_loss", "mask", "sh_reg"}
    assert pkg["render"].shape == (3, 48, 48)
    ocam = O.make_camera(sc.camera, 4, 0.1)
    t, r, vs, cc = O.project(ocam, sc.indices, sc.vertices)
    keys, vals, ranges = O.bin_tiles(ocam, t, r, vs)
    img, dist, nc, _ = O.composite_fwd(ocam, vals, ranges, sc.indices, sc.vertices,
                                       sc.features * (t > 0)[:, None], np.full((48, 48, 2), 0.5, np.float32))
    err = np.abs(pkg["render"].detach().cpu().numpy() - img[..., :3].transpose(2, 0, 1)).max()
    print(f"render() image err vs FP64 oracle {err:.3e}")
    assert err <= 1e-5
    assert model.feats.grad is not None and torch.isfinite(model.feats.grad).all()
    assert model.vertices.grad is not None and torch.isfinite(model.vertices.grad).all()


def test_render_mesh_loaded_from_ply(oracle_lib, tmp_path):
    """A mesh saved in the reference's PLY layout, loaded, turned into per-view features (SH up to
    degree 3) and rendered on the GPU; image against the FP64 oracle on the same features."""
    import torch
    from radiance_meshes_amd import meshio, scenes
    from radiance_meshes_amd.rasterizer import AlphaBlendTiledRender, RenderGrid, vertex_and_tile_shader
    O = oracle_lib
    sc = scenes.ball_scene(n_points=400, seed=3, size=40, tile_size=4, n_features=7, dtype=np.float32)
    rng = np.random.default_rng(0)
    T = len(sc.indices)
    p = tmp_path / "ckpt.ply"
    meshio.save_mesh_ply(p, sc.vertices, sc.indices, sc.features[:, 0], sc.features[:, 1:4], sc.features[:, 4:7],
                         rng.normal(0, 0.1, (T, 15, 3)))
    mesh = meshio.load_mesh_ply(p)
    feats = meshio.view_features(mesh, sc.camera.cam_pos, sh_deg=3)
    verts = torch.from_numpy(mesh["vertices"]).cuda()
    idx = torch.from_numpy(mesh["indices"]).cuda()
    H = W = 40
    tcam = {k: (v.cuda() if hasattr(v, "cuda") else v) for k, v in sc.camera.tcam(4, sc.min_t).items()}
    rg = RenderGrid(H, W, 4, 4)
    print(f"PLY mesh T={T}", flush=True)
    sidx, ranges, *_ = vertex_and_tile_shader(idx, verts, tcam, rg)
    img, _, _ = AlphaBlendTiledRender.apply(sidx, ranges, idx, verts, feats, rg, tcam, torch.full((H, W, 2), 0.5).cuda())
    torch.cuda.synchronize()
    cam = O.make_camera(sc.camera, 4, sc.min_t)
    o_img, _, _, _ = O.composite_fwd(cam, sidx.cpu().numpy(), ranges.cpu().numpy(), mesh["indices"], mesh["vertices"],
                                     feats.cpu().numpy().astype(np.float64), np.full((H, W, 2), 0.5, np.float32))
    err = np.abs(img.cpu().numpy()[..., :3] - o_img[..., :3]).max()
    print(f"PLY-loaded mesh: image err vs FP64 oracle {err:.2e}")
    assert err <= 1e-5


class _CircumModel(_Model):
    """Cell values that depend on the circumcenters, as the reference's field does when
    ablate_circumsphere is off (models/ingp_color.py:320-325: the circumcenter reaches the field
    through the query point and offset_normalize): values = features + cc @ A."""

    def __init__(self, sc, A):
        import torch
        super().__init__(sc)
        self.A = torch.from_numpy(A).float().cuda()

    def get_cell_values(self, camera, mask=None, all_circumcenters=None):
        import torch
        return torch.zeros((int(mask.sum()), 0, 3), device="cuda"), self.feats[mask] + all_circumcenters @ self.A


def test_render_circumcenter_gradient_chain(oracle_lib):
    """render() with circumcenter-dependent cell values: the vertex gradient is the compositing
    adjoint plus the projection adjoint of d loss / d circumcenter (VertexShader.backward);
    checked against the FP64 oracle chain (orc_composite_bwd, then orc_power_bwd per tet)."""
    import torch
    from radiance_meshes_amd.render import render
    from radiance_meshes_amd import scenes
    O = oracle_lib
    sc = scenes.ball_scene(n_points=400, seed=4, size=48, tile_size=4, n_features=7, dtype=np.float32)
    A = np.random.default_rng(5).normal(0, 0.05, (3, 7))
    cam, model = _Camera(sc.camera), _CircumModel(sc, A)
    H = W = 48
    Wt = np.random.default_rng(6).standard_normal((3, H, W))
    print("render() with circumcenter-dependent values", flush=True)
    pkg = render(cam, model, tile_size=4, min_t=0.1)
    (pkg["render"] * torch.from_numpy(Wt).float().cuda()).sum().backward()
    torch.cuda.synchronize()
    ocam = O.make_camera(sc.camera, 4, 0.1)
    t, r, vs, cc = O.project(ocam, sc.indices, sc.vertices)
    keys, vals, ranges = O.bin_tiles(ocam, t, r, vs)
    mask = t > 0
    f64 = (sc.features.astype(np.float64) + cc.astype(np.float64) @ A) * mask[:, None]
    jit = np.full((H, W, 2), 0.5, np.float32)
    img, dist, nc, margin = O.composite_fwd(ocam, vals, ranges, sc.indices, sc.vertices, f64, jit)
    err = np.abs(pkg["render"].detach().cpu().numpy() - img[..., :3].transpose(2, 0, 1)).max()
    print(f"image err vs FP64 oracle {err:.3e}")
    assert err <= 1e-5
    gimg = np.zeros((H, W, 4))
    gimg[..., :3] = Wt.transpose(1, 2, 0)
    gv, gf = O.composite_bwd(ocam, vals, ranges, sc.indices, sc.vertices, f64, jit, img, dist, nc, gimg,
                             np.zeros((H, W, 5)))
    gcc = (gf @ A.T) * mask[:, None]
    gv = gv + O.project_bwd(ocam, sc.indices, sc.vertices.astype(np.float64), t, np.zeros((len(t), 3)), gcc)
    g = model.vertices.grad.cpu().numpy().astype(np.float64)
    rel = np.linalg.norm(g - gv) / np.linalg.norm(gv)
    gproj = np.abs(gv - O.composite_bwd(ocam, vals, ranges, sc.indices, sc.vertices, f64, jit, img, dist, nc,
                                        gimg, np.zeros((H, W, 5)))[0]).max()
    print(f"vertex grad rel L2 err vs FP64 chain {rel:.2e} (projection part up to {gproj:.3e})")
    assert gproj > 0
    assert rel <= 1e-3

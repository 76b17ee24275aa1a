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

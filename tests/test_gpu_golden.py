"""GPU path on the golden fixtures (restaged reference test scenes): bit-exact projection and
binning, image within 1e-5 of the FP64 oracle, gradients within 1e-3 relative."""
import numpy as np
import pytest

from tests._golden import load

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["single_tet", "multi_tet"])
def test_gpu_matches_golden(name):
    import torch
    from radiance_meshes_amd.rasterizer import AlphaBlendTiledRender, RenderGrid, VertexShader, vertex_and_tile_shader
    z = load(name)
    print(f"golden {name}", flush=True)
    dev = "cuda"
    verts = torch.from_numpy(z["vertices"]).to(dev).requires_grad_(True)
    idx = torch.from_numpy(z["indices"]).to(dev)
    feats = torch.from_numpy(z["features"]).to(dev).requires_grad_(True)
    cam = z["camera"]
    tcam = {k: (v.to(dev) if hasattr(v, "to") else v) for k, v in cam.tcam(z["tile"], z["min_t"]).items()}
    rg = RenderGrid(cam.height, cam.width, z["tile"], z["tile"])
    tiles, rect, vs, cc = VertexShader.apply(idx, verts, tcam, rg)
    np.testing.assert_array_equal(tiles.cpu().numpy(), z["tiles"])
    np.testing.assert_array_equal(rect.cpu().numpy(), z["rect"])
    np.testing.assert_array_equal(vs[:, 2].detach().cpu().numpy().view(np.int32), z["depth"].view(np.int32))
    sidx, ranges, vs, cc, mask, rect = vertex_and_tile_shader(idx, verts, tcam, rg)
    np.testing.assert_array_equal(sidx.cpu().numpy(), z["sorted_idx"])
    np.testing.assert_array_equal(ranges.cpu().numpy(), z["ranges"])
    jit = torch.from_numpy(z["jitter"]).to(dev)
    img, dist, _ = AlphaBlendTiledRender.apply(sidx, ranges, idx, verts, feats, rg, tcam, jit)
    loss = (img * torch.from_numpy(z["w_img"]).float().to(dev)).sum() + \
        (dist * torch.from_numpy(z["w_dist"]).float().to(dev)).sum()
    loss.backward()
    err = np.abs(img.detach().cpu().numpy() - z["img"]).max()
    print(f"image max-abs err {err:.3e}")
    assert err <= 1e-5 * max(1.0, np.abs(z["img"]).max())
    for got, ref, what in [(verts.grad, z["grad_vertices"], "vertices"), (feats.grad, z["grad_features"], "features")]:
        rel = np.linalg.norm(got.cpu().numpy() - ref) / np.linalg.norm(ref)
        print(f"grad {what} rel err {rel:.3e}")
        assert rel <= 1e-3

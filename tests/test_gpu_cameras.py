"""GPU parity for the other camera models the reference supports (camera.slang:241-346,357-508):
a distorted pinhole (k1, k2, p1, p2), the fisheye (camera_type 1), a camera inside the mesh, and
a view with nothing in it (empty tile lists)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from tests._util import device_inputs, f32_bits, grid_of, jitter, scene, tcam_of, with_camera

CASES = [("distorted", 4, 7), ("distorted", 8, 13), ("fisheye", 4, 13), ("fisheye", 16, 7), ("inside", 4, 13),
         ("inside", 8, 7)]


@pytest.mark.parametrize("kind,tile,nf", CASES)
def test_camera_models_match_oracle(oracle_lib, kind, tile, nf):
    import torch
    from radiance_meshes_amd.rasterizer import AlphaBlendTiledRender, VertexShader, vertex_and_tile_shader
    O = oracle_lib
    sc = with_camera(scene("ball", tile, nf), kind)
    print(f"case {kind} tile={tile} nf={nf}", flush=True)
    verts, idx, feats, _ = device_inputs(sc)
    tcam, rg = tcam_of(sc), grid_of(sc)
    cam = O.make_camera(sc.camera, sc.tile_size, sc.min_t)
    # projection and binning bit-exact
    tiles, rect, vs, cc = VertexShader.apply(idx, verts, tcam, rg)
    o_tiles, o_rect, o_vs, o_cc = O.project(cam, sc.indices, sc.vertices)
    assert (o_tiles > 0).sum() > 100
    np.testing.assert_array_equal(tiles.cpu().numpy(), o_tiles)
    np.testing.assert_array_equal(rect.cpu().numpy(), o_rect)
    np.testing.assert_array_equal(f32_bits(vs.cpu().numpy()), f32_bits(o_vs))
    o_keys, o_vals, o_ranges = O.bin_tiles(cam, o_tiles, o_rect, o_vs)
    verts.requires_grad_(True)
    feats.requires_grad_(True)
    sidx, ranges, *_ = vertex_and_tile_shader(idx, verts, tcam, rg)
    np.testing.assert_array_equal(sidx.cpu().numpy(), o_vals)
    np.testing.assert_array_equal(ranges.cpu().numpy(), o_ranges)
    # compositing forward and backward against the FP64 oracle
    jit = torch.from_numpy(jitter(sc, 3)).cuda()
    img, dist, _ = AlphaBlendTiledRender.apply(sidx, ranges, idx, verts, feats, rg, tcam, jit)
    H, W = sc.camera.height, sc.camera.width
    rng = np.random.default_rng(8)
    w_img = rng.standard_normal((H, W, 4))
    w_dist = rng.standard_normal((H, W, 5)) * 0.1
    ((img * torch.from_numpy(w_img).float().cuda()).sum() + (dist * torch.from_numpy(w_dist).float().cuda()).sum()).backward()
    torch.cuda.synchronize()
    o_img, o_dist, o_nc, margin = O.composite_fwd(cam, o_vals, o_ranges, sc.indices, sc.vertices, sc.features,
                                                  jitter(sc, 3))
    keep = margin >= 1e-4
    err = np.abs(img.detach().cpu().numpy().astype(np.float64) - o_img)[keep][:, :3].max()
    print(f"{kind}: rgb max-abs err vs FP64 {err:.2e}, near-threshold pixels {int((~keep).sum())}")
    assert err <= 1e-5
    gv, gf = O.composite_bwd(cam, o_vals, o_ranges, sc.indices, sc.vertices, sc.features, jitter(sc, 3), o_img,
                             o_dist, o_nc, w_img, w_dist)
    for name, a, b in (("vertex", verts.grad, gv), ("feature", feats.grad, gf)):
        a = a.cpu().numpy().astype(np.float64)
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(f"{kind}: {name} gradient rel L2 err vs FP64 adjoint {rel:.2e}")
        assert rel <= 1e-3


def test_view_with_nothing_in_it(oracle_lib):
    """A camera looking away from the mesh: every tile list is empty; the image is the background
    (colour 0, log-transmittance 0), the distortion image 0, and the gradients are zero."""
    import torch
    from radiance_meshes_amd.rasterizer import AlphaBlendTiledRender, vertex_and_tile_shader
    sc = with_camera(scene("ball", 4, 13), "away")
    verts, idx, feats, jit = device_inputs(sc)
    verts.requires_grad_(True)
    feats.requires_grad_(True)
    tcam, rg = tcam_of(sc), grid_of(sc)
    sidx, ranges, vs, cc, mask, rect = vertex_and_tile_shader(idx, verts, tcam, rg)
    assert sidx.numel() == 0 and int(mask.sum()) == 0
    assert int(ranges.abs().sum()) == 0
    img, dist, alive = AlphaBlendTiledRender.apply(sidx, ranges, idx, verts, feats, rg, tcam, jit)
    (img.sum() + dist.sum()).backward()
    torch.cuda.synchronize()
    assert float(img.abs().max()) == 0.0 and float(dist.abs().max()) == 0.0
    assert float(verts.grad.abs().max()) == 0.0 and float(feats.grad.abs().max()) == 0.0

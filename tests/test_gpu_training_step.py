"""The training-step workload at full size against the FP64 oracle.

Workload (bench.py's): 1,000,000 points of the seeded 32-component Gaussian mixture plus 8 corners
at +-20, tetrahedralised (here on the GPU with rm_delaunay, whose parity with scipy is covered by
test_gpu_delaunay.py) and laid out along a Morton curve; 1280x840 orbit view (radius 6, 50 deg),
4x4 tiles, min_t 0.1; field: hash grid L16 F2 2^19 base 16 growth 1.38 at the visible tets'
centroids, MLP 32-64-64-13 in FP32, features through Field.features (13 channels). The hash table
is drawn U(-1, 1) (seed 3) so that the features vary across the mesh.

Checked here:
* projection (tiles, rect, depth, circumcenter) and binning (sorted tet list, ranges) bit-exact
  against orc_project / orc_bin over every tet;
* the image on 64 seeded random tiles within 1e-5 (max-abs, rgb) of orc_composite_fwd on the
  same features; pixels whose termination decision is within 1e-4 of the -5.5 threshold are
  counted (at most 2% of the sampled pixels) and bounded by the transmittance left at the
  threshold times the largest colour magnitude in their tiles;
* the loss sum(img[..., :3] * Wt), Wt ~ N(0, 1) on the sampled tiles' pixels (0 elsewhere and at
  near-threshold pixels), backpropagated through compositing, the feature scatter, the MLP and
  the hash grid; feature, vertex, MLP-weight and hash-table gradients against the FP64 chain
  (orc_composite_bwd on the sampled tiles, then the activation, MLP and hash-grid adjoints in
  numpy FP64 on the GPU's query points): per-element relative error <= 1e-3 with a floor of
  1e-4 of the group's largest gradient (MLP weights: or within the FP32 summation bound of the
  element's terms, for elements that cancel).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H, W, TILE, N_TILES = 840, 1280, 4, 64


def _selu_prime(h):
    from oracle.oracle import SELU_ALPHA, SELU_SCALE
    return np.where(h > 0, SELU_SCALE, SELU_SCALE * SELU_ALPHA * np.exp(np.minimum(h, 0)))


def _rel(a, b, floor_frac=1e-4):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    floor = floor_frac * np.abs(b).max()
    return np.abs(a - b) / (np.abs(b) + floor)


@pytest.fixture(scope="module")
def step(oracle_lib):
    import torch
    from radiance_meshes_amd import field as fieldmod, scenes
    from radiance_meshes_amd.delaunay import delaunay_tets
    from radiance_meshes_amd.layout import spatial_layout
    from radiance_meshes_amd.rasterizer import AlphaBlendTiledRender, RenderGrid, vertex_and_tile_shader
    O = oracle_lib
    print("training-step workload: 1M GMM points, GPU Delaunay", flush=True)
    pts = scenes.gmm_points(1_000_000, 0).astype(np.float32)
    tets = delaunay_tets(torch.from_numpy(pts).cuda()).cpu().numpy()
    pts, tets, _, _ = spatial_layout(pts, tets)
    T = len(tets)
    print(f"mesh: {len(pts)} points, {T} tets", flush=True)
    cam = scenes.orbit_camera(6.0, 30.0, 20.0, W, H, 50.0)
    tcam = {k: (v.cuda() if hasattr(v, "cuda") else v) for k, v in cam.tcam(TILE, 0.1).items()}
    rg = RenderGrid(H, W, TILE, TILE)
    fld = fieldmod.Field(n_features=13, precision="fp32", scale=2.0, device="cuda")
    with torch.no_grad():
        fld.grid.table.copy_(torch.rand(fld.grid.table.shape, generator=torch.Generator().manual_seed(3)) * 2 - 1)
    vertices = torch.nn.Parameter(torch.from_numpy(pts).cuda())
    indices = torch.from_numpy(tets).cuda()
    jitter = torch.full((H, W, 2), 0.5, device="cuda")

    sidx, ranges, vs, cc, mask, rect = vertex_and_tile_shader(indices, vertices, tcam, rg)
    feats = fld.features(vertices, indices, mask)
    feats.retain_grad()
    img, dist, _ = AlphaBlendTiledRender.apply(sidx, ranges, indices, vertices, feats, rg, tcam, jitter)
    torch.cuda.synchronize()
    print(f"GPU forward done: {int(mask.sum())} visible tets, {sidx.numel()} pairs", flush=True)

    # oracle forward on seeded random tiles with the GPU's FP32 features
    ocam = O.make_camera(cam, TILE, 0.1)
    G = ocam.gh * ocam.gw
    tiles = np.sort(np.random.default_rng(17).choice(G, N_TILES, replace=False)).astype(np.int32)
    f_gpu = feats.detach().cpu().numpy().astype(np.float64)
    s_np, r_np = sidx.cpu().numpy(), ranges.cpu().numpy()
    jit = np.full((H, W, 2), 0.5, np.float32)
    o_img, o_dist, o_nc, margin = O.composite_fwd(ocam, s_np, r_np, tets, pts, f_gpu, jit, tiles=tiles)
    pix = np.zeros((H, W), bool)
    for t in tiles:
        ty, tx = divmod(int(t), ocam.gw)
        pix[ty * TILE:(ty + 1) * TILE, tx * TILE:(tx + 1) * TILE] = True
    near = pix & (margin < 1e-4)
    Wt = np.random.default_rng(18).standard_normal((H, W, 3)) * (pix & ~near)[..., None]
    loss = (img[..., :3] * torch.from_numpy(Wt).float().cuda()).sum()
    loss.backward()
    torch.cuda.synchronize()
    print("GPU backward done", flush=True)
    gimg = np.zeros((H, W, 4))
    gimg[..., :3] = Wt
    gv, gf = O.composite_bwd(ocam, s_np, r_np, tets, pts, f_gpu, jit, o_img, o_dist, o_nc, gimg,
                             np.zeros((H, W, 5)), tiles=tiles)
    tet_ids = torch.nonzero(mask).reshape(-1).to(torch.int32)
    x = fieldmod.field_points(vertices, indices, tet_ids, fld.scale)
    return dict(O=O, ocam=ocam, pts=pts, tets=tets, fld=fld, vertices=vertices, feats=feats, img=img,
                sidx=s_np, ranges=r_np, vs=vs.detach().cpu().numpy(), cc=cc.detach().cpu().numpy(),
                rect=rect.cpu().numpy(), mask=mask.cpu().numpy(), tiles=tiles, pix=pix, near=near,
                o_img=o_img, gv=gv, gf=gf, tet_ids=tet_ids.cpu().numpy(), x=x.cpu().numpy())


def test_projection_and_binning_bit_exact(step):
    s = step
    O = s["O"]
    print("oracle projection and binning over every tet", flush=True)
    o_tiles, o_rect, o_vs, o_cc = O.project(s["ocam"], s["tets"], s["pts"])
    np.testing.assert_array_equal(s["mask"], o_tiles > 0)
    np.testing.assert_array_equal(s["rect"], o_rect)
    np.testing.assert_array_equal(s["vs"].view(np.int32), o_vs.view(np.int32))
    np.testing.assert_array_equal(s["cc"].view(np.int32), o_cc.view(np.int32))
    keys, vals, ranges = O.bin_tiles(s["ocam"], o_tiles, o_rect, o_vs)
    print(f"{int((o_tiles > 0).sum())} visible tets, {len(vals)} pairs", flush=True)
    np.testing.assert_array_equal(s["sidx"], vals)
    np.testing.assert_array_equal(s["ranges"], ranges)


def test_image_on_sampled_tiles(step):
    s = step
    g = s["img"].detach().cpu().numpy().astype(np.float64)
    pix, near = s["pix"], s["near"]
    err = np.abs(g - s["o_img"])[..., :3]
    far = pix & ~near
    e_far = err[far].max()
    print(f"{int(pix.sum())} pixels of {N_TILES} tiles: rgb max-abs err vs FP64 {e_far:.3e} "
          f"(mean {err[far].mean():.2e}); {int(near.sum())} near-threshold pixels", flush=True)
    assert e_far <= 1e-5
    logt = np.abs(g[..., 3] - s["o_img"][..., 3])[far] / np.maximum(1, np.abs(s["o_img"][..., 3][far]))
    assert logt.max() <= 1e-5
    assert near.sum() <= 0.02 * pix.sum()
    if near.any():
        # past the threshold the pixel keeps at most exp(-5.5 + 1e-4) of its transmittance, so a
        # different stopping tet changes it by at most that times the largest colour magnitude
        f = s["feats"].detach().cpu().numpy().astype(np.float64)
        tl = np.unique(s["sidx"][np.concatenate([np.arange(*s["ranges"][t]) for t in s["tiles"]])])
        v = s["pts"].astype(np.float64)[s["tets"][tl]]
        diam = np.linalg.norm(v[:, :, None] - v[:, None], axis=-1).max((1, 2))
        cmax = (np.abs(f[tl, 1:4]) + np.abs(f[tl, 4:13]).reshape(-1, 3, 3).sum(2) * diam[:, None]).max()
        bound = np.exp(-5.5 + 1e-4) * cmax + 1e-5
        print(f"near-threshold pixels: max err {err[near].max():.3e} (bound {bound:.3e})")
        assert err[near].max() <= bound


def test_compositing_gradients(step):
    s = step
    gf_gpu = s["feats"].grad.cpu().numpy().astype(np.float64)
    gv_gpu = s["vertices"].grad.cpu().numpy().astype(np.float64)
    for name, sl in (("density", slice(0, 1)), ("colour", slice(1, 4)), ("gradient", slice(4, 13))):
        rel = _rel(gf_gpu[:, sl], s["gf"][:, sl])
        nz = int((s["gf"][:, sl] != 0).sum())
        print(f"feature grad {name}: {nz} non-zero, max per-element rel err {rel.max():.2e}, "
              f"99.9th pct {np.percentile(rel[s['gf'][:, sl] != 0], 99.9):.2e}")
        assert nz > 1000
        assert rel.max() <= 1e-3, name
    rel = _rel(gv_gpu, s["gv"])
    print(f"vertex grad: {int((s['gv'] != 0).sum())} non-zero, max per-element rel err {rel.max():.2e}")
    assert rel.max() <= 1e-3


def _fp64_field(s, rows):
    """FP64 restatement of the field on the GPU's query points of the listed visible-tet rows:
    hash encoding (oracle.hashgrid_fwd), MLP with SELU (models/ingp_color.py:141-151), head
    activation (density exp, colour + 0.5, gradient rows as they are)."""
    O = s["O"]
    g = s["fld"].grid
    m = s["fld"].mlp
    w = [t.detach().cpu().numpy().astype(np.float64) for t in (m.w1, m.b1, m.w2, m.b2, m.w3, m.b3)]
    x = s["x"][rows]
    table = g.table.detach().cpu().numpy()
    h0 = O.hashgrid_fwd(x, table, g.res, g.offs, g.n_feat, g.log2_size)
    h1 = h0 @ w[0].T + w[1]
    a1 = O.selu(h1)
    h2 = a1 @ w[2].T + w[3]
    a2 = O.selu(h2)
    raw = a2 @ w[4].T + w[5]
    return dict(w=w, x=x, h0=h0, h1=h1, a1=a1, h2=h2, a2=a2, raw=raw)


def test_field_forward(step):
    s = step
    rows = np.random.default_rng(19).choice(len(s["tet_ids"]), 20_000, replace=False)
    f = _fp64_field(s, rows)
    feats = s["feats"].detach().cpu().numpy().astype(np.float64)[s["tet_ids"][rows]]
    dens_rel = np.abs(feats[:, 0] - np.exp(np.minimum(f["raw"][:, 0], np.log(1e20)))) / np.exp(f["raw"][:, 0])
    col = np.abs(feats[:, 1:4] - (f["raw"][:, 1:4] + 0.5))
    grd = np.abs(feats[:, 4:13] - f["raw"][:, 4:13])
    print(f"field forward on 20k queries: density max rel err {dens_rel.max():.2e}, colour max abs "
          f"{col.max():.2e}, gradient rows max abs {grd.max():.2e}")
    assert dens_rel.max() <= 1e-5 and col.max() <= 1e-5 and grd.max() <= 1e-5


def test_field_gradients(step):
    """The chain from the feature gradients down to the MLP weights and the hash table, in FP64 on
    the queries whose features received a gradient (every other query contributes exactly 0)."""
    import torch
    s = step
    O = s["O"]
    gf = s["gf"][s["tet_ids"]]                          # FP64 feature grads of the visible rows
    rows = np.nonzero(np.abs(gf).sum(1) > 0)[0]
    print(f"field chain on {len(rows)} queries with a feature gradient", flush=True)
    f = _fp64_field(s, rows)
    gfr = gf[rows]
    graw = gfr.copy()
    graw[:, 0] = gfr[:, 0] * np.exp(f["raw"][:, 0]) * (f["raw"][:, 0] < np.log(1e20))
    w1, b1, w2, b2, w3, b3 = f["w"]
    ref = {"w3": graw.T @ f["a2"], "b3": graw.sum(0)}
    gh2 = (graw @ w3) * _selu_prime(f["h2"])
    ref.update(w2=gh2.T @ f["a1"], b2=gh2.sum(0))
    gh1 = (gh2 @ w2) * _selu_prime(f["h1"])
    ref.update(w1=gh1.T @ f["h0"], b1=gh1.sum(0))
    gh0 = gh1 @ w1
    # magnitude of each element's terms: the same chain on absolute values. An element that is the
    # sum of 13k terms cancelling to 1/1000 of their magnitude cannot come out of FP32 sums to 1e-3
    # of itself; for those the bound is FP32 summation rounding, 64 * 2^-24 of the terms' magnitude
    mg = np.abs(graw)
    mh2 = (mg @ np.abs(w3)) * _selu_prime(f["h2"])
    mh1 = (mh2 @ np.abs(w2)) * _selu_prime(f["h1"])
    mag = {"w3": mg.T @ np.abs(f["a2"]), "b3": mg.sum(0), "w2": mh2.T @ np.abs(f["a1"]), "b2": mh2.sum(0),
           "w1": mh1.T @ np.abs(f["h0"]), "b1": mh1.sum(0)}
    m = s["fld"].mlp
    for name in ("w1", "b1", "w2", "b2", "w3", "b3"):
        g = getattr(m, name).grad.cpu().numpy().astype(np.float64)
        rel = _rel(g, ref[name])
        cond = np.abs(g - ref[name]) <= 64 * 2.0 ** -24 * mag[name]
        print(f"MLP {name}: max per-element rel err vs FP64 {rel.max():.2e}; elements past 1e-3 within the FP32 "
              f"summation bound: {int(((rel > 1e-3) & cond).sum())} of {rel.size}")
        assert ((rel <= 1e-3) | cond).all(), name
        assert (rel <= 1e-3).mean() >= 0.999, name
    g = s["fld"].grid
    gt = O.hashgrid_bwd(f["x"], gh0, g.res, g.offs, g.n_feat, g.log2_size, int(g.offs[-1]))
    gt_gpu = g.table.grad.cpu().numpy().astype(np.float64)
    nz = np.abs(gt).sum(1) > 0
    rel = _rel(gt_gpu, gt)
    print(f"hash table: {int(nz.sum())} rows with a gradient, max per-element rel err {rel.max():.2e}, "
          f"rows without one all zero on the GPU: {not gt_gpu[~nz].any()}")
    assert nz.sum() > 10_000
    assert rel.max() <= 1e-3
    assert np.abs(gt_gpu[~nz]).max() <= 1e-4 * np.abs(gt).max()
    torch.cuda.synchronize()

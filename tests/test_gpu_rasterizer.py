"""GPU parity of the rasterizer against the CPU oracle: projection and binning bit-exact,
compositing within 1e-5 of the FP64 oracle, backward against the FP64 adjoint."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from tests._util import device_inputs, f32_bits, grid_of, jitter, scene, tcam_of

# ("plain", 16, 13) is the 10k-point workload exactly: 256x256, 16x16 tiles, 3x3 colour gradient
CASES = [("plain", 16, 7), ("plain", 16, 13), ("plain", 4, 7), ("plain", 8, 13), ("ball", 16, 7), ("ball", 4, 13)]


def _project(sc):
    import torch
    from radiance_meshes_amd.rasterizer import VertexShader
    verts, idx, feats, jit = device_inputs(sc)
    tiles, rect, vs, cc = VertexShader.apply(idx, verts, tcam_of(sc), grid_of(sc))
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (tiles, rect, vs, cc)]


@pytest.mark.parametrize("kind,tile,nf", CASES)
def test_projection_bit_exact(oracle_lib, kind, tile, nf):
    O = oracle_lib
    sc = scene(kind, tile, nf)
    print(f"case {kind} tile={tile} T={len(sc.indices)}", flush=True)
    g_tiles, g_rect, g_vs, g_cc = _project(sc)
    cam = O.make_camera(sc.camera, sc.tile_size, sc.min_t)
    o_tiles, o_rect, o_vs, o_cc = O.project(cam, sc.indices, sc.vertices)
    assert (g_tiles > 0).sum() > 0
    np.testing.assert_array_equal(g_tiles, o_tiles)
    np.testing.assert_array_equal(g_rect, o_rect)
    np.testing.assert_array_equal(f32_bits(g_vs), f32_bits(o_vs))
    np.testing.assert_array_equal(f32_bits(g_cc), f32_bits(o_cc))


@pytest.mark.parametrize("kind,tile,nf", CASES)
@pytest.mark.parametrize("fused", [True, False])
def test_binning_bit_exact(oracle_lib, kind, tile, nf, fused):
    import torch
    from radiance_meshes_amd.rasterizer import vertex_and_tile_shader
    O = oracle_lib
    sc = scene(kind, tile, nf)
    print(f"case {kind} tile={tile} fused={fused}", flush=True)
    verts, idx, feats, jit = device_inputs(sc)
    sidx, ranges, vs, cc, mask, rect = vertex_and_tile_shader(idx, verts, tcam_of(sc), grid_of(sc), fused=fused)
    torch.cuda.synchronize()
    cam = O.make_camera(sc.camera, sc.tile_size, sc.min_t)
    o_tiles, o_rect, o_vs, o_cc = O.project(cam, sc.indices, sc.vertices)
    o_keys, o_vals, o_ranges = O.bin_tiles(cam, o_tiles, o_rect, o_vs)
    np.testing.assert_array_equal(sidx.cpu().numpy(), o_vals)
    np.testing.assert_array_equal(ranges.cpu().numpy(), o_ranges)
    np.testing.assert_array_equal(mask.cpu().numpy(), o_tiles > 0)


def test_sort_by_keys_matches_stable_sort():
    import torch
    from radiance_meshes_amd.rasterizer import sort_by_keys_cub
    rng = np.random.default_rng(3)
    n = 300_001
    tile = rng.integers(0, 4000, n).astype(np.int64)
    depth = rng.integers(0, 1 << 20, n).astype(np.int64)   # many duplicate keys
    keys = (tile << 32) | depth
    vals = np.arange(n, dtype=np.int32)[::-1].copy()
    print(f"sort n={n}", flush=True)
    ko, vo = sort_by_keys_cub.sort_by_keys(torch.from_numpy(keys).cuda(), torch.from_numpy(vals).cuda(), 12)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(ko.cpu().numpy(), keys[order])
    np.testing.assert_array_equal(vo.cpu().numpy(), vals[order])


def _composite(sc, jit_seed=None):
    import torch
    from radiance_meshes_amd.rasterizer import AlphaBlendTiledRender, vertex_and_tile_shader
    verts, idx, feats, _ = device_inputs(sc)
    jit = torch.from_numpy(jitter(sc, jit_seed)).cuda()
    tcam = tcam_of(sc)
    rg = grid_of(sc)
    sidx, ranges, vs, cc, mask, rect = vertex_and_tile_shader(idx, verts, tcam, rg)
    img, dist, alive = AlphaBlendTiledRender.apply(sidx, ranges, idx, verts, feats, rg, tcam, jit)
    torch.cuda.synchronize()
    return sidx, ranges, img, dist


def _oracle_fwd(O, sc, jit_seed=None):
    cam = O.make_camera(sc.camera, sc.tile_size, sc.min_t)
    t, r, vs, cc = O.project(cam, sc.indices, sc.vertices)
    keys, vals, ranges = O.bin_tiles(cam, t, r, vs)
    img, dist, nc, margin = O.composite_fwd(cam, vals, ranges, sc.indices, sc.vertices, sc.features,
                                            jitter(sc, jit_seed))
    return cam, vals, ranges, img, dist, nc, margin


@pytest.mark.parametrize("kind,tile,nf", CASES)
def test_composite_forward_fp64(oracle_lib, kind, tile, nf):
    O = oracle_lib
    sc = scene(kind, tile, nf)
    print(f"case {kind} tile={tile} nf={nf}", flush=True)
    sidx, ranges, img, dist = _composite(sc, jit_seed=7)
    cam, vals, o_ranges, o_img, o_dist, o_nc, margin = _oracle_fwd(O, sc, jit_seed=7)
    g_img = img.cpu().numpy().astype(np.float64)
    # pixels whose termination decision sits within FP32 noise of the -5.5 threshold are
    # reported, not compared (their n_contrib differs by construction)
    flip = margin < 1e-4
    err = np.abs(g_img - o_img)[~flip]
    rgb_err = err[:, :3].max()
    logt_err = (np.abs(g_img[..., 3] - o_img[..., 3]) / np.maximum(1, np.abs(o_img[..., 3])))[~flip].max()
    print(f"rgb max-abs err vs FP64 = {rgb_err:.3e}, logT rel err = {logt_err:.3e}, "
          f"near-threshold pixels = {int(flip.sum())}/{flip.size}")
    assert rgb_err <= 1e-5
    assert logt_err <= 1e-5
    # near-threshold pixels: few, and off by at most what is left past the threshold (transmittance
    # exp(-5.5 + 1e-4) times the largest colour magnitude of the scene's tets)
    assert flip.sum() <= 0.01 * flip.size
    if flip.any():
        f = sc.features.astype(np.float64)
        v = sc.vertices.astype(np.float64)[sc.indices]
        diam = np.linalg.norm(v[:, :, None] - v[:, None], axis=-1).max((1, 2))
        grd = np.abs(f[:, 4:]).reshape(len(f), -1, 3).sum(2).max(1)
        cmax = (np.abs(f[:, 1:4]).max(1) + grd * diam).max()
        e_flip = np.abs(g_img - o_img)[flip][:, :3].max()
        print(f"near-threshold pixels: max err {e_flip:.3e} (bound {np.exp(-5.5 + 1e-4) * cmax:.3e})")
        assert e_flip <= np.exp(-5.5 + 1e-4) * cmax + 1e-5
    d_err = (np.abs(dist.cpu().numpy() - o_dist) / np.maximum(1, np.abs(o_dist)))[~flip].max()
    print(f"distortion rel err = {d_err:.3e}")
    assert d_err <= 1e-4


@pytest.mark.parametrize("kind,tile,nf", [("ball", 16, 7), ("ball", 4, 7), ("ball", 8, 13), ("plain", 16, 7),
                                          ("plain", 16, 13), ("plain", 4, 13)])
def test_composite_backward_fp64(oracle_lib, kind, tile, nf):
    import torch
    from radiance_meshes_amd.rasterizer import AlphaBlendTiledRender, vertex_and_tile_shader
    O = oracle_lib
    sc = scene(kind, tile, nf)
    print(f"case {kind} tile={tile} nf={nf}", flush=True)
    verts, idx, feats, _ = device_inputs(sc)
    verts.requires_grad_(True)
    feats.requires_grad_(True)
    jit = torch.from_numpy(jitter(sc, 5)).cuda()
    tcam, rg = tcam_of(sc), grid_of(sc)
    H, W = sc.camera.height, sc.camera.width
    rng = np.random.default_rng(11)
    w_img = rng.standard_normal((H, W, 4))
    w_dist = rng.standard_normal((H, W, 5)) * 0.1
    # pixels whose termination sits within FP32 noise of the -5.5 threshold end their walk at a tet
    # that depends on rounding (the forward test bounds how many there are): they carry no loss
    # weight, so both sides differentiate the same function
    cam, vals, o_ranges, o_img, o_dist, o_nc, margin = _oracle_fwd(O, sc, jit_seed=5)
    w_img[margin < 1e-4] = 0
    w_dist[margin < 1e-4] = 0
    print(f"{int((margin < 1e-4).sum())} near-threshold pixels without loss weight", flush=True)
    sidx, ranges, vs, cc, mask, rect = vertex_and_tile_shader(idx, verts, tcam, rg)
    img, dist, alive = AlphaBlendTiledRender.apply(sidx, ranges, idx, verts, feats, rg, tcam, jit)
    loss = (img * torch.from_numpy(w_img).float().cuda()).sum() + (dist * torch.from_numpy(w_dist).float().cuda()).sum()
    loss.backward()
    torch.cuda.synchronize()
    gv, gf = O.composite_bwd(cam, vals, o_ranges, sc.indices, sc.vertices, sc.features, jitter(sc, 5),
                             o_img, o_dist, o_nc, w_img, w_dist)
    # conditioning of each gradient element under FP32 rounding: the same FP64 adjoint with every
    # vertex coordinate moved by one FP32 ulp of max(|v|, |camera|) in a random direction (rays
    # grazing a tet make its segment length, and so its colour weight, sensitive at this level)
    sc2 = _perturbed(sc, seed=17)
    cam2, vals2, r2, img2, dist2, nc2, _ = _oracle_fwd(O, sc2, jit_seed=5)
    gv2, gf2 = O.composite_bwd(cam2, vals2, r2, sc2.indices, sc2.vertices, sc2.features, jitter(sc, 5),
                               img2, dist2, nc2, w_img, w_dist)
    g_v = verts.grad.cpu().numpy().astype(np.float64)
    g_f = feats.grad.cpu().numpy().astype(np.float64)
    groups = {"vertex": (g_v, gv, gv2), "density": (g_f[:, :1], gf[:, :1], gf2[:, :1]),
              "colour": (g_f[:, 1:4], gf[:, 1:4], gf2[:, 1:4]), "gradient": (g_f[:, 4:], gf[:, 4:], gf2[:, 4:])}
    for name, (a, b, b2) in groups.items():
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        # per element: within 1e-3 of the element (floor 1e-4 of the group's largest value), or within
        # four times the change one ulp of input rounding makes to the exact adjoint
        cond = np.abs(b2 - b)
        pe = np.abs(a - b) / (np.abs(b) + 1e-4 * np.abs(b).max() + 4e3 * cond)
        worst = np.unravel_index(np.argmax(np.abs(a - b) / (np.abs(b) + 1e-4 * np.abs(b).max())), b.shape)
        print(f"{name}: rel L2 err vs FP64 adjoint = {rel:.3e}, max per-element {pe.max():.3e} "
              f"(|ref| = {np.linalg.norm(b):.3e}; largest plain relative error at {worst}: gpu {a[worst]:.6e} "
              f"fp64 {b[worst]:.6e} perturbed fp64 {b2[worst]:.6e})")
        assert rel <= 1e-3, name
        assert pe.max() <= 1e-3, name


def _perturbed(sc, seed):
    """sc with each vertex coordinate moved by +-2^-24 max(|v|_inf, |camera|_inf) (one FP32 ulp at the
    magnitude of camera-relative coordinates), in FP64."""
    import dataclasses
    rng = np.random.default_rng(seed)
    v = sc.vertices.astype(np.float64)
    scale = np.maximum(np.abs(v).max(1, keepdims=True), np.abs(np.asarray(sc.camera.cam_pos, np.float64)).max())
    v = v + rng.choice([-1.0, 1.0], v.shape) * scale * 2.0 ** -24
    return dataclasses.replace(sc, vertices=v)


@pytest.mark.parametrize("kind,tile,nf", [("plain", 16, 7), ("ball", 8, 13)])
def test_composite_forward_fp32_planes(oracle_lib, kind, tile, nf):
    """The all-FP32 intersection mode: same images up to the FP32 conditioning of grazing
    ray/face crossings (the reference's own precision)."""
    from radiance_meshes_amd.rasterizer import set_intersect_precision
    O = oracle_lib
    sc = scene(kind, tile, nf)
    set_intersect_precision(False)
    try:
        sidx, ranges, img, dist = _composite(sc, jit_seed=7)
    finally:
        set_intersect_precision(True)
    cam, vals, o_ranges, o_img, o_dist, o_nc, margin = _oracle_fwd(O, sc, jit_seed=7)
    flip = margin < 1e-3
    err = np.abs(img.cpu().numpy().astype(np.float64) - o_img)[~flip][:, :3]
    print(f"fp32 planes: rgb max-abs {err.max():.3e}, 99.9th pct {np.percentile(err, 99.9):.3e}")
    assert np.percentile(err, 99.9) <= 5e-5
    assert err.max() <= 2e-3


@pytest.mark.parametrize("kind,tile,nf", CASES)
@pytest.mark.parametrize("fp64", [True, False])
def test_composite_backward_rows_match_direct(kind, tile, nf, fp64):
    """rm_composite_bwd_ws (per-tet rows of summed plane adjoints, converted once per tet) against
    rm_composite_bwd (conversion per tile and tet): the same gradients up to FP32 summation order."""
    import torch
    from radiance_meshes_amd.rasterizer import _cam, set_intersect_precision, vertex_and_tile_shader
    from radiance_meshes_amd._capi import check, lib, ptr, stream_of
    sc = scene(kind, tile, nf)
    print(f"case {kind} tile={tile} nf={nf} fp64={fp64}", flush=True)
    verts, idx, feats, _ = device_inputs(sc)
    jit = torch.from_numpy(jitter(sc, 5)).cuda()
    tcam, rg = tcam_of(sc), grid_of(sc)
    H, W = sc.camera.height, sc.camera.width
    sidx, ranges, vs, cc, mask, rect = vertex_and_tile_shader(idx, verts, tcam, rg)
    cam = _cam(tcam, rg)
    L = lib()
    st = stream_of(verts.device)
    img = torch.empty((H, W, 4), device="cuda")
    dist = torch.empty((H, W, 5), device="cuda")
    nc = torch.empty((H, W), device="cuda", dtype=torch.int32)
    gen = torch.Generator(device="cuda").manual_seed(3)
    gimg = torch.randn((H, W, 4), device="cuda", generator=gen)
    gdist = torch.randn((H, W, 5), device="cuda", generator=gen) * 0.1
    set_intersect_precision(fp64)
    try:
        check(L.rm_composite_fwd(cam, ptr(sidx), sidx.numel(), ptr(ranges), ptr(idx), ptr(verts), ptr(feats), nf,
                                 ptr(jit), ptr(img), ptr(dist), ptr(nc), st), "rm_composite_fwd")
        out = []
        for ws_path in (False, True):
            gv = torch.zeros_like(verts)
            gf = torch.zeros_like(feats)
            if ws_path:
                T = idx.shape[0]
                wsb = L.rm_composite_bwd_workspace_size(T)
                ws = torch.full((wsb,), 7, device="cuda", dtype=torch.uint8)   # stale contents
                check(L.rm_composite_bwd_ws(cam, ptr(sidx), sidx.numel(), ptr(ranges), ptr(idx), T, ptr(verts),
                                            ptr(feats), nf, ptr(jit), ptr(img), ptr(dist), ptr(nc), ptr(gimg),
                                            ptr(gdist), ptr(gv), ptr(gf), ptr(ws), wsb, st), "rm_composite_bwd_ws")
            else:
                check(L.rm_composite_bwd(cam, ptr(sidx), sidx.numel(), ptr(ranges), ptr(idx), ptr(verts), ptr(feats),
                                         nf, ptr(jit), ptr(img), ptr(dist), ptr(nc), ptr(gimg), ptr(gdist), ptr(gv),
                                         ptr(gf), st), "rm_composite_bwd")
            torch.cuda.synchronize()
            out.append((gv.cpu().double().numpy(), gf.cpu().double().numpy()))
    finally:
        set_intersect_precision(True)
    for name, a, b in (("vertex", out[1][0], out[0][0]), ("feature", out[1][1], out[0][1])):
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(f"{name}: rows vs direct rel L2 = {rel:.3e} (|direct| = {np.linalg.norm(b):.3e})")
        assert np.linalg.norm(b) > 0
        assert rel <= 2e-5, name
    # tets no tile list names keep exactly zero gradients
    listed = np.zeros(idx.shape[0], bool)
    listed[sidx.cpu().numpy()] = True
    assert not out[1][1][~listed].any()


def test_differentiation_check_central_differences(oracle_lib):
    """The differentiation workload: 2k points uniform in the unit ball (~12k tets), FP64 scene
    parameters, 64x64 view; gradients of sum(image * W), W ~ N(0,1), w.r.t. per-tet density,
    colour and vertex positions from the HIP backward, against FP64 central finite differences
    (h = 1e-6) of the oracle's forward on 256 randomly chosen parameters; tolerance 1e-3 relative
    (with a floor of 1e-2 of the largest derivative for parameters whose derivative is ~0)."""
    import torch
    from radiance_meshes_amd.rasterizer import AlphaBlendTiledRender, vertex_and_tile_shader
    O = oracle_lib
    sc = scene("ball", 8, 7)
    V64 = sc.vertices.astype(np.float64)
    F64 = sc.features.astype(np.float64)
    H, W = sc.camera.height, sc.camera.width
    print(f"differentiation check T={len(sc.indices)} {H}x{W}", flush=True)
    rng = np.random.default_rng(12)
    Wt = rng.standard_normal((H, W, 3))
    verts, idx, feats, jit = device_inputs(sc)
    verts.requires_grad_(True)
    feats.requires_grad_(True)
    tcam, rg = tcam_of(sc), grid_of(sc)
    sidx, ranges, vs, cc, mask, rect = vertex_and_tile_shader(idx, verts, tcam, rg)
    img, dist, _ = AlphaBlendTiledRender.apply(sidx, ranges, idx, verts, feats, rg, tcam, jit)
    (img[..., :3] * torch.from_numpy(Wt).float().cuda()).sum().backward()
    torch.cuda.synchronize()
    g_v = verts.grad.cpu().numpy().astype(np.float64)
    g_f = feats.grad.cpu().numpy().astype(np.float64)
    cam = O.make_camera(sc.camera, sc.tile_size, sc.min_t)
    vals, o_ranges = sidx.cpu().numpy(), ranges.cpu().numpy()
    j = jitter(sc)

    def loss(Vx, Fx):
        im, _, _, _ = O.composite_fwd(cam, vals, o_ranges, sc.indices, Vx, Fx, j)
        return (im[..., :3] * Wt).sum()

    used_t = np.unique(vals)
    used_v = np.unique(sc.indices[used_t])
    h = 1e-6
    rows = []
    for k in range(256):
        kind = k % 3
        if kind == 0:      # density
            t, c = rng.choice(used_t), 0
        elif kind == 1:    # colour (base rgb or gradient)
            t, c = rng.choice(used_t), int(rng.integers(1, 7))
        if kind < 2:
            Fp, Fm = F64.copy(), F64.copy()
            Fp[t, c] += h
            Fm[t, c] -= h
            rows.append((kind, g_f[t, c], (loss(V64, Fp) - loss(V64, Fm)) / (2 * h)))
        else:              # vertex position
            v, c = rng.choice(used_v), int(rng.integers(0, 3))
            Vp, Vm = V64.copy(), V64.copy()
            Vp[v, c] += h
            Vm[v, c] -= h
            rows.append((kind, g_v[v, c], (loss(Vp, F64) - loss(Vm, F64)) / (2 * h)))
    a = np.array(rows)
    for kind, name in enumerate(("density", "colour", "vertex")):
        r = a[a[:, 0] == kind]
        scale = np.abs(r[:, 2]).max()
        err = np.abs(r[:, 1] - r[:, 2]) / np.maximum(np.abs(r[:, 2]), 1e-2 * scale)
        print(f"{name}: {len(r)} params, max rel err vs FP64 central differences {err.max():.2e}, "
              f"median {np.median(err):.2e}")
        assert err.max() <= 1e-3, name


@pytest.mark.parametrize("kind,tile", [("plain", 16), ("ball", 8)])
def test_projection_backward_fp64(oracle_lib, kind, tile):
    """vertex_shader.bwd (tile_shader_slang.py:127-150): a loss reaching vs[:,2] (the circumsphere
    power) and the circumcenters, sum(w * vs[:,2]) + sum(u * cc), backpropagated through
    VertexShader.apply, against the oracle's FP64 adjoint of the same outputs; per-element
    relative error with a floor of 1e-6 of the largest gradient."""
    import torch
    from radiance_meshes_amd.rasterizer import VertexShader
    O = oracle_lib
    sc = scene(kind, tile, 7)
    print(f"projection backward {kind} tile={tile} T={len(sc.indices)}", flush=True)
    verts, idx, feats, jit = device_inputs(sc)
    verts.requires_grad_(True)
    T = idx.shape[0]
    rng = np.random.default_rng(31)
    w = rng.standard_normal(T)
    u = rng.standard_normal((T, 3))
    tiles, rect, vs, cc = VertexShader.apply(idx, verts, tcam_of(sc), grid_of(sc))
    loss = (vs[:, 2] * torch.from_numpy(w).float().cuda()).sum() + (cc * torch.from_numpy(u).float().cuda()).sum()
    loss.backward()
    torch.cuda.synchronize()
    g = verts.grad.cpu().numpy().astype(np.float64)
    cam = O.make_camera(sc.camera, sc.tile_size, sc.min_t)
    o_tiles, _, _, _ = O.project(cam, sc.indices, sc.vertices)
    np.testing.assert_array_equal(tiles.cpu().numpy(), o_tiles)
    gvs = np.zeros((T, 3))
    gvs[:, 2] = w
    o = O.project_bwd(cam, sc.indices, sc.vertices.astype(np.float64), o_tiles, gvs, u)
    floor = 1e-6 * np.abs(o).max()
    rel = np.abs(g - o) / (np.abs(o) + floor)
    print(f"grad_vertices: max per-element rel err vs FP64 adjoint {rel.max():.2e} "
          f"(median {np.median(rel):.2e}, |g|max {np.abs(o).max():.3e})")
    assert np.abs(o).max() > 0
    assert rel.max() <= 1e-3


@pytest.mark.parametrize("kind,tile,nf", [("plain", 4, 7), ("plain", 8, 13), ("ball", 4, 13), ("ball", 8, 7)])
@pytest.mark.parametrize("fp64", [True, False])
def test_flagged_pair_matches_workspace_path(kind, tile, nf, fp64):
    """rm_composite_fwd_flags writes the same image, distortion and n_contrib as rm_composite_fwd
    (bit for bit), here with the tiles taken longest list first (rm_tile_order) at 4x4 tiles;
    rm_composite_bwd_flags, which walks only the entries some pixel of the tile blended, gives the
    gradients of rm_composite_bwd_ws (per element within 1e-5 of the largest: the two differ only
    in the order of float additions); on the scene an entry is skipped for."""
    import torch
    from radiance_meshes_amd.rasterizer import _cam, set_intersect_precision, tile_order, vertex_and_tile_shader
    from radiance_meshes_amd._capi import check, lib, ptr, stream_of
    sc = scene(kind, tile, nf)
    print(f"flagged pair {kind} tile={tile} nf={nf} fp64={fp64}", flush=True)
    verts, idx, feats, _ = device_inputs(sc)
    jit = torch.from_numpy(jitter(sc, 5)).cuda()
    tcam, rg = tcam_of(sc), grid_of(sc)
    H, W = sc.camera.height, sc.camera.width
    sidx, ranges, vs, cc, mask, rect = vertex_and_tile_shader(idx, verts, tcam, rg)
    cam = _cam(tcam, rg)
    L = lib()
    st = stream_of(verts.device)
    P, T = sidx.numel(), idx.shape[0]
    outs = []
    gen = torch.Generator(device="cuda").manual_seed(3)
    gimg = torch.randn((H, W, 4), device="cuda", generator=gen)
    gdist = torch.randn((H, W, 5), device="cuda", generator=gen) * 0.1
    wsb = L.rm_composite_bwd_workspace_size(T)
    set_intersect_precision(fp64)
    try:
        for flagged in (False, True):
            img = torch.empty((H, W, 4), device="cuda")
            dist = torch.empty((H, W, 5), device="cuda")
            nc = torch.empty((H, W), device="cuda", dtype=torch.int32)
            flags = torch.full((P,), 7, device="cuda", dtype=torch.uint8)
            ws = torch.empty((wsb,), device="cuda", dtype=torch.uint8)
            gv = torch.zeros_like(verts)
            gf = torch.zeros_like(feats)
            if flagged:
                order = tile_order(ranges) if tile == 4 else None
                if order is not None:   # a permutation of the tiles, non-increasing list lengths
                    n_of = (ranges[:, 1] - ranges[:, 0]).cpu().numpy()
                    o = order.cpu().numpy()
                    assert np.array_equal(np.sort(o), np.arange(len(o))) and (np.diff(n_of[o]) <= 0).all()
                check(L.rm_composite_fwd_flags(cam, ptr(sidx), P, ptr(ranges), ptr(idx), ptr(verts), ptr(feats), nf,
                                               ptr(jit), ptr(img), ptr(dist), ptr(nc), ptr(flags), ptr(order), st),
                      "fwd_flags")
                check(L.rm_composite_bwd_flags(cam, ptr(sidx), P, ptr(ranges), ptr(idx), T, ptr(verts), ptr(feats), nf,
                                               ptr(jit), ptr(img), ptr(dist), ptr(nc), ptr(flags), ptr(order),
                                               ptr(gimg), ptr(gdist), ptr(gv), ptr(gf), ptr(ws), wsb, st),
                      "bwd_flags")
            else:
                check(L.rm_composite_fwd(cam, ptr(sidx), P, ptr(ranges), ptr(idx), ptr(verts), ptr(feats), nf,
                                         ptr(jit), ptr(img), ptr(dist), ptr(nc), st), "fwd")
                check(L.rm_composite_bwd_ws(cam, ptr(sidx), P, ptr(ranges), ptr(idx), T, ptr(verts), ptr(feats), nf,
                                            ptr(jit), ptr(img), ptr(dist), ptr(nc), ptr(gimg), ptr(gdist), ptr(gv),
                                            ptr(gf), ptr(ws), wsb, st), "bwd_ws")
            torch.cuda.synchronize()
            outs.append([t.cpu().numpy() for t in (img, dist, nc, gv, gf, flags)])
    finally:
        set_intersect_precision(True)
    for a, b in zip(outs[0][:3], outs[1][:3]):
        np.testing.assert_array_equal(a.view(np.int32), b.view(np.int32))
    fl = outs[1][5]
    n_ent = int((fl <= 1).sum())
    print(f"{n_ent} of {P} entries walked, {int((fl == 1).sum())} flagged")
    assert int((fl == 0).sum()) > 0 and int((fl == 1).sum()) > 0 and set(np.unique(fl)) <= {0, 1, 7}
    for name, a, b in (("vertex", outs[1][3], outs[0][3]), ("feature", outs[1][4], outs[0][4])):
        err = np.abs(a.astype(np.float64) - b).max() / np.abs(b).max()
        print(f"{name}: flagged vs workspace path, max abs diff / max |grad| = {err:.2e}")
        assert err <= 1e-5, name

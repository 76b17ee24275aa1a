"""Batched serving path: every view rendered in a mixed batch (several scenes, two resolutions,
captured-graph replays) is bit-identical to the same view rendered alone, through the batch path
and through the single-view path; one batched view against the FP64 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from radiance_meshes_amd import scenes
from tests._util import scene

RADIUS = (3.0, 3.0, 6.0)
RES_A = (66, 90)     # not a multiple of the tile: ragged edge tiles
RES_B = (48, 64)


def _server(use_graphs, max_slots=4):
    import torch
    from radiance_meshes_amd.serving import RenderServer
    # capacity probe closer than any test view and a wide margin: these tests never overflow
    srv = RenderServer(tile=4, n_features=13, min_t=0.1, hbm_budget_bytes=4 << 30, max_slots=max_slots,
                       use_graphs=use_graphs, pair_margin=4.0, probe_radius=2.5,
                       on_overflow="raise")
    for kind, kw in (("plain", {}), ("ball", {}), ("gmm", dict(n_points=20000, width=160, height=120))):
        sc = scene(kind, 4, 13, **kw)
        srv.add_scene(torch.from_numpy(np.ascontiguousarray(sc.vertices, np.float32)),
                      torch.from_numpy(np.ascontiguousarray(sc.indices, np.int32)),
                      torch.from_numpy(np.ascontiguousarray(sc.features, np.float32)))
    return srv


def _views(res, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        s = int(k % 3)
        cam = scenes.orbit_camera(RADIUS[s], rng.uniform(0, 360), rng.uniform(-30, 40), res[1], res[0], 50.0)
        out.append((s, cam))
    return out


def _single_full(srv, sid, cam):
    """Single-view path with pixel-centre rays: (image, distortion, n_contrib)."""
    import torch
    from radiance_meshes_amd._capi import check, lib, ptr, stream_of
    from radiance_meshes_amd.rasterizer import RenderGrid, _cam, vertex_and_tile_shader
    sc = srv.scenes[sid]
    rg = RenderGrid(cam.height, cam.width, srv.tile, srv.tile)
    tcam = cam.tcam(srv.tile, srv.min_t)
    sidx, ranges, *_ = vertex_and_tile_shader(sc.indices, sc.vertices, tcam, rg)
    H, W = cam.height, cam.width
    jit = torch.full((H, W, 2), 0.5, device="cuda")
    img = torch.empty((H, W, 4), device="cuda")
    dist = torch.empty((H, W, 5), device="cuda")
    nc = torch.empty((H, W), device="cuda", dtype=torch.int32)
    check(lib().rm_composite_fwd(_cam(tcam, rg), ptr(sidx), sidx.numel(), ptr(ranges), ptr(sc.indices),
                                 ptr(sc.vertices), ptr(sc.features), 13, ptr(jit), ptr(img), ptr(dist), ptr(nc),
                                 stream_of()), "rm_composite_fwd")
    torch.cuda.synchronize()
    return img, dist, nc


def _assert_same(a, b, what):
    import torch
    for x, y, name in zip(a, b, ("image", "distortion", "n_contrib")):
        assert x.shape == y.shape, (what, name)
        if not torch.equal(x, y):
            d = (x.double() - y.double()).abs().max().item()
            raise AssertionError(f"{what}: {name} differs (max abs {d:.3e})")


@pytest.mark.parametrize("use_graphs", [False, True])
def test_mixed_batch_matches_single_views(use_graphs):
    import torch
    print(f"serving mixed batch use_graphs={use_graphs}", flush=True)
    srv = _server(use_graphs)
    views_a = _views(RES_A, 7, 1)     # two batches of the 4-slot bucket: 4 + 3 views
    views_b = _views(RES_B, 3, 2)
    ref = {id(c): _single_full(srv, s, c) for s, c in views_a + views_b}
    for rep in range(3):
        got_a = srv.render_views(views_a, keep_aux=True)
        got_b = srv.render_views(views_b, keep_aux=True)
        for (s, c), g in zip(views_a + views_b, got_a + got_b):
            assert len(g) == 3, "pair capacity overflowed"
            _assert_same(g, ref[id(c)], f"replay {rep} scene {s} {c.height}x{c.width}")
        # eager work between replays: single-view renders and unrelated allocations
        _single_full(srv, *views_b[0])
        torch.randn(1 << 20, device="cuda").sum().item()
    for s, c in views_a[:3] + views_b[:1]:
        alone = srv.render_views([(s, c)], keep_aux=True)[0]
        assert len(alone) == 3, "pair capacity overflowed"
        _assert_same(alone, ref[id(c)], f"alone scene {s}")
    assert any(int(r[2].sum()) > 0 for r in ref.values())


def test_batched_view_against_oracle(oracle_lib):
    O = oracle_lib
    srv = _server(False, max_slots=2)
    views = _views(RES_A, 2, 5)
    got = srv.render_views(views, keep_aux=True)
    assert all(len(g) == 3 for g in got), "pair capacity overflowed"
    for (s, c), (img, dist, nc) in zip(views, got):
        sc = srv.scenes[s]
        V = sc.vertices.cpu().numpy()
        I = sc.indices.cpu().numpy()
        Fe = sc.features.cpu().numpy()
        cam = O.make_camera(c, 4, 0.1)
        t, r, vs, cc = O.project(cam, I, V)
        keys, vals, ranges = O.bin_tiles(cam, t, r, vs)
        jit = np.full((c.height, c.width, 2), 0.5, np.float32)
        o_img, o_dist, o_nc, margin = O.composite_fwd(cam, vals, ranges, I, V, Fe, jit)
        keep = margin >= 1e-4
        err = np.abs(img.cpu().numpy().astype(np.float64) - o_img)[keep]
        print(f"scene {s}: rgb max err {err[:, :3].max():.3e}, logT max err {err[:, 3].max():.3e}")
        assert err[:, :3].max() <= 1e-5
        assert (np.abs(img.cpu().numpy()[..., 3] - o_img[..., 3]) / np.maximum(1, np.abs(o_img[..., 3])))[keep].max() <= 1e-5
        np.testing.assert_array_equal(nc.cpu().numpy()[keep], o_nc[keep])


def test_serve_trace_completes_every_request():
    from radiance_meshes_amd.serving import poisson_trace
    srv = _server(True)
    trace = poisson_trace(3, [RES_A, RES_B], rate_hz=400.0, duration_s=0.25, seed=3, radius=3.0)
    for res in (RES_A, RES_B):
        srv.render_views(_views(res, 1, 9))          # create and capture both buckets
    print(f"trace of {len(trace)} requests", flush=True)
    st = srv.serve(trace, mode="batched")
    print(st)
    assert st["overflows"] == 0
    assert st["frames"] == len(trace) and all(r.done >= r.arrival for r in trace)
    for r in trace[:4]:
        ref = _single_full(srv, r.scene, r.camera)[0]
        assert np.array_equal(r.image.cpu().numpy(), ref.cpu().numpy())


def test_batch_right_after_single_view_work():
    """A bucket created and launched while a single-view render's kernels are still queued on the
    current stream: the bucket's buffers may be blocks those kernels were using (the caching
    allocator hands them out when the host drops them), so the batch stream must wait for the
    current stream. Every batched view equals the same view rendered alone."""
    import torch
    from radiance_meshes_amd.serving import RenderServer
    srv = RenderServer(tile=4, n_features=13, min_t=0.1, hbm_budget_bytes=8 << 30, max_slots=4,
                       use_graphs=True, pair_margin=4.0, probe_radius=2.5, on_overflow="raise")
    sc = scene("gmm", 4, 13, n_points=100000, width=320, height=240)
    srv.add_scene(torch.from_numpy(np.ascontiguousarray(sc.vertices, np.float32)),
                  torch.from_numpy(np.ascontiguousarray(sc.indices, np.int32)),
                  torch.from_numpy(np.ascontiguousarray(sc.features, np.float32)))
    print(f"scene T={len(sc.indices)}", flush=True)
    for k, (H, W) in enumerate([(240, 320), (200, 256), (120, 160), (240, 320)]):
        cams = [scenes.orbit_camera(3.0, 40.0 * j + 15 * k, 10.0, W, H, 50.0) for j in range(3)]
        srv.render_single(0, cams[0])            # left running on the current stream
        out = srv.render_views([(0, c) for c in cams], keep_aux=True)
        torch.cuda.synchronize()
        for j, c in enumerate(cams):
            _assert_same(out[j], _single_full(srv, 0, c), f"{W}x{H} view {j}")


def test_overflow_fallback_keeps_aux_outputs():
    """A batch whose pairs exceed the bucket and whose workspace may not grow (HBM budget) is
    rendered view by view; with keep_aux it still returns (image, distortion, n_contrib), equal to
    the single-view path's outputs."""
    import torch
    from radiance_meshes_amd.serving import RenderServer
    srv = RenderServer(tile=4, n_features=13, min_t=0.1, hbm_budget_bytes=4 << 30, max_slots=4,
                       use_graphs=False, pair_margin=0.05, probe_radius=6.0, on_overflow="grow")
    sc = scene("plain", 4, 13)
    srv.add_scene(torch.from_numpy(np.ascontiguousarray(sc.vertices, np.float32)),
                  torch.from_numpy(np.ascontiguousarray(sc.indices, np.int32)),
                  torch.from_numpy(np.ascontiguousarray(sc.features, np.float32)))
    views = [(0, scenes.orbit_camera(2.5, 40.0 * k, 10.0, RES_A[1], RES_A[0], 50.0)) for k in range(3)]
    b = srv.bucket(*RES_A)
    srv.budget = sum(x.bytes for x in srv.buckets.values())   # no room to grow
    print(f"overflow fallback: bucket pair capacity {b.pair_cap}", flush=True)
    out = srv.render_views(views, keep_aux=True)
    torch.cuda.synchronize()
    assert len(out) == 3
    for (s, c), o in zip(views, out):
        assert len(o) == 3
        _assert_same(o, _single_full(srv, s, c), "fallback view")


def test_add_scene_after_buckets_exist():
    """Scenes added after a bucket was created are served too: the bucket is rebuilt for the new
    scene set, and its views match the single-view path."""
    import torch
    print("add_scene after a bucket exists", flush=True)
    srv = _server(False, max_slots=2)
    first = srv.render_views(_views(RES_B, 2, 7), keep_aux=True)
    torch.cuda.synchronize()
    assert len(first) == 2
    sc = scene("ball", 4, 13, n_points=4000, seed=5)
    sid = srv.add_scene(torch.from_numpy(np.ascontiguousarray(sc.vertices, np.float32)),
                        torch.from_numpy(np.ascontiguousarray(sc.indices, np.int32)),
                        torch.from_numpy(np.ascontiguousarray(sc.features, np.float32)))
    views = [(sid, scenes.orbit_camera(3.0, 25.0, 15.0, RES_B[1], RES_B[0], 50.0)),
             (0, scenes.orbit_camera(3.0, 70.0, -5.0, RES_B[1], RES_B[0], 50.0))]
    out = srv.render_views(views, keep_aux=True)
    torch.cuda.synchronize()
    for (s, c), o in zip(views, out):
        _assert_same(o, _single_full(srv, s, c), f"scene {s} after add_scene")


@pytest.mark.parametrize("use_graphs", [False, True])
def test_model_scene_views_match_render(tmp_path, use_graphs):
    """A checkpoint loaded with Model.load_ckpt served as a view-dependent scene: each view's
    features come from activate_output at its camera centre. Batched views (two poses mixed with a
    plain scene, replayed twice) equal the single-view path bit for bit, and render() of the same
    model (the drop-in path) within 1e-6."""
    import torch
    from radiance_meshes_amd.ingp import Model
    from radiance_meshes_amd.render import render
    from radiance_meshes_amd.serving import RenderServer
    from tests.test_gpu_ckpt import _Camera, _write_checkpoint
    _write_checkpoint(tmp_path)
    m = Model.load_ckpt(tmp_path, "cuda")
    srv = RenderServer(tile=4, n_features=7, min_t=0.1, hbm_budget_bytes=2 << 30, max_slots=4,
                       use_graphs=use_graphs, pair_margin=4.0, probe_radius=2.0, on_overflow="raise")
    sc = scene("plain", 4, 7)
    srv.add_scene(torch.from_numpy(np.ascontiguousarray(sc.vertices, np.float32)),
                  torch.from_numpy(np.ascontiguousarray(sc.indices, np.int32)),
                  torch.from_numpy(np.ascontiguousarray(sc.features, np.float32)))
    sid = srv.add_model(m)
    cams = [scenes.look_at_camera(p, (0, 0, 0), (0, 1, 0), 64, 64, 50.0) for p in ((0.4, 0.5, 2.4), (-2.0, 0.3, 1.2))]
    views = [(sid, cams[0]), (0, scenes.orbit_camera(3.0, 30.0, 10.0, 64, 64, 50.0)), (sid, cams[1])]
    print(f"model scene ({m.indices.shape[0]} tets) batched with a plain scene, graphs={use_graphs}", flush=True)
    for rep in range(2):
        out = srv.render_views(views, keep_aux=True)
        torch.cuda.synchronize()
        for (s, cam), o in zip(views, out):
            single = srv.render_single(s, cam, aux=True)
            torch.cuda.synchronize()
            _assert_same(o, single, f"replay {rep} scene {s}")
    for cam, o in zip(cams, (out[0], out[2])):
        with torch.no_grad():
            ref = render(_Camera(cam, 4), m, tile_size=4, min_t=0.1)["render"]
        d = (o[0][..., :3].permute(2, 0, 1) - ref).abs().max().item()
        print(f"served view vs render(): max abs {d:.3e}")
        assert d <= 1e-6
    # the two poses see different colours (view-dependent SH)
    assert (out[0][0] - out[2][0]).abs().max().item() > 1e-3


@pytest.mark.parametrize("capacity", [False, True])
def test_mixed_resolution_plan_matches_single_views(capacity):
    """One rm_batch_render over a plan whose views differ in resolution (ragged tile and block
    grids, one view wider than one 8x8-tile block row, one narrower): every view's image,
    distortion and contributor counts equal the single-view path's. With `capacity` the launch
    geometry is larger than the plan (tets, tiles, blocks and workgroups doubled, as a captured
    graph of a larger class would launch)."""
    import ctypes as C
    import torch
    from radiance_meshes_amd._capi import RmBatchInfo, check, lib, ptr, stream_of
    print(f"mixed-resolution plan capacity={capacity}", flush=True)
    srv = _server(False)
    res = [(66, 90), (20, 28), (130, 200), (48, 64), (36, 36)]
    rng = np.random.default_rng(11)
    views = [(k % 3, scenes.orbit_camera(RADIUS[k % 3], rng.uniform(0, 360), rng.uniform(-30, 40), W, H, 50.0))
             for k, (H, W) in enumerate(res)]
    plan_host = torch.empty(srv.plan_info(views).plan_bytes, dtype=torch.uint8, pin_memory=True)
    info = srv.plan_info(views, plan_host)
    plan_dev = plan_host.cuda()
    if capacity:
        launch = RmBatchInfo(n_views=2 * info.n_views, total_tets=2 * info.total_tets,
                             total_tiles=2 * info.total_tiles, total_pixels=2 * info.total_pixels,
                             total_workgroups=2 * info.total_workgroups, plan_bytes=info.plan_bytes,
                             tile=info.tile, n_features=info.n_features,
                             total_bin_blocks=2 * info.total_bin_blocks)
    else:
        launch = info
    cap = 1 << 22
    ws = torch.empty(lib().rm_batch_workspace_size(launch, cap), dtype=torch.uint8, device="cuda")
    npx = int(launch.total_pixels)
    img = torch.empty((npx, 4), device="cuda")
    dist = torch.empty((npx, 5), device="cuda")
    nc = torch.empty((npx,), device="cuda", dtype=torch.int32)
    counts = torch.zeros(2, device="cuda", dtype=torch.int64)
    check(lib().rm_batch_render(ptr(plan_dev), C.byref(launch), cap, ptr(img), ptr(dist), ptr(nc), ptr(counts),
                                ptr(ws), ws.numel(), stream_of()), "rm_batch_render")
    torch.cuda.synchronize()
    assert 0 < int(counts[1]) <= cap
    off = 0
    for (s, c), (H, W) in zip(views, res):
        got = (img[off:off + H * W].view(H, W, 4), dist[off:off + H * W].view(H, W, 5), nc[off:off + H * W].view(H, W))
        _assert_same(got, _single_full(srv, s, c), f"view {H}x{W} scene {s}")
        off += H * W
    assert off == int(info.total_pixels)

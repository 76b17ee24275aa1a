"""rm_composite_bwd_noalloc (the compositing backward with its per-tet rows in a static device
array) against rm_composite_bwd_ws (the same kernels with a caller workspace), through the C ABI:
two calls in a row on one GPU, the second with a smaller input (fewer tets, another tile size and
feature layout), so the second call runs on the rows the first left behind."""
import numpy as np
import pytest

from tests._util import device_inputs, grid_of, jitter, scene, tcam_of


def test_noalloc_rejects_oversize():
    """Sizes beyond the static array are refused before any device work (no GPU needed)."""
    from radiance_meshes_amd import _capi
    L = _capi.lib()
    mx = L.rm_composite_bwd_noalloc_max_tets()
    assert mx == 1 << 19
    cam = _capi.RmCamera()
    rc = L.rm_composite_bwd_noalloc(cam, None, 0, None, None, mx + 1, None, None, 7, None, None, None, None, None,
                                    None, None, None, None)
    assert rc == 5 and b"exceeds the static scratch" in L.rm_last_error()
    rc = L.rm_composite_bwd_noalloc(cam, None, 0, None, None, -1, None, None, 7, None, None, None, None, None,
                                    None, None, None, None)
    assert rc == 1


def _grads(sc, noalloc):
    import torch
    from radiance_meshes_amd._capi import check, lib, ptr, stream_of
    from radiance_meshes_amd.rasterizer import _cam, vertex_and_tile_shader
    verts, idx, feats, _ = device_inputs(sc)
    jit = torch.from_numpy(jitter(sc, 5)).cuda()
    tcam, rg = tcam_of(sc), grid_of(sc)
    H, W = sc.camera.height, sc.camera.width
    nf = feats.shape[1]
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
    check(L.rm_composite_fwd(cam, ptr(sidx), sidx.numel(), ptr(ranges), ptr(idx), ptr(verts), ptr(feats), nf,
                             ptr(jit), ptr(img), ptr(dist), ptr(nc), st), "rm_composite_fwd")
    T = idx.shape[0]
    gv = torch.zeros_like(verts)
    gf = torch.zeros_like(feats)
    if noalloc:
        check(L.rm_composite_bwd_noalloc(cam, ptr(sidx), sidx.numel(), ptr(ranges), ptr(idx), T, ptr(verts),
                                         ptr(feats), nf, ptr(jit), ptr(img), ptr(dist), ptr(nc), ptr(gimg),
                                         ptr(gdist), ptr(gv), ptr(gf), st), "rm_composite_bwd_noalloc")
    else:
        wsb = L.rm_composite_bwd_workspace_size(T)
        ws = torch.full((wsb,), 7, device="cuda", dtype=torch.uint8)
        check(L.rm_composite_bwd_ws(cam, ptr(sidx), sidx.numel(), ptr(ranges), ptr(idx), T, ptr(verts), ptr(feats),
                                    nf, ptr(jit), ptr(img), ptr(dist), ptr(nc), ptr(gimg), ptr(gdist), ptr(gv),
                                    ptr(gf), ptr(ws), wsb, st), "rm_composite_bwd_ws")
    torch.cuda.synchronize()
    return gv.cpu().double().numpy(), gf.cpu().double().numpy()


@pytest.mark.gpu
def test_noalloc_matches_workspace_path():
    """Per-element agreement to 1e-5 of each gradient group's largest value (the two paths run the
    same kernels; only the order of float atomic additions differs), on a 65k-tet scene and then
    on a 12k-tet scene."""
    for kind, tile, nf in (("plain", 16, 7), ("ball", 8, 13)):
        sc = scene(kind, tile, nf)
        print(f"no-alloc backward: {kind} T={len(sc.indices)} tile={tile} F={nf}", flush=True)
        ref = _grads(sc, noalloc=False)
        got = _grads(sc, noalloc=True)
        for name, a, b in (("vertex", got[0], ref[0]), ("feature", got[1], ref[1])):
            scale = np.abs(b).max()
            err = np.abs(a - b).max() / scale
            print(f"  {name}: max abs diff / max |grad| = {err:.2e}")
            assert scale > 0
            assert err <= 1e-5, name

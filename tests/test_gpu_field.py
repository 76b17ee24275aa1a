"""GPU parity of the hash-grid encode and the MFMA field MLP against the FP64 field oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _grid(seed=0, n=20000, L=16, F=2, log2=19, base=16, growth=1.38):
    from radiance_meshes_amd.field import level_table
    res, offs = level_table(L, base, growth, log2)
    rng = np.random.default_rng(seed)
    table = rng.standard_normal((int(offs[-1]), F)).astype(np.float32)
    x = rng.uniform(-0.05, 1.05, (n, 3)).astype(np.float32)   # includes points clamped to the box
    return x, table, res, offs, L, F, log2


@pytest.mark.parametrize("L,F,log2", [(16, 2, 19), (8, 8, 19), (3, 4, 19), (16, 8, 17)])
def test_hashgrid_forward_matches_fp64(oracle_lib, L, F, log2):
    """L * F <= 64 runs the level-major kernel (64 points per workgroup, 20000 = 312 x 64 + 32
    leaves a ragged last group; L = 3 leaves a wave without a level), L * F = 128 the point-major one."""
    import torch
    from radiance_meshes_amd.field import _HashEncode
    O = oracle_lib
    x, table, res, offs, L, F, log2 = _grid(L=L, F=F, log2=log2)
    print(f"hashgrid fwd n={len(x)} levels={L} dense={[int(r) ** 3 <= 2 ** log2 for r in res]}", flush=True)
    out = _HashEncode.apply(torch.from_numpy(x).cuda(), torch.from_numpy(table).cuda(), res, offs, L, F, log2)
    ref = O.hashgrid_fwd(x, table.astype(np.float64), res, offs, F, log2)
    err = np.abs(out.cpu().numpy() - ref).max()
    print(f"max abs err vs FP64 = {err:.3e}")
    assert err <= 1e-5


def test_hashgrid_backward_matches_fp64_adjoint(oracle_lib):
    import torch
    from radiance_meshes_amd.field import _HashEncode
    O = oracle_lib
    x, table, res, offs, L, F, log2 = _grid(seed=1, n=5000)
    g = np.random.default_rng(2).standard_normal((len(x), L * F))
    tt = torch.from_numpy(table).cuda().requires_grad_(True)
    out = _HashEncode.apply(torch.from_numpy(x).cuda(), tt, res, offs, L, F, log2)
    (out * torch.from_numpy(g).float().cuda()).sum().backward()
    ref = O.hashgrid_bwd(x, g, res, offs, F, log2, table.shape[0])
    got = tt.grad.cpu().numpy()
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"table grad rel L2 err vs FP64 = {rel:.3e}")
    assert rel <= 1e-5
    # the adjoint identity <enc(table), g> is linear: check against the forward oracle too
    lhs = (O.hashgrid_fwd(x, table.astype(np.float64), res, offs, F, log2) * g).sum()
    assert abs(lhs - (ref * table).sum()) <= 1e-9 * max(1.0, abs(lhs))


def _mlp_params(seed, n_out=13):
    rng = np.random.default_rng(seed)
    w1 = rng.standard_normal((64, 32)) / np.sqrt(32)
    w2 = rng.standard_normal((64, 64)) / np.sqrt(64)
    w3 = rng.standard_normal((n_out, 64)) / np.sqrt(64)
    b1, b2, b3 = rng.standard_normal(64) * 0.1, rng.standard_normal(64) * 0.1, rng.standard_normal(n_out) * 0.1
    return [a.astype(np.float32) for a in (w1, b1, w2, b2, w3, b3)]


@pytest.mark.parametrize("precision,tol", [(0, 2e-5), (1, 5e-2)])
@pytest.mark.parametrize("n", [16, 1000, 40001, 200003])   # the last: several grid-stride passes
def test_mlp_forward_matches_fp64(oracle_lib, precision, tol, n):
    import torch
    from radiance_meshes_amd.field import _MLP
    O = oracle_lib
    p = _mlp_params(3)
    x = np.random.default_rng(4).standard_normal((n, 32)).astype(np.float32)
    print(f"mlp fwd n={n} precision={'bf16' if precision else 'fp32'}", flush=True)
    y = _MLP.apply(torch.from_numpy(x).cuda(), *[torch.from_numpy(a).cuda() for a in p], precision)
    ref = O.mlp_fwd(x.astype(np.float64), *[a.astype(np.float64) for a in p])
    err = np.abs(y.cpu().numpy() - ref).max() / np.abs(ref).max()
    print(f"max err / max|y| vs FP64 = {err:.3e}")
    assert err <= tol


def test_mlp_small_preactivations(oracle_lib):
    """Pre-activations clustered around 0 (inputs ~ N(0, 1e-4), zero biases): the SELU branch for
    small negative inputs keeps its relative precision (e^h - 1 by its series there), so outputs
    of size ~1e-4 still agree with FP64 to 1e-5 relative."""
    import torch
    from radiance_meshes_amd.field import _MLP
    O = oracle_lib
    p = _mlp_params(8)
    for k in (1, 3, 5):
        p[k] = np.zeros_like(p[k])
    x = (np.random.default_rng(9).standard_normal((4096, 32)) * 1e-4).astype(np.float32)
    print("mlp fwd, pre-activations near 0", flush=True)
    y = _MLP.apply(torch.from_numpy(x).cuda(), *[torch.from_numpy(a).cuda() for a in p], 0)
    ref = O.mlp_fwd(x.astype(np.float64), *[a.astype(np.float64) for a in p])
    rel = np.linalg.norm(y.cpu().numpy() - ref) / np.linalg.norm(ref)
    print(f"|y| ~ {np.abs(ref).mean():.1e}: rel L2 err vs FP64 {rel:.2e}")
    assert rel <= 1e-5


@pytest.mark.parametrize("precision,tol", [(0, 2e-5), (1, 5e-2)])
@pytest.mark.parametrize("n", [3001, 100003])
def test_mlp_backward_matches_fp64(precision, tol, n):
    """n = 100003 exceeds one grid-stride pass of the resident workgroups (at most 5 per CU, 16
    queries per wave), so waves carry their weight-gradient tiles over several query tiles; the
    last tile is ragged."""
    import torch
    from radiance_meshes_amd.field import _MLP
    p = _mlp_params(5)
    x = np.random.default_rng(6).standard_normal((n, 32)).astype(np.float32)
    gy = np.random.default_rng(7).standard_normal((n, 13))
    print(f"mlp bwd n={n} precision={precision}", flush=True)
    xs = torch.from_numpy(x).cuda().requires_grad_(True)
    ps = [torch.from_numpy(a).cuda().requires_grad_(True) for a in p]
    y = _MLP.apply(xs, *ps, precision)
    (y * torch.from_numpy(gy).float().cuda()).sum().backward()
    # FP64 reference gradients with torch autograd on the CPU
    xd = torch.from_numpy(x).double().requires_grad_(True)
    pd = [torch.from_numpy(a).double().requires_grad_(True) for a in p]
    sel = torch.nn.SELU()
    yd = sel(sel(xd @ pd[0].T + pd[1]) @ pd[2].T + pd[3]) @ pd[4].T + pd[5]
    (yd * torch.from_numpy(gy)).sum().backward()
    for name, a, b in [("x", xs.grad, xd.grad)] + [(f"p{i}", ps[i].grad, pd[i].grad) for i in range(6)]:
        rel = (a.cpu().double() - b).norm() / b.norm()
        print(f"grad {name}: rel L2 err vs FP64 = {rel:.3e}")
        assert rel <= tol, name


def test_hashgrid_backward_clustered_points(oracle_lib):
    """Morton-sorted points in a few tight clusters: 256-point chunks share most corner slots at
    every level, so the workgroup aggregation merges many updates per slot before the atomics."""
    import torch
    from radiance_meshes_amd.field import _HashEncode
    from radiance_meshes_amd.layout import morton_codes
    O = oracle_lib
    _, table, res, offs, L, F, log2 = _grid(seed=3, n=1)
    rng = np.random.default_rng(4)
    centres = rng.uniform(0.2, 0.8, (6, 3))
    x = (centres[rng.integers(0, 6, 30000)] + rng.normal(0, 0.01, (30000, 3))).astype(np.float32)
    x = x[np.argsort(morton_codes(x, np.zeros(3), np.ones(3)), kind="stable")]
    g = rng.standard_normal((len(x), L * F))
    print(f"clustered hashgrid bwd n={len(x)}", flush=True)
    tt = torch.from_numpy(table).cuda().requires_grad_(True)
    out = _HashEncode.apply(torch.from_numpy(x).cuda(), tt, res, offs, L, F, log2)
    (out * torch.from_numpy(g).float().cuda()).sum().backward()
    ref = O.hashgrid_bwd(x, g, res, offs, F, log2, table.shape[0])
    got = tt.grad.cpu().numpy()
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"rel L2 err vs FP64 adjoint = {rel:.3e}, max abs = {np.abs(got - ref).max():.3e}")
    assert rel <= 1e-5


def _glue_case(nf=13, seed=5):
    import torch
    from tests._util import device_inputs, scene
    sc = scene("ball", 16, nf)
    verts, idx, feats, _ = device_inputs(sc)
    rng = np.random.default_rng(seed)
    mask = torch.from_numpy(rng.uniform(size=idx.shape[0]) < 0.6).cuda()
    tet_ids = torch.nonzero(mask).reshape(-1).to(torch.int32)
    return verts, idx, mask, tet_ids


def test_field_points_match_torch_contraction():
    import torch
    from radiance_meshes_amd.field import contract, field_points
    verts, idx, mask, tet_ids = _glue_case()
    for scale in (0.3, 2.0):   # 0.3 puts most centroids outside the unit ball (contracted branch)
        got = field_points(verts, idx, tet_ids, scale)
        ref = contract(verts[idx[mask].long()].mean(dim=1), 0.0, scale)
        err = (got - ref).abs().max().item()
        print(f"scale {scale}: max abs diff vs torch = {err:.3e}")
        assert err <= 1e-6   # centroid summation order differs from torch's mean by an ulp
        assert float(got.min()) >= 0.0 and float(got.max()) <= 1.0


@pytest.mark.parametrize("nf", [7, 13])
def test_field_scatter_matches_torch(nf):
    import torch
    from radiance_meshes_amd.field import _FieldScatter, activate
    verts, idx, mask, tet_ids = _glue_case(nf)
    T = idx.shape[0]
    rng = np.random.default_rng(6)
    raw_np = rng.standard_normal((tet_ids.shape[0], nf)).astype(np.float32) * 3
    raw_np[::97, 0] = 60.0          # density argument above the clamp
    raw = torch.from_numpy(raw_np).cuda().requires_grad_(True)
    out = _FieldScatter.apply(raw, tet_ids, T)
    raw_ref = raw.detach().clone().requires_grad_(True)
    ref = torch.zeros((T, nf), device="cuda")
    ref = ref.index_put((tet_ids.long(),), activate(raw_ref, nf))
    rel = ((out - ref).abs() / ref.abs().clamp_min(1.0)).max().item()
    print(f"nf={nf}: max rel diff forward = {rel:.3e}")
    assert rel <= 1e-6
    w = torch.from_numpy(rng.standard_normal((T, nf)).astype(np.float32)).cuda()
    (out * w).sum().backward()
    (ref * w).sum().backward()
    g_rel = ((raw.grad - raw_ref.grad).abs() / raw_ref.grad.abs().clamp_min(1.0)).max().item()
    print(f"nf={nf}: max rel diff backward = {g_rel:.3e}")
    assert g_rel <= 1e-6


def test_hashgrid_backward_workspace_path_matches_atomic_path():
    """The slot-ownership backward of the fine hashed levels and the atomic backward of every level
    give the same table gradient (up to float summation order) on the benchmark's level table."""
    import torch
    from radiance_meshes_amd._capi import check, lib, ptr, stream_of
    x, table, res, offs, L, F, log2 = _grid(seed=8, n=50000)
    g = np.random.default_rng(9).standard_normal((len(x), L * F)).astype(np.float32)
    xd, gd = torch.from_numpy(x).cuda(), torch.from_numpy(g).cuda()
    n = len(x)
    wsb = lib().rm_hashgrid_bwd_workspace_size(n, res.ctypes.data, offs.ctypes.data, L, F, log2)
    print(f"workspace {wsb} bytes, partitioned levels: {[int(r) for r in res if int(r) ** 3 > 2 ** log2 and r >= 256]}",
          flush=True)
    assert wsb > 0
    ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
    a = torch.zeros(table.shape, device="cuda")
    b = torch.zeros(table.shape, device="cuda")
    check(lib().rm_hashgrid_bwd_ws(ptr(xd), n, ptr(gd), res.ctypes.data, offs.ctypes.data, L, F, log2, ptr(a), ptr(ws),
                                   wsb, stream_of()), "rm_hashgrid_bwd_ws")
    check(lib().rm_hashgrid_bwd(ptr(xd), n, ptr(gd), res.ctypes.data, offs.ctypes.data, L, F, log2, ptr(b),
                                stream_of()), "rm_hashgrid_bwd")
    torch.cuda.synchronize()
    rel = ((a - b).norm() / b.norm()).item()
    print(f"rel L2 difference = {rel:.3e}")
    assert rel <= 1e-6

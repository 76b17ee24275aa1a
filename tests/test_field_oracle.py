"""Pins the FP64 field oracle: dense levels reproduce trilinear functions exactly, hashed levels
use the reference's XOR-prime hash bits, and the MLP matches torch's Linear/SELU in float64."""
import numpy as np
import pytest
import torch


def test_dense_level_reproduces_linear_field(oracle_lib):
    O = oracle_lib
    res, offs = O.hashgrid_levels(3, 4, 1.5, 19)       # resolutions 4, 6, 9: all dense
    assert all(r ** 3 <= 2 ** 19 for r in res)
    a, b = np.array([0.3, -1.2, 0.7]), 0.25
    table = np.zeros((int(offs[-1]), 1))
    for l, r in enumerate(res):
        idx = np.arange(r ** 3)
        g = np.stack([idx % r, (idx // r) % r, idx // (r * r)], 1).astype(float)
        table[offs[l]:offs[l + 1], 0] = g @ a + b       # linear in the (unwrapped) grid coordinate
    rng = np.random.default_rng(0)
    x = rng.uniform(0, 1, (500, 3))
    out = O.hashgrid_fwd(x, table, res, offs, 1, 19)
    for l, r in enumerate(res):
        fl, fr = O._grid_cells(x, r)                    # the reference's float32 cell position
        pos = fl + fr
        inner = np.all((fl >= 0) & (fl + 1 <= r - 1), 1)  # no modulo wrap
        assert inner.sum() > 50
        np.testing.assert_allclose(out[inner, l], (pos[inner] @ a) + b, rtol=1e-12, atol=1e-12)


def test_hash_bits_match_reference_formula(oracle_lib):
    O = oracle_lib
    res, offs = O.hashgrid_levels(1, 200, 1.0, 10)     # 200^3 > 2^10: hashed
    table = np.arange(int(offs[-1]), dtype=np.float64)[:, None]
    x = np.array([[0.0, 0.0, 0.0]])                      # pos = 0.5: corners (0|1, 0|1, 0|1), weights 1/8
    out = O.hashgrid_fwd(x, table, res, offs, 1, 10)
    expect = 0.0
    for c in range(8):
        i, j, k = (c >> 2) & 1, (c >> 1) & 1, c & 1
        h = ((i * 1) ^ (j * 2654435761) ^ (k * 805459861)) & ((1 << 10) - 1)
        expect += h / 8
    np.testing.assert_allclose(out[0, 0], expect, rtol=1e-12)


def test_level_resolutions():
    from radiance_meshes_amd.field import level_table
    res, offs = level_table(16, 16, 1.38, 19)
    assert list(res[:4]) == [16, 22, 30, 42]
    assert offs[1] == 16 ** 3 and offs[-1] - offs[-2] == 2 ** 19


def test_mlp_oracle_matches_torch_fp64(oracle_lib):
    O = oracle_lib
    rng = np.random.default_rng(1)
    x = rng.standard_normal((64, 32))
    ps = [rng.standard_normal(s) for s in [(64, 32), (64,), (64, 64), (64,), (13, 64), (13,)]]
    ref = O.mlp_fwd(x, *ps)
    t = [torch.from_numpy(a) for a in ps]
    sel = torch.nn.SELU()
    y = sel(sel(torch.from_numpy(x) @ t[0].T + t[1]) @ t[2].T + t[3]) @ t[4].T + t[5]
    np.testing.assert_allclose(ref, y.numpy(), rtol=1e-12, atol=1e-12)


def test_hashgrid_adjoint_identity(oracle_lib):
    """<enc(table), g> == <table, enc^T(g)> for random tables and gradients (linearity in the table)."""
    O = oracle_lib
    res, offs = O.hashgrid_levels(6, 8, 1.7, 12)       # mixes dense and hashed levels
    rng = np.random.default_rng(3)
    x = rng.uniform(-0.1, 1.1, (300, 3))
    table = rng.standard_normal((int(offs[-1]), 2))
    g = rng.standard_normal((300, 12))
    lhs = (O.hashgrid_fwd(x, table, res, offs, 2, 12) * g).sum()
    rhs = (table * O.hashgrid_bwd(x, g, res, offs, 2, 12, table.shape[0])).sum()
    np.testing.assert_allclose(lhs, rhs, rtol=1e-12)


def test_hashgrid_position_rounding_is_float32(oracle_lib):
    """The cell position follows the reference's float32 arithmetic (x*res + 0.5 rounded twice)."""
    O = oracle_lib
    x = np.array([[0.7071067690849304, 0.1, 0.9]], np.float32)
    fl, fr = O._grid_cells(x, 2006)
    pos = np.float32(np.float32(x[0, 0]) * np.float32(2006)) + np.float32(0.5)
    assert fl[0, 0] == np.floor(pos) and fr[0, 0] == pos - np.floor(pos)


# ------------------------------------------------------------------ the reference model's field


def test_circumspheres_are_equidistant(oracle_lib):
    """Every vertex of each tet lies on its circumsphere; the center is invariant under vertex
    permutation (so no vertex is treated specially)."""
    O = oracle_lib
    rng = np.random.default_rng(1)
    tets = rng.standard_normal((2000, 4, 3))
    c, r = O.circumspheres(tets)
    d = np.linalg.norm(tets - c[:, None, :], axis=2)
    np.testing.assert_allclose(d, np.repeat(r[:, None], 4, 1), rtol=1e-9)
    c2, r2 = O.circumspheres(tets[:, [2, 0, 3, 1]])
    np.testing.assert_allclose(c2, c, rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(r2, r, rtol=1e-9)


def test_contraction_closed_forms(oracle_lib):
    """Inside the unit ball the contraction is the identity and keeps std; outside it keeps the
    direction with |z| = 2 - 1/|x|, and det_13 is the cube root of the Jacobian determinant of
    x -> z (central differences)."""
    O = oracle_lib
    rng = np.random.default_rng(2)
    x = rng.standard_normal((400, 3)) * rng.choice([0.3, 3.0, 40.0], (400, 1))
    std = rng.uniform(0.1, 2.0, 400)
    z, s = O.contract_mean_std(x, std)
    mag = np.linalg.norm(x, axis=1)
    ins = mag <= 1
    assert ins.sum() > 50 and (~ins).sum() > 200
    np.testing.assert_array_equal(z[ins], x[ins])
    np.testing.assert_array_equal(s[ins], std[ins])
    np.testing.assert_allclose(np.linalg.norm(z[~ins], axis=1), 2 - 1 / mag[~ins], rtol=1e-12)
    np.testing.assert_allclose((z[~ins] * x[~ins]).sum(1), np.linalg.norm(z[~ins], axis=1) * mag[~ins], rtol=1e-12)
    for i in np.nonzero(~ins & (mag < 1e5))[0][:40]:
        h = 1e-6 * mag[i]
        J = np.stack([(O.contract_mean_std(x[i] + h * e, 1.0)[0] - O.contract_mean_std(x[i] - h * e, 1.0)[0]) / (2 * h)
                      for e in np.eye(3)], 1)
        np.testing.assert_allclose(s[i] / std[i], np.cbrt(np.linalg.det(J)), rtol=1e-6)


def test_downweight_groups_follow_the_reference_reshape(oracle_lib):
    """The reference scales the [N, L*F] encoding as reshape(-1, F, L) * scale[:, None, :]; the
    oracle's downweight (and the column rule k = i mod L it documents) agree with torch's reshape,
    also for L != F."""
    O = oracle_lib
    rng = np.random.default_rng(3)
    for L, F in [(8, 8), (16, 2), (4, 8)]:
        enc = rng.standard_normal((50, L * F))
        cr = rng.uniform(0, 2, 50)
        s = O.downweight_scale(cr, L, 0.35, 2)
        t = (torch.from_numpy(enc).reshape(-1, F, L) * torch.from_numpy(s).reshape(-1, 1, L)).reshape(-1, L * F)
        np.testing.assert_array_equal(O.downweight(enc, cr, L, F, 0.35, 2), t.numpy())
        cols = np.arange(L * F) % L
        np.testing.assert_array_equal(O.downweight(enc, cr, L, F, 0.35, 2), enc * s[:, cols])


def test_downweight_scale_values(oracle_lib):
    """n = 0 keeps the feature (scale exactly 1: safe_sqrt's tiny floor); the scale falls with n
    and with cr; approx_erf stays within 0.007 of erf (its stated accuracy); a closed form at
    8 n cr scale_multi = 1 gives approx_erf(1) = sqrt(1 - e^(-4/pi))."""
    from scipy.special import erf
    O = oracle_lib
    cr = np.array([0.0, 1e-6, 0.01, 0.3, 2.0, 1e3])
    s = O.downweight_scale(cr, 8, 0.35, 2)
    np.testing.assert_array_equal(s[:, 0], 1.0)
    assert np.all(np.diff(s[1:], axis=1) <= 0) and np.all(np.diff(s[:, 1:], axis=0) <= 0)
    np.testing.assert_array_equal(s[0], 1.0)          # cr = 0: every group at the floor
    xs = np.linspace(-3, 3, 2001)
    assert np.abs(O.approx_erf(xs) - erf(xs)).max() <= 0.007
    one = O.downweight_scale(np.array([1 / (8 * 0.35)]), 2, 0.35, 2)[0, 1]
    assert abs(one - np.sqrt(1 - np.exp(-4 / np.pi))) <= 1e-15


def test_field_query_oracle_on_unit_tets(oracle_lib):
    """A regular tet centered at c with circumradius R: the centroid and circumcenter anchors
    coincide; inside the unit ball x = (n/2 + 1)/2 and cr = R / scaling exactly as stated."""
    O = oracle_lib
    v = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], np.float64) / np.sqrt(3)   # R = 1
    c = np.array([0.2, -0.1, 0.3])
    verts = 0.25 * v + c
    for anchor in ("centroid", "circumcenter"):
        x, cr, anc = O.field_query(verts, np.array([[0, 1, 2, 3]]), np.zeros(3), 2.0, anchor)
        np.testing.assert_allclose(anc[0], c, atol=1e-15)
        np.testing.assert_allclose(x[0], (c / 2.0 / 2 + 1) / 2, atol=1e-15)
        np.testing.assert_allclose(cr[0], 0.25 / 2.0, rtol=1e-14)


def test_reference_level_table():
    """L = 8, base 64, per-level scale 2 (finest 64 * 2^8): growth 2^(8/7) in float32 and the
    floors of base * b^l; sizes min(2^23, res^3): levels 0-1 dense, the rest hashed."""
    from radiance_meshes_amd.ingp import reference_level_table
    res, offs = reference_level_table(8, 64, 64 * 2 ** 8, 23)
    exact = np.floor(64 * 2.0 ** (8 * np.arange(8) / 7))
    b32 = np.float32(np.exp(np.float32(np.log(np.float32(16384)) - np.log(np.float32(64))) / np.float32(7)))
    print("res", res.tolist(), "float64 floors", exact.astype(int).tolist(), "b", b32)
    assert res[0] == 64 and res[-1] in (16383, 16384) and np.all(np.abs(res - exact) <= 1)
    sizes = np.diff(offs)
    assert sizes[0] == 64 ** 3 and sizes[1] == int(res[1]) ** 3 and np.all(sizes[2:] == 2 ** 23)

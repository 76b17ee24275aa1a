"""On-disk mesh format and per-view feature assembly (radiance_meshes_amd.meshio); CPU only."""
import math

import numpy as np
import pytest
import torch

from radiance_meshes_amd import meshio


def _mesh(T=50, V=30, K=15, seed=0):
    rng = np.random.default_rng(seed)
    return dict(vertices=rng.normal(size=(V, 3)).astype(np.float32),
                indices=rng.integers(0, V, (T, 4)).astype(np.int32),
                density=rng.uniform(0, 5, T).astype(np.float32),
                base=rng.uniform(0, 1, (T, 3)).astype(np.float32),
                grd=rng.normal(0, 0.1, (T, 3)).astype(np.float32),
                sh_rest=rng.normal(0, 0.2, (T, K, 3)).astype(np.float32))


def test_save_load_round_trip(tmp_path):
    m = _mesh()
    p = tmp_path / "ckpt.ply"
    meshio.save_mesh_ply(p, m["vertices"], m["indices"], m["density"], m["base"], m["grd"], m["sh_rest"])
    d = meshio.read_ply(p)
    assert set(d) == {"vertex", "tetrahedron"}
    names = set(d["tetrahedron"])
    assert {"indices", "s", "grd_x", "grd_y", "grd_z", "sh_0_r", "sh_0_g", "sh_0_b", "sh_15_b"} <= names
    L = meshio.load_mesh_ply(p)
    np.testing.assert_array_equal(L["vertices"], m["vertices"])
    np.testing.assert_array_equal(L["indices"], m["indices"])
    np.testing.assert_array_equal(L["density"], m["density"])
    np.testing.assert_array_equal(L["grd"], m["grd"])
    np.testing.assert_array_equal(L["sh_rest"], m["sh_rest"])
    np.testing.assert_allclose(L["sh0"] * meshio.SH_C0 + 0.5, m["base"], atol=1e-6)


def test_ascii_and_variable_lists(tmp_path):
    p = tmp_path / "a.ply"
    p.write_text("ply\nformat ascii 1.0\ncomment hand-written\nelement vertex 3\nproperty float x\n"
                 "property float y\nproperty float z\nelement face 2\nproperty list uchar int vertex_indices\n"
                 "property uchar flag\nend_header\n0 0 0\n1 0 0\n0 1 2.5\n3 0 1 2 7\n4 0 1 2 1 9\n")
    d = meshio.read_ply(p)
    np.testing.assert_array_equal(d["vertex"]["z"], [0, 0, 2.5])
    assert [list(v) for v in d["face"]["vertex_indices"]] == [[0, 1, 2], [0, 1, 2, 1]]
    np.testing.assert_array_equal(d["face"]["flag"], [7, 9])


def test_binary_scalar_index_columns(tmp_path):
    """Tetrahedra written as four scalar index columns load the same."""
    m = _mesh(K=0)
    p = tmp_path / "s.ply"
    tet = {f"vertex_index_{k}": m["indices"][:, k] for k in range(4)}
    tet.update({"s": m["density"], "grd_x": m["grd"][:, 0], "grd_y": m["grd"][:, 1], "grd_z": m["grd"][:, 2],
                "sh_0_r": m["base"][:, 0], "sh_0_g": m["base"][:, 1], "sh_0_b": m["base"][:, 2]})
    v = m["vertices"]
    meshio.write_ply(p, {"vertex": {"x": v[:, 0], "y": v[:, 1], "z": v[:, 2]}, "tetrahedron": tet})
    L = meshio.load_mesh_ply(p)
    np.testing.assert_array_equal(L["indices"], m["indices"])
    assert L["sh_rest"].shape == (len(m["indices"]), 0, 3)


def _fib_sphere(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    r = np.sqrt(1 - z * z)
    phi = math.pi * (1 + 5 ** 0.5) * i
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)


def test_sh_basis_orthonormal_and_matches_scipy():
    """The degree 1-3 basis is orthonormal on the sphere (with Y_00 = SH_C0) and equals
    sqrt(2) Im / Re of scipy's complex spherical harmonics (which carry the Condon-Shortley phase)
    for m < 0 / m > 0: the sign convention of the graphics SH tables (Y_1,-1 = -c y)."""
    from scipy import special
    d = _fib_sphere(40000)
    B = meshio.sh_basis(torch.from_numpy(d), 3).numpy()
    full = np.concatenate([np.full((len(d), 1), meshio.SH_C0), B], 1)
    gram = full.T @ full * (4 * math.pi / len(d))
    np.testing.assert_allclose(gram, np.eye(16), atol=2e-3)
    polar = np.arccos(np.clip(d[:, 2], -1, 1))
    azim = np.arctan2(d[:, 1], d[:, 0])
    col = 0
    for l in range(1, 4):
        for m in range(-l, l + 1):
            Y = special.sph_harm_y(l, abs(m), polar, azim)
            if m < 0:
                ref = math.sqrt(2) * Y.imag
            elif m == 0:
                ref = Y.real
            else:
                ref = math.sqrt(2) * Y.real
            np.testing.assert_allclose(B[:, col], ref, atol=1e-10, err_msg=f"l={l} m={m}")
            col += 1


def test_view_features():
    m = _mesh(K=15)
    mesh = dict(vertices=m["vertices"], indices=m["indices"], density=m["density"], grd=m["grd"],
                sh0=meshio.rgb_to_sh(m["base"]).astype(np.float32), sh_rest=m["sh_rest"])
    cam = np.array([0.3, -2.0, 4.0], np.float32)
    f = meshio.view_features(mesh, cam, sh_deg=3, device="cpu").double().numpy()
    cen = m["vertices"][m["indices"]].astype(np.float64).mean(1)
    d = cen - cam
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    B = meshio.sh_basis(torch.from_numpy(d), 3).numpy()
    v = m["vertices"][m["indices"]].astype(np.float64)
    off = v[:, 0] - v.mean(1)                          # the field anchored at the centroid
    v0col = m["base"].astype(np.float64) + (m["grd"].astype(np.float64) * off).sum(1, keepdims=True)
    raw = v0col + np.einsum("tk,tkc->tc", B, m["sh_rest"].astype(np.float64))
    col = np.log1p(np.exp(10 * raw)) / 10
    np.testing.assert_allclose(f[:, 0], m["density"], rtol=1e-6)
    np.testing.assert_allclose(f[:, 1:4], col, atol=2e-6)
    np.testing.assert_allclose(f[:, 4:7], m["grd"], rtol=1e-6)
    f0 = meshio.view_features(mesh, cam, sh_deg=0, device="cpu").double().numpy()
    np.testing.assert_allclose(f0[:, 1:4], np.log1p(np.exp(10 * v0col)) / 10, atol=2e-6)


@pytest.mark.parametrize("anchor", ["centroid", "circumcenter"])
def test_view_features_of_reference_save2ply_fields(oracle_lib, tmp_path, anchor):
    """A PLY holding exactly the fields the reference's save2ply writes (models/base_model.py:107-151:
    s = density, grd = offset_normalize's normalised gradient, sh_0 = RGB2SH(rgb) with rgb the raw
    colour at the anchor, the higher SH bands as they are) renders through view_features to the
    features the reference's activate_output (utils/model_util.py:93-105) forms from the raw field
    outputs with the same anchor -- here the FP64 oracle restatement of activate_output."""
    O = oracle_lib
    from radiance_meshes_amd import scenes
    sc = scenes.ball_scene(n_points=120, seed=4, dtype=np.float32)
    m = {"vertices": sc.vertices, "indices": sc.indices}
    T = len(m["indices"])
    rng = np.random.default_rng(5)
    rgb = rng.uniform(0.1, 0.9, (T, 3))
    grd_raw = rng.normal(0, 0.3, (T, 3))
    sh = rng.normal(0, 0.1, (T, 15, 3))
    dens = rng.uniform(0.5, 3.0, T)
    cam = np.array([0.3, -2.0, 4.0])
    V = m["vertices"].astype(np.float64)
    tets = V[m["indices"]]
    if anchor == "centroid":
        cc = tets.mean(1)
    else:
        cc = np.stack([O.power_of_circumsphere(*tets[t], np.zeros(3))[1] for t in range(T)])
    ref = O.activate_output(cam, dens, rgb, grd_raw, sh, m["indices"], cc, V, 3)
    p = tmp_path / "ckpt.ply"
    meshio.save_mesh_ply(p, m["vertices"], m["indices"], dens, rgb, ref[:, 4:7], sh)   # save2ply's fields
    f = meshio.view_features(meshio.load_mesh_ply(p), cam.astype(np.float32), sh_deg=3, device="cpu",
                             anchor=anchor).double().numpy()
    # flat tets (zero volume) have no circumcenter; compare the others
    vol6 = np.abs(np.einsum("ij,ij->i", np.cross(tets[:, 1] - tets[:, 0], tets[:, 2] - tets[:, 0]),
                            tets[:, 3] - tets[:, 0]))
    ok = vol6 > 1e-9
    assert ok.mean() > 0.99
    err = np.abs(f - ref)[ok].max(0)
    print(f"{anchor}: max abs err per channel {err} over {int(ok.sum())} of {T} tets")
    assert err.max() <= 2e-5
    # the former reading (colour at vertex 0 taken as SH2RGB(sh_0), no gradient term) differs
    shifted = np.abs((ref[ok, 4:7] * (tets[ok, 0] - cc[ok])).sum(1)).max()
    assert shifted > 1e-2

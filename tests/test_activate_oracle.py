"""Pins the FP64 activate_output oracle (oracle.activate_output) on the CPU: closed forms for the
degree-0 colour, the normalised gradient and its zero-radius guard, and the SH part against the
same basis written independently (radiance_meshes_amd.meshio, itself pinned against scipy)."""
import numpy as np
import torch

from radiance_meshes_amd import meshio


def _inputs(T=40, V=60, K=15, seed=0):
    rng = np.random.default_rng(seed)
    return dict(camera_center=rng.normal(size=3) * 3, density=rng.uniform(0, 3, (T, 1)),
                rgb=rng.uniform(0, 1, (T, 3)), grd=rng.normal(0, 0.3, (T, 3)), sh=rng.normal(0, 0.2, (T, K, 3)),
                indices=rng.integers(0, V, (T, 4)), circumcenters=rng.normal(size=(T, 3)),
                vertices=rng.normal(size=(V, 3)))


def test_degree_zero_closed_form(oracle_lib):
    O = oracle_lib
    a = _inputs()
    f = O.activate_output(deg=0, **a)
    tets = a["vertices"][a["indices"]]
    off = tets[:, 0] - a["circumcenters"]
    r = np.linalg.norm(off, axis=1, keepdims=True)
    ng = a["grd"] * a["rgb"].min(1, keepdims=True) / r
    base = a["rgb"] + (ng * off).sum(1, keepdims=True)
    np.testing.assert_allclose(f[:, 0:1], a["density"])
    np.testing.assert_allclose(f[:, 4:7], ng, rtol=1e-14)
    np.testing.assert_allclose(f[:, 1:4], np.log1p(np.exp(10 * base)) / 10, rtol=1e-12)


def test_higher_degrees_add_the_sh_terms(oracle_lib):
    O = oracle_lib
    a = _inputs()
    f0 = O.activate_output(deg=0, **a)
    base = np.log(np.expm1(10 * f0[:, 1:4])) / 10      # invert the softplus
    for deg in (1, 2, 3):
        f = O.activate_output(deg=deg, **a)
        md = a["vertices"][a["indices"]].mean(1) - a["camera_center"]
        d = md / np.linalg.norm(md, axis=1, keepdims=True)
        B = meshio.sh_basis(torch.from_numpy(d), deg).numpy()
        raw = base + np.einsum("tk,tkc->tc", B, a["sh"][:, :B.shape[1]])
        np.testing.assert_allclose(f[:, 1:4], np.log1p(np.exp(10 * raw)) / 10, rtol=1e-10, atol=1e-12)


def test_zero_radius_gives_zero_gradient_feature(oracle_lib):
    O = oracle_lib
    a = _inputs(T=5)
    a["circumcenters"][2] = a["vertices"][a["indices"][2, 0]]
    f = O.activate_output(deg=1, **a)
    assert np.all(f[2, 4:7] == 0)

"""Pins the FP64 error-vote oracle (orc_tet_err) without retyping it: a brute-force ray walk
(membership sampling for which tets a ray crosses and in what order, exact chords from the
intersection routine that tests/test_oracle.py pins against brute force) rebuilds each pixel's
votes, and the whole-image votes must equal the sum of single-pixel votes. Nothing here touches a
GPU."""
import numpy as np
import pytest

from radiance_meshes_amd import scenes


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


def _scene(density_scale):
    sc = scenes.ball_scene(n_points=40, seed=4, size=12, tile_size=4, dtype=np.float64)
    f = sc.features.copy()
    f[:, 0] *= density_scale
    return sc, f


def _bin(O, sc, tile):
    cam = O.make_camera(sc.camera, tile, sc.min_t)
    t, r, vs, cc = O.project(cam, sc.indices, sc.vertices.astype(np.float32))
    keys, vals, ranges = O.bin_tiles(cam, t, r, vs)
    return cam, vals, ranges


def _brute_votes(O, cam, V, idx, f, x, y, n=40_000, tmax=8.0):
    """Votes of pixel (x, y) by walking its ray: tets in order of their first sampled point
    inside, chord from the pinned intersection routine, log-transmittance starting at 1."""
    o, d = O.get_ray(cam, x + 0.5, y + 0.5)
    t = np.linspace(0.0, tmax, n)
    p = o[None] + t[:, None] * d[None]
    first = {}
    for k, q in enumerate(idx):
        v = V[q]
        M = np.stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]], 1)
        lam = np.linalg.solve(M, (p - v[0]).T).T
        inside = (lam >= 0).all(1) & (lam.sum(1) <= 1)
        if inside.any():
            first[k] = t[inside][0]
    A, out = 1.0, {}
    for k in sorted(first, key=first.get):
        hit, e, xx = O.intersect(V[idx[k]].reshape(12), o, d)
        assert hit
        x_ = max(f[k, 0], 0.0) * max(xx - e, 0.0)
        if not x_ > 0:
            continue
        w = np.exp(A) * x_
        out[k] = (w, o + e * d, o + xx * d)
        A -= x_
        if A < -5.5:
            break
    return out


@pytest.mark.parametrize("density_scale", [0.3, 4.0])
def test_single_pixel_votes_match_brute_force_walk(O, density_scale):
    sc, f = _scene(density_scale)
    V, idx = sc.vertices.astype(np.float64), sc.indices
    cam, vals, ranges = _bin(O, sc, 1)      # one tile per pixel: a pixel's list is its tile's
    H, W = sc.camera.height, sc.camera.width
    rng = np.random.default_rng(2)
    checked = 0
    for pix in rng.choice(H * W, 14, replace=False):
        y, x = divmod(int(pix), W)
        r1 = np.zeros_like(ranges)
        r1[pix] = ranges[pix]
        err, cnt, margin = O.tet_err(cam, vals, r1, idx, V, f, np.ones((H, W)), np.full((H, W), 0.25))
        ref = _brute_votes(O, cam, V, idx, f, x, y)
        got = np.nonzero(cnt[:, 0])[0]
        assert set(got) == set(ref), (pix, sorted(got), sorted(ref))
        for k, (w, p1, p2) in ref.items():
            assert cnt[k, 0] == 1
            np.testing.assert_allclose(err[k, 0], w, rtol=1e-9)
            np.testing.assert_allclose(err[k, [1, 2, 12]], [w, w, w], rtol=1e-12)
            np.testing.assert_allclose(err[k, 5], 0.25 * w, rtol=1e-12)
            np.testing.assert_allclose(err[k, 6:9], w * p1, rtol=1e-9, atol=1e-12)
            np.testing.assert_allclose(err[k, 9:12], w * p2, rtol=1e-9, atol=1e-12)
            np.testing.assert_allclose(err[k, 13:16], w * 0.5 * (p1 + p2), rtol=1e-9, atol=1e-12)
            assert cnt[k, 1] == min(int(np.floor(65535.0 * w)), 2 ** 31 - 1)
            checked += 1
        assert np.all(err[:, 3:5] == 0)
    assert checked > 10


def test_termination_after_the_crossing_blend(O):
    """Dense medium: optical depths recovered from the votes starting at log T = 1 equal
    sigma * chord of the pinned intersection; a pixel stops right after the blend that takes
    log T below -5.5, so later tets of its list that its ray crosses get no vote."""
    sc, f = _scene(40.0)
    V, idx = sc.vertices.astype(np.float64), sc.indices
    cam, vals, ranges = _bin(O, sc, 1)
    H, W = sc.camera.height, sc.camera.width
    n_stop = 0
    for pix in range(0, H * W, 5):
        y, x = divmod(pix, W)
        r1 = np.zeros_like(ranges)
        r1[pix] = ranges[pix]
        err, cnt, _ = O.tet_err(cam, vals, r1, idx, V, f, np.ones((H, W)), np.ones((H, W)))
        lst = [int(k) for k in vals[ranges[pix, 0]:ranges[pix, 1]]]
        o, d = O.get_ray(cam, x + 0.5, y + 0.5)
        A, stopped = 1.0, False
        for k in lst:
            hit, e, xx = O.intersect(V[idx[k]].reshape(12), o, d)
            x_ = max(f[k, 0], 0.0) * max(xx - e, 0.0) if hit else 0.0
            if stopped or not x_ > 0:
                assert cnt[k, 0] == 0
                continue
            assert cnt[k, 0] == 1
            np.testing.assert_allclose(err[k, 0] / np.exp(A), x_, rtol=1e-9)
            A -= x_
            stopped = A < -5.5
        n_stop += stopped
    assert n_stop > 3


@pytest.mark.parametrize("tile", [4, 8])
def test_image_votes_are_the_sum_of_pixel_votes(O, tile):
    sc, f = _scene(2.0)
    V, idx = sc.vertices.astype(np.float64), sc.indices
    H, W = sc.camera.height, sc.camera.width
    rng = np.random.default_rng(5)
    pe = rng.uniform(-0.5, 0.5, (H, W))
    se = rng.uniform(0, 1, (H, W))
    cam, vals, ranges = _bin(O, sc, tile)
    err, cnt, _ = O.tet_err(cam, vals, ranges, idx, V, f, pe, se)
    cam1, vals1, ranges1 = _bin(O, sc, 1)
    acc = np.zeros_like(err)
    acc_c = np.zeros_like(cnt)
    for pix in range(H * W):
        r1 = np.zeros_like(ranges1)
        r1[pix] = ranges1[pix]
        e1, c1, _ = O.tet_err(cam1, vals1, r1, idx, V, f, pe, se)
        acc += e1
        acc_c[:, 0] += c1[:, 0]
        acc_c[:, 1] = np.maximum(acc_c[:, 1], c1[:, 1])
    np.testing.assert_allclose(err, acc, rtol=1e-10, atol=1e-13)
    np.testing.assert_array_equal(cnt, acc_c)
    assert cnt[:, 0].sum() > 50


def test_ssim_map_against_independent_filtering():
    """render_err's SSIM map: identity and symmetry, and equality with the same statistics formed
    by scipy's correlation under a zero-padded normalised Gaussian window."""
    import torch
    from scipy import ndimage
    from radiance_meshes_amd.render_err import gaussian_blur, ssim
    rng = np.random.default_rng(1)
    a = rng.uniform(0, 1, (3, 20, 24))
    b = np.clip(a + rng.normal(0, 0.1, a.shape), 0, 1)
    ta, tb = torch.from_numpy(a), torch.from_numpy(b)
    np.testing.assert_allclose(ssim(ta, ta).numpy(), 1.0, atol=1e-12)
    np.testing.assert_allclose(ssim(ta, tb).numpy(), ssim(tb, ta).numpy(), atol=1e-12)
    x = np.arange(11) - 5
    g = np.exp(-x ** 2 / (2 * 1.5 ** 2))
    g /= g.sum()
    k = np.outer(g, g)

    def blur(u):
        return np.stack([ndimage.correlate(c, k, mode="constant") for c in u])

    m1, m2 = blur(a), blur(b)
    v1, v2, c12 = blur(a * a) - m1 ** 2, blur(b * b) - m2 ** 2, blur(a * b) - m1 * m2
    ref = ((2 * m1 * m2 + 1e-4) * (2 * c12 + 9e-4)) / ((m1 ** 2 + m2 ** 2 + 1e-4) * (v1 + v2 + 9e-4))
    np.testing.assert_allclose(ssim(ta, tb).numpy(), ref, atol=1e-12)
    x5 = np.arange(5) - 2
    g5 = np.exp(-x5 ** 2 / (2 * 1.5 ** 2))
    g5 /= g5.sum()
    ref_b = np.stack([ndimage.correlate(c, np.outer(g5, g5), mode="constant") for c in a])
    np.testing.assert_allclose(gaussian_blur(ta).numpy(), ref_b, atol=1e-12)

"""The CPU oracle's rasterizer functions under AddressSanitizer and UndefinedBehaviorSanitizer.

`make -C oracle sanitize` builds oracle/_san/rmesh_oracle_san (oracle/sanitize_main.c linked with
oracle/rmesh_oracle.c, -fsanitize=address,undefined, no recovery): it runs projection, binning,
the FP64 compositing forward and backward, the error votes and the quadrature on one scene file
and writes every output. A sanitizer finding aborts the run (non-zero exit, report on stderr); the
outputs must equal those of the regular library on the same scene (same source, same FP flags).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from tests._util import jitter, scene

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
MAGIC = 0x52534E45


def _write_scene(path, O, sc, rng):
    cam = O.make_camera(sc.camera, sc.tile_size, sc.min_t)
    H, W = sc.camera.height, sc.camera.width
    T, V, F = len(sc.indices), len(sc.vertices), sc.features.shape[1]
    pix = np.sort(rng.choice(H * W, 6, replace=False)).astype(np.int64)
    arrays = dict(
        idx=np.ascontiguousarray(sc.indices, np.int32),
        verts=np.ascontiguousarray(sc.vertices, np.float32),
        feats=np.ascontiguousarray(sc.features, np.float64),
        jit=jitter(sc, 5),
        gimg=rng.standard_normal((H, W, 4)),
        gdist=rng.standard_normal((H, W, 5)) * 0.1,
        perr=rng.uniform(-0.5, 0.5, (H, W)),
        serr=rng.uniform(0, 1, (H, W)),
        pix=pix)
    with open(path, "wb") as f:
        f.write(np.array([MAGIC, T, V, F, H, W, len(pix)], np.int64).tobytes())
        f.write(bytes(cam))
        for a in arrays.values():
            f.write(np.ascontiguousarray(a).tobytes())
    return cam, arrays


def _read_outputs(path, cam, T, V, F, H, W, n_pix):
    buf = open(path, "rb").read()
    off = 0

    def take(dtype, n):
        nonlocal off
        a = np.frombuffer(buf, dtype, count=n, offset=off)
        off += a.nbytes
        return a

    P = int(take(np.int64, 1)[0])
    G = cam.gh * cam.gw
    out = dict(P=P, tiles=take(np.int32, T), rect=take(np.int32, 4 * T).reshape(T, 4),
               vs=take(np.float32, 3 * T).reshape(T, 3), cc=take(np.float32, 3 * T).reshape(T, 3),
               keys=take(np.int64, P), vals=take(np.int32, P), ranges=take(np.int32, 2 * G).reshape(G, 2),
               img=take(np.float64, H * W * 4).reshape(H, W, 4), dist=take(np.float64, H * W * 5).reshape(H, W, 5),
               nc=take(np.int32, H * W).reshape(H, W), margin=take(np.float64, H * W).reshape(H, W),
               gv=take(np.float64, V * 3).reshape(V, 3), gf=take(np.float64, T * F).reshape(T, F),
               err=take(np.float64, T * 16).reshape(T, 16), count=take(np.int32, 2 * T).reshape(T, 2),
               quad=take(np.float64, 4 * n_pix).reshape(n_pix, 4))
    assert off == len(buf)
    return out


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None, reason="needs gcc and make")
@pytest.mark.parametrize("kind,tile,nf", [("ball", 4, 13), ("ball", 8, 7)])
def test_oracle_clean_under_asan_and_ubsan(oracle_lib, tmp_path, kind, tile, nf):
    O = oracle_lib
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True)
    sc = scene(kind, tile, nf)
    rng = np.random.default_rng(21)
    scene_path, out_path = str(tmp_path / "scene.bin"), str(tmp_path / "out.bin")
    cam, a = _write_scene(scene_path, O, sc, rng)
    env = dict(os.environ, OMP_NUM_THREADS="4", ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(ORACLE, "_san", "rmesh_oracle_san"), scene_path, out_path], env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    T, V, F = len(sc.indices), len(sc.vertices), sc.features.shape[1]
    H, W = sc.camera.height, sc.camera.width
    s = _read_outputs(out_path, cam, T, V, F, H, W, len(a["pix"]))

    tiles, rect, vs, cc = O.project(cam, sc.indices, sc.vertices)
    np.testing.assert_array_equal(s["tiles"], tiles)
    np.testing.assert_array_equal(s["rect"], rect)
    np.testing.assert_array_equal(s["vs"].view(np.int32), vs.view(np.int32))
    np.testing.assert_array_equal(s["cc"].view(np.int32), cc.view(np.int32))
    keys, vals, ranges = O.bin_tiles(cam, tiles, rect, vs)
    assert s["P"] == len(vals) > 0
    np.testing.assert_array_equal(s["keys"], keys)
    np.testing.assert_array_equal(s["vals"], vals)
    np.testing.assert_array_equal(s["ranges"], ranges)
    img, dist, nc, margin = O.composite_fwd(cam, vals, ranges, sc.indices, sc.vertices, a["feats"], a["jit"])
    np.testing.assert_allclose(s["img"], img, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(s["dist"], dist, rtol=1e-12, atol=1e-14)
    np.testing.assert_array_equal(s["nc"], nc)
    gv, gf = O.composite_bwd(cam, vals, ranges, sc.indices, sc.vertices, a["feats"], a["jit"], img, dist, nc,
                             a["gimg"], a["gdist"])
    np.testing.assert_allclose(s["gv"], gv, rtol=1e-10, atol=1e-12 * np.abs(gv).max())
    np.testing.assert_allclose(s["gf"], gf, rtol=1e-10, atol=1e-12 * np.abs(gf).max())
    err, cnt, _ = O.tet_err(cam, vals, ranges, sc.indices, sc.vertices, a["feats"], a["perr"], a["serr"])
    np.testing.assert_allclose(s["err"], err, rtol=1e-12, atol=1e-14)
    np.testing.assert_array_equal(s["count"], cnt)
    q = O.quadrature(cam, a["jit"], a["pix"], 12.0, 32, sc.indices, sc.vertices, a["feats"])
    np.testing.assert_allclose(s["quad"], q, rtol=1e-12, atol=1e-14)
    assert np.abs(img).max() > 0 and np.abs(gf).max() > 0 and cnt[:, 0].sum() > 0

#!/usr/bin/env python
"""Training-step benchmark of the MI355X tetrahedral rasterizer + hash-grid field.

One step = one pass of the hot path over one synthetic view: per-tet projection and tile binning,
field query at the visible tets' centroids (hash grid 16 x 2 features, 2^19 entries/level, base 16,
growth 1.38, MLP 32-64-64-13), per-tile compositing of a 1280x840 image, L2 loss against a seeded
target, backward through compositing, MLP and hash grid, and an Adam update of the field and the
vertices. The mesh is 1M points from a seeded 32-component Gaussian mixture plus 8 corners at +-20,
Delaunay-tetrahedralised on the CPU before timing starts.

Prints one JSON line (rank 0): metric fwd+bwd Mrays/s (rays = pixels per step / step time), the
roofline of the dominant kernel (HIP events around its launches inside the timed region; HBM and
vector-ALU issue figures), and a CPU baseline: the same rasterizer restated in C (oracle/, FP64
compositing, OpenMP) on the host's cores, projection and binning of the whole mesh plus the
compositing forward and backward of a sample of tiles, next to the reference's CPU method (per-ray
midpoint quadrature testing every tet at every sample, jaxutil/tetra_quad.py) on a few rays; the
GPU image of the step is compared with the FP64 one on the sampled tiles (image_err_vs_fp64).
The line also carries forward-only frame rates at 1080p/2k/4k on the same mesh, the per-tet
error-vote timing, the BF16-vs-FP64 field report, the GPU
Delaunay tetrahedralisation of the benchmark's points (wall time per call, set difference with the
benchmark mesh, scipy/Qhull on a 100k-point sample as the CPU reference) and the serving leg: three device-resident meshes (250k, 1M, 3M points) served a Poisson trace of
mixed-resolution views (256x256 .. 1920x1080, 200/s for --serve-seconds, then a saturating
1000/s trace) batched by resolution bucket and per frame, with p50/p99 latency, frames/s and the
batched-vs-per-frame image difference.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
N_CU, CLOCK_HZ = 256, 2.4e9    # MI355X compute units and peak engine clock
FP32_VALU_PEAK_TFLOPS = 78.6   # non-packed FP32 vector FMA rate (packed FP32 ops are not used)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--points", type=int, default=1_000_000)
    p.add_argument("--width", type=int, default=1280)
    p.add_argument("--height", type=int, default=840)
    p.add_argument("--tile", type=int, default=4)
    p.add_argument("--sweep-tiles", default="", help="comma-separated extra tile sizes timed on the same mesh "
                   "(one JSON line each on stderr)")
    p.add_argument("--precision", default="fp32", choices=["fp32", "bf16"])
    p.add_argument("--intersect", default="fp64", choices=["fp64", "fp32"])
    p.add_argument("--cpu-baseline-tiles", type=int, default=4096,
                   help="tiles of the view composited (forward + backward) by the CPU restatement")
    p.add_argument("--cpu-baseline-rays", type=int, default=16, help="rays of the quadrature (reference method)")
    p.add_argument("--cpu-baseline-samples", type=int, default=64)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-forward-fps", action="store_true",
                   help="skip the forward-only 1080p/2k/4k frame rates on the benchmark mesh")
    p.add_argument("--no-precision-report", action="store_true",
                   help="skip the BF16-vs-FP64 field MLP report on the benchmark mesh")
    p.add_argument("--no-layout", action="store_true", help="keep the Delaunay order of vertices and tets")
    p.add_argument("--no-delaunay", action="store_true", help="skip the GPU Delaunay leg")
    p.add_argument("--delaunay-only", action="store_true",
                   help="only triangulate the benchmark's GMM points (float32) on the GPU, 3 timed calls "
                        "after one warm-up, and print one JSON line (for profiling; no CPU mesh needed)")
    p.add_argument("--delaunay-cpu-points", type=int, default=100_000,
                   help="points of the scipy (Qhull) CPU reference sample in the Delaunay leg")
    p.add_argument("--mesh-cache", default="", help="npz file to load the mesh from / save it to")
    p.add_argument("--profile-steps", type=int, default=0,
                   help="run only this many steps after warmup, without the CPU baseline (for rocprof)")
    p.add_argument("--serve-seconds", type=float, default=10.0,
                   help="length of the Poisson request trace of the serving leg (0: leg off)")
    p.add_argument("--serve-rate", type=float, default=200.0, help="request arrival rate (1/s)")
    p.add_argument("--serve-points", default="250000,1000000,3000000", help="points of the served meshes")
    p.add_argument("--serve-slots", type=int, default=8, help="views per batch of a resolution bucket")
    p.add_argument("--serve-hbm-gib", type=float, default=160.0, help="HBM budget of the server (GiB)")
    p.add_argument("--serve-only", action="store_true", help="skip the training-step leg")
    p.add_argument("--serve-warmup-batches", type=int, default=40,
                   help="random batches per resolution rendered before the trace (graph capture per class)")
    p.add_argument("--serve-saturate-rate", type=float, default=1000.0,
                   help="also replay a short trace at this rate (1/s) to measure each mode's capacity")
    p.add_argument("--serve-saturate-seconds", type=float, default=2.0)
    p.add_argument("--calibrate-copy", action="store_true",
                   help="only run a 256 MiB device-to-device copy 5 times (calibrates the HBM byte counters)")
    return p.parse_args()


def _mesh(n_points, layout=True):
    """GMM points (seed 0) -> Delaunay tets (-> Morton layout); runs in a worker process."""
    from radiance_meshes_amd import scenes
    t0 = time.time()
    pts = scenes.gmm_points(n_points, 0)
    tets = scenes.tetrahedralize(pts)
    if layout:
        from radiance_meshes_amd.layout import spatial_layout
        pts, tets, _, _ = spatial_layout(pts, tets)
    return n_points, pts.astype(np.float32), tets, time.time() - t0


def build_workload(args):
    from radiance_meshes_amd import scenes
    cam = scenes.orbit_camera(6.0, 30.0, 20.0, args.width, args.height, 50.0)
    if args.mesh_cache and os.path.exists(args.mesh_cache):
        z = np.load(args.mesh_cache)
        if int(z["points"]) == args.points and bool(z["layout"]) == (not args.no_layout):
            log(f"[bench] mesh from {args.mesh_cache}")
            return z["pts"], z["tets"], cam
    t0 = time.time()
    pts = scenes.gmm_points(args.points, 0)
    log(f"[bench] {len(pts)} points; tetrahedralising on the CPU ...")
    tets = scenes.tetrahedralize(pts)
    log(f"[bench] {len(tets)} tets ({time.time() - t0:.1f} s)")
    if not args.no_layout:
        from radiance_meshes_amd.layout import spatial_layout
        pts, tets, _, _ = spatial_layout(pts, tets)
    pts = pts.astype(np.float32)
    if args.mesh_cache:
        np.savez(args.mesh_cache, pts=pts, tets=tets, points=args.points, layout=not args.no_layout)
    return pts, tets, cam


def count_evaluations(indices, vertices, tcam, rg, jitter, sidx, ranges, feats, out=None):
    """Sum over pixels of n_contrib: (pixel, tet) evaluations of one forward pass (the image is
    left in out["img"] when out is given)."""
    import torch
    from radiance_meshes_amd._capi import check, lib, ptr, stream_of
    from radiance_meshes_amd.rasterizer import _cam
    H, W = rg.image_height, rg.image_width
    img = torch.empty((H, W, 4), device="cuda")
    dist = torch.empty((H, W, 5), device="cuda")
    nc = torch.empty((H, W), device="cuda", dtype=torch.int32)
    check(lib().rm_composite_fwd(_cam(tcam, rg), ptr(sidx), sidx.numel(), ptr(ranges), ptr(indices), ptr(vertices),
                                 ptr(feats), feats.shape[1], ptr(jitter), ptr(img), ptr(dist), ptr(nc),
                                 stream_of()), "rm_composite_fwd")
    if out is not None:
        out["img"] = img
    return int(nc.sum().item())


def time_tet_err(indices, vertices, tcam, rg, sidx, ranges, feats, reps=3):
    """Per-tet error votes (render_err's calc_tet_err) of the benchmark view, timed with HIP events
    on the stream they run on; pixel and SSIM errors are seeded uniform fields."""
    import torch
    from radiance_meshes_amd.render_err import calc_tet_err
    H, W = rg.image_height, rg.image_width
    g = torch.Generator(device="cuda").manual_seed(2)
    pe = torch.rand((H, W), device="cuda", generator=g) - 0.5
    se = torch.rand((H, W), device="cuda", generator=g)
    T = indices.shape[0]
    err = torch.zeros((T, 16), device="cuda")
    cnt = torch.zeros((T, 2), device="cuda", dtype=torch.int32)
    calc_tet_err(sidx, ranges, indices, vertices, feats, rg, tcam, pe, se, err, cnt)   # warm-up
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * reps)]
    st = torch.cuda.current_stream()
    for i in range(reps):
        err.zero_()
        cnt.zero_()
        ev[2 * i].record(st)
        calc_tet_err(sidx, ranges, indices, vertices, feats, rg, tcam, pe, se, err, cnt)
        ev[2 * i + 1].record(st)
    torch.cuda.synchronize()
    ms = [ev[2 * i].elapsed_time(ev[2 * i + 1]) for i in range(reps)]
    voted_tets = int((cnt[:, 0] > 0).sum().item())
    # algorithmic output: one 64-byte vote row and one 8-byte count row per voted tet
    return {"mean_ms": round(sum(ms) / reps, 4), "voted_tets": voted_tets, "votes": int(cnt[:, 0].sum().item())}


def delaunay_leg(args, pts, tets, reps=3):
    """The triangulation step of update_triangulation (models/ingp_color.py:523-544) on the
    benchmark's points: rm_delaunay, wall time per call including the host's read of the tet count
    (the reference's gdel3d/scipy calls also hand back host arrays); the result compared as a set
    with the benchmark mesh (scipy/Qhull on the float64 points before their float32 rounding);
    scipy timed on a sample of the same points as the CPU reference (the reference's
    high_precision branch)."""
    import torch
    from radiance_meshes_amd.delaunay import delaunay_tets
    P = torch.from_numpy(np.ascontiguousarray(pts, np.float32)).cuda()
    t, info = delaunay_tets(P, return_info=True)   # warm-up: workspace, code objects
    torch.cuda.synchronize()
    wall = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        t, info = delaunay_tets(P, return_info=True)
        torch.cuda.synchronize()
        wall.append(time.perf_counter() - t0)
    ms = 1e3 * sum(wall) / reps

    def canon(a):
        a = np.sort(np.asarray(a, np.int64), axis=1)
        return np.ascontiguousarray(a).view(np.dtype((np.void, 32))).ravel()
    g, r = canon(t.cpu().numpy()), canon(tets)
    gpu_only = int(np.setdiff1d(g, r, assume_unique=True).size)
    ref_only = int(np.setdiff1d(r, g, assume_unique=True).size)
    del P, t
    torch.cuda.empty_cache()
    from scipy.spatial import Delaunay
    n_cpu = min(args.delaunay_cpu_points, len(pts))
    sample = np.asarray(pts[np.random.default_rng(0).choice(len(pts), n_cpu, replace=False)], np.float64)
    t0 = time.perf_counter()
    Delaunay(sample)
    cpu_s = time.perf_counter() - t0
    log(f"[bench] delaunay: {ms:.1f} ms per call on {len(pts)} points; scipy {cpu_s:.2f} s on {n_cpu}")
    return {"points": int(len(pts)), "tets": int(info["n_tets"]), "ms_per_call": round(ms, 2),
            "points_per_s": round(len(pts) / (ms * 1e-3)), "info": info,
            "vs_benchmark_mesh": {"gpu_only_tets": gpu_only, "mesh_only_tets": ref_only,
                                  "note": "mesh: scipy/Qhull on the float64 points; GPU: their float32 values"},
            "cpu_baseline": {"kind": "reference", "what": "scipy.spatial.Delaunay (Qhull), one thread",
                             "points": n_cpu, "seconds": round(cpu_s, 3),
                             "points_per_s": round(n_cpu / cpu_s)}}


def tet_err_line(votes, P, G, N, workload, tile, precision):
    """The error-vote kernel's figures: algorithmic bytes, achieved GB/s and, from the profile's
    SQ_INSTS_VALU pass, its VALU issue fraction (the kernel is bound by vector-ALU work)."""
    algo = 96 * P + 8 * G + 8 * N + 72 * votes["voted_tets"]
    out = {**votes, "algorithmic_bytes": algo, "achieved_gbs": round(algo / (votes["mean_ms"] * 1e-3) / 1e9, 2)}
    insts = measured_traffic(workload, tile, precision, "tet_err", key="valu_insts")
    if insts:
        out["valu"] = {"insts_per_launch": insts,
                       "frac": round(insts / (N_CU * CLOCK_HZ * votes["mean_ms"] * 1e-3), 4)}
    return out


def measured_traffic(workload, tile, precision, span, key="traffic"):
    """HBM bytes per launch of `span`'s kernel from the rocprofv3 FETCH_SIZE/WRITE_SIZE passes
    recorded for this workload (profiles/traffic_*.json), or None when none was recorded; with
    key="valu_insts", the span's VALU wave-instructions per launch (SQ_INSTS_VALU pass)."""
    here = os.path.dirname(os.path.abspath(__file__))
    path = os.path.join(here, "profiles", "traffic_gmm1M_1280x840.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("workload") != workload or t.get("tile") != tile or t.get("precision") != precision:
        return None
    return t.get(key, {}).get(span)


def run_config(args, pts, tets, cam, tile):
    import torch
    from radiance_meshes_amd import field as fieldmod, timing
    from radiance_meshes_amd.rasterizer import (AlphaBlendTiledRender, RenderGrid, distortion_loss,
                                                set_intersect_precision, vertex_and_tile_shader)
    dev = "cuda"
    torch.manual_seed(0)
    vertices = torch.nn.Parameter(torch.from_numpy(pts).to(dev))
    indices = torch.from_numpy(tets).to(dev)
    T = indices.shape[0]
    H, W = args.height, args.width
    tcam = {k: (v.to(dev) if hasattr(v, "to") else v) for k, v in cam.tcam(tile, 0.1).items()}
    rg = RenderGrid(H, W, tile, tile)
    field = fieldmod.Field(n_features=13, precision=args.precision, scale=2.0, device=dev)
    target = torch.rand((H, W, 3), generator=torch.Generator().manual_seed(1)).to(dev)
    jitter = torch.full((H, W, 2), 0.5, device=dev)
    opt = torch.optim.Adam([{"params": field.parameters(), "lr": 1e-3},
                            {"params": [vertices], "lr": 1e-6}], fused=True)
    set_intersect_precision(args.intersect == "fp64")
    stats = {}

    def step():
        opt.zero_grad(set_to_none=True)
        sidx, ranges, vs, cc, mask, rect = vertex_and_tile_shader(indices, vertices, tcam, rg)
        feats = field.features(vertices, indices, mask)
        img, dist, _ = AlphaBlendTiledRender.apply(sidx, ranges, indices, vertices, feats, rg, tcam, jitter)
        loss = ((img[..., :3] - target) ** 2).mean() + 0.01 * distortion_loss(dist).mean()
        loss.backward()
        opt.step()
        stats["P"] = sidx.numel()
        stats["mask"] = mask
        stats["last"] = (sidx, ranges, feats.detach())
        return loss

    log("[bench] warmup")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    steps = args.profile_steps or args.steps
    log(f"[bench] timing {steps} steps")
    timing.reset()
    timing.enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    timing.enable(False)
    kern = timing.summary()
    ms = dt / steps * 1e3
    mrays = H * W / (dt / steps) / 1e6
    n_vis = int(stats["mask"].sum().item())
    # one more forward, without gradients, on the parameters as the timed steps left them: the
    # evaluation count and what the CPU leg checks (binning, features, image) all come from it
    with torch.no_grad():
        sidx_l, ranges_l, _, _, mask_l, _ = vertex_and_tile_shader(indices, vertices, tcam, rg)
        feats_l = field.features(vertices, indices, mask_l)
    last_img = {}
    n_eval = count_evaluations(indices, vertices.detach(), tcam, rg, jitter, sidx_l, ranges_l, feats_l, out=last_img)
    P = int(stats["P"])
    G = rg.grid_height * rg.grid_width
    N = H * W
    votes = time_tet_err(indices, vertices.detach(), tcam, rg, sidx_l, ranges_l, feats_l, reps=3)
    V = vertices.shape[0]

    # algorithmic bytes per launch (FP32 values, int32 indices)
    L, Fh = field.grid.n_levels, field.grid.n_feat
    algo = {
        "project_fwd": 16 * T + 12 * V + 44 * T,
        "bin_count": 4 * T,
        "bin_sort": (24 * T) + (12 * P + 24 * P + 8 * P) + 8 * G,
        "composite_fwd": 96 * P + 8 * G + 8 * N + 40 * N,
        "composite_bwd": 96 * P + 8 * G + 8 * N + 40 * N + 36 * N + 76 * P,
        "hashgrid_fwd": n_vis * (12 + L * 8 * Fh * 4 + 4 * L * Fh),
        "hashgrid_bwd": n_vis * (12 + 4 * L * Fh + L * 8 * Fh * 4),
        "mlp_fwd": n_vis * (128 + 52),
        "mlp_bwd": n_vis * (128 + 52 + 128),
    }
    dominant = max(kern, key=lambda k: kern[k]["total_ms"])
    d = kern[dominant]
    achieved = algo.get(dominant, 0) / (d["mean_ms"] * 1e-3) / 1e9
    workload = f"gmm_{args.points}pts_{W}x{H}_fwd_bwd_train_step"
    roofline = {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": measured_traffic(workload, tile, args.precision, dominant),
                "mean_ms": round(d["mean_ms"], 4),
                "algorithmic_bytes": int(algo.get(dominant, 0))}
    # the compositing kernels are bound by vector-ALU issue, not HBM: wave-instructions per launch
    # (rocprofv3 SQ_INSTS_VALU pass of this workload) over the issue peak of one VALU
    # wave-instruction per CU per cycle at the peak engine clock
    insts = measured_traffic(workload, tile, args.precision, dominant, key="valu_insts")
    if insts:
        peak = N_CU * CLOCK_HZ * d["mean_ms"] * 1e-3
        roofline["valu"] = {"insts_per_launch": insts, "issue_peak_per_s": N_CU * CLOCK_HZ,
                            "frac": round(insts / peak, 4)}
    total_algo = sum(algo[k] * kern[k]["launches"] / steps for k in algo if k in kern)
    breakdown = {k: round(v["total_ms"] / steps, 4) for k, v in sorted(kern.items(), key=lambda kv: -kv[1]["total_ms"])}

    # what the CPU leg compares against: this step's binning, features, image and loss target
    last = {"sidx": sidx_l.cpu().numpy(), "ranges": ranges_l.cpu().numpy(), "feats": feats_l.cpu().numpy(),
            "img": last_img["img"].cpu().numpy(), "target": target.cpu().numpy(),
            "vertices": vertices.detach().cpu().numpy()}
    return {"last": last, "value": round(mrays, 3), "ms_per_step": round(ms, 3), "steps": steps, "roofline": roofline,
            "hbm_frac_all_kernels": round(total_algo / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
            "breakdown_ms": breakdown,
            "tet_err": tet_err_line(votes, P, G, N, workload, tile, args.precision),
            "config": {"workload": workload, "points": int(V),
                       "tets": int(T), "visible_tets": n_vis, "tile_pairs": P, "evaluations": n_eval,
                       "width": W, "height": H, "tile": tile,
                       "field": f"hashgrid L16 F2 2^19 base16 x1.38 + mlp 32-64-64-13 {args.precision}",
                       "intersect": args.intersect, "layout": "delaunay" if args.no_layout else "morton",
                       "parallelism": "single_gpu"}}


def precision_report(pts, tets, cam, tile):
    """Field MLP with BF16 operands (FP32 accumulation) against an FP64 evaluation of the same
    weights on the same hash encodings, over every visible tet centroid of the benchmark view;
    images composited in FP32 from either feature set. The hash table is drawn U(-1, 1) (seed 3)
    so that the encodings, and the field, vary across the mesh; MLP weights N(0, 1/fan_in)."""
    import torch
    import torch.nn.functional as Fn
    from radiance_meshes_amd import field as fieldmod
    from radiance_meshes_amd._capi import check, lib, ptr, stream_of
    from radiance_meshes_amd.rasterizer import RenderGrid, _cam, vertex_and_tile_shader
    dev = "cuda"
    vertices = torch.from_numpy(pts).to(dev)
    indices = torch.from_numpy(tets).to(dev)
    H, W = cam.height, cam.width
    tcam = {k: (v.to(dev) if hasattr(v, "to") else v) for k, v in cam.tcam(tile, 0.1).items()}
    rg = RenderGrid(H, W, tile, tile)
    fld = fieldmod.Field(n_features=13, precision="fp32", scale=2.0, device=dev)
    with torch.no_grad():
        fld.grid.table.copy_(torch.rand(fld.grid.table.shape, generator=torch.Generator().manual_seed(3)) * 2 - 1)
        sidx, ranges, vs, cc, mask, rect = vertex_and_tile_shader(indices, vertices, tcam, rg)
        tet_ids = torch.nonzero(mask).reshape(-1).to(torch.int32)
        h = fld.grid(fieldmod.field_points(vertices, indices, tet_ids, fld.scale))
        m = fld.mlp
        raw32 = m(h)
        m.precision = 1
        raw16 = m(h)
        m.precision = 0
        w = [t.double() for t in (m.w1, m.b1, m.w2, m.b2, m.w3, m.b3)]
        raw64 = torch.empty((h.shape[0], 13), dtype=torch.float64, device=dev)
        for a in range(0, h.shape[0], 1 << 20):
            z = Fn.selu(h[a:a + (1 << 20)].double() @ w[0].T + w[1])
            z = Fn.selu(z @ w[2].T + w[3])
            raw64[a:a + (1 << 20)] = z @ w[4].T + w[5]

        def params(raw):
            d = torch.exp(raw[:, 0].double()), torch.exp(raw64[:, 0])
            dens_rel = ((d[0] - d[1]).abs() / d[1]).float()
            col = (raw[:, 1:].double() - raw64[:, 1:]).abs().float()
            return {"density_rel_max": float(dens_rel.max()), "density_rel_mean": float(dens_rel.mean()),
                    "colour_abs_max": float(col.max()), "colour_abs_mean": float(col.mean())}

        def image(raw):
            feats = torch.zeros((indices.shape[0], 13), device=dev)
            feats[tet_ids.long()] = fieldmod.activate(raw.float(), 13)
            img = torch.empty((H, W, 4), device=dev)
            dist = torch.empty((H, W, 5), device=dev)
            nc = torch.empty((H, W), device=dev, dtype=torch.int32)
            jit = torch.full((H, W, 2), 0.5, device=dev)
            check(lib().rm_composite_fwd(_cam(tcam, rg), ptr(sidx), sidx.numel(), ptr(ranges), ptr(indices),
                                         ptr(vertices), ptr(feats), 13, ptr(jit), ptr(img), ptr(dist), ptr(nc),
                                         stream_of()), "rm_composite_fwd")
            return img[..., :3].double()

        ref = image(raw64)
        out = {"queries": int(h.shape[0]), "table_init": "U(-1,1) seed 3", "reference": "fp64 mlp, same weights"}
        for name, raw in (("bf16", raw16), ("fp32", raw32)):
            e = (image(raw) - ref).abs()
            out[name] = {**params(raw), "pixel_abs_max": float(e.max()), "pixel_abs_mean": float(e.mean())}
    torch.cuda.synchronize()
    return out


def full_features(pts, tets):
    """13-channel features of every tet from the training step's field (default initialisation)."""
    import torch
    from radiance_meshes_amd import field as fieldmod
    fld = fieldmod.Field(n_features=13, precision="fp32", scale=2.0, device="cuda")
    with torch.no_grad():
        v = torch.from_numpy(pts).cuda()
        i = torch.from_numpy(tets).cuda()
        f = fld.features(v, i, torch.ones((len(tets),), dtype=torch.bool, device="cuda"))
    return f.cpu().numpy()


def calibrate_copy():
    """A plain 256 MiB device-to-device copy, 5 launches: under `rocprofv3 --pmc FETCH_SIZE` /
    `WRITE_SIZE` it gives the counters' reading for 256 MiB read + 256 MiB written."""
    import torch
    src = torch.empty(64 << 20, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": src.numel() * 4, "launches": 5}), flush=True)


class BusySampler:
    """Mean of the GPU activity the driver reports (sysfs gpu_busy_percent of the card torch uses,
    matched by PCI bus id when several cards are listed), sampled every `period` s on a thread
    between start() and stop(); None where the file is not readable."""

    def __init__(self, period=0.01):
        import glob
        import threading
        self.period, self.samples, self._stop = period, [], threading.Event()
        self.path = None
        paths = sorted(glob.glob("/sys/class/drm/card*/device/gpu_busy_percent"))
        if len(paths) > 1:
            try:
                import torch
                pr = torch.cuda.get_device_properties(0)
                bus = f"{pr.pci_domain_id:04x}:{pr.pci_bus_id:02x}:{pr.pci_device_id:02x}"
                paths = [q for q in paths if os.path.realpath(os.path.dirname(q)).split("/")[-1].startswith(bus)]
            except Exception:
                paths = []
        if len(paths) == 1:
            try:
                open(paths[0]).read()
                self.path = paths[0]
            except OSError:
                self.path = None
        self._thread = None

    def _run(self):
        while not self._stop.is_set():
            try:
                self.samples.append(float(open(self.path).read()))
            except (OSError, ValueError):
                pass
            self._stop.wait(self.period)

    def start(self):
        import threading
        self.samples = []
        self._stop.clear()
        if self.path:
            self._thread = threading.Thread(target=self._run, daemon=True)
            self._thread.start()
        return self

    def stop(self):
        self._stop.set()
        if self._thread is not None:
            self._thread.join()
        return round(float(np.mean(self.samples)), 1) if self.samples else None


def serve_leg(args, meshes):
    """Batched serving against per-frame rendering on one Poisson trace (forward only, fp32).
    meshes: [(n_points, pts, tets)], registered as device-resident scenes with seeded features."""
    import torch
    from radiance_meshes_amd import scenes
    from radiance_meshes_amd.serving import RenderServer, Request, poisson_trace
    srv = RenderServer(tile=args.tile, n_features=13, min_t=0.1,
                       hbm_budget_bytes=int(args.serve_hbm_gib * (1 << 30)), max_slots=args.serve_slots,
                       use_graphs=True)
    tets_of = []
    for n, pts, tets in meshes:
        feats = scenes.tet_features(len(tets), 1, 13)
        srv.add_scene(torch.from_numpy(pts), torch.from_numpy(tets), torch.from_numpy(feats))
        tets_of.append(len(tets))
    resolutions = [(256, 256), (512, 512), (768, 1024), (1080, 1920)]   # (H, W)
    # warm-up outside the trace: every bucket created (graph captured), every kernel loaded
    wrng = np.random.default_rng(1)
    for H, W in resolutions:
        log(f"[bench] serving warm-up {W}x{H}: bucket, graph capture per capacity class, batched and single views")
        views = [(s, scenes.orbit_camera(6.0, 10.0 * s, 10.0, W, H, 50.0)) for s in range(len(meshes))]
        srv.render_views(views)
        for s, c in views:
            srv.render_single(s, c)
        b = srv.buckets[(H, W)]
        # batches of the trace's composition (1..slots views of random scenes) capture the classes
        # the trace will use, so no capture (and its device-wide synchronisation) happens inside it
        for _ in range(args.serve_warmup_batches):
            n = int(wrng.integers(1, b.slots + 1))
            srv.render_views([(int(wrng.integers(len(meshes))),
                               scenes.orbit_camera(6.0, wrng.uniform(0, 360), wrng.uniform(-20, 40), W, H, 50.0))
                              for _ in range(n)])
        torch.cuda.synchronize()
        log(f"[bench]   {W}x{H}: {b.slots} slots, pair capacity {b.pair_cap}, {b.bytes / 2**30:.1f} GiB, "
            f"{len(b.graphs)} capacity classes captured")
    torch.cuda.synchronize()
    trace = poisson_trace(len(meshes), resolutions, args.serve_rate, args.serve_seconds, seed=0, radius=6.0)
    log(f"[bench] serving: {len(trace)} requests over {args.serve_seconds} s at {args.serve_rate}/s")
    keep = lambda r: r.rid % 25 == 0   # noqa: E731  (results kept for the batched/per-frame comparison)
    runs = {}
    busy = {}
    sampler = BusySampler()
    log(f"[bench] GPU activity from {sampler.path}" if sampler.path else "[bench] GPU activity: no readable sysfs file")
    for mode in ("batched", "per_frame"):
        reqs = [Request(r.rid, r.scene, r.camera, r.arrival) for r in trace]
        sampler.start()
        runs[mode] = (srv.serve(reqs, mode=mode, keep=keep,
                                progress=lambda d, n, t, m=mode: log(f"[bench]   {m}: {d}/{n} done at {t:.1f} s")),
                      reqs)
        busy[mode] = sampler.stop()
        log(f"[bench] {mode}: {json.dumps(runs[mode][0])}, GPU busy {busy[mode]}%")
    # every kept frame: batched result against the per-frame result of the same request
    max_err, n_cmp = 0.0, 0
    for rb, rp in zip(runs["batched"][1], runs["per_frame"][1]):
        if rb.image is not None and rp.image is not None:
            max_err = max(max_err, float((rb.image - rp.image).abs().max().item()))
            n_cmp += 1
    sb, sp = runs["batched"][0], runs["per_frame"][0]
    saturated = None
    if args.serve_saturate_rate > 0:
        strace = poisson_trace(len(meshes), resolutions, args.serve_saturate_rate, args.serve_saturate_seconds,
                               seed=5, radius=6.0)
        log(f"[bench] saturating trace: {len(strace)} requests over {args.serve_saturate_seconds} s "
            f"at {args.serve_saturate_rate}/s")
        saturated = {"requests": len(strace), "rate_hz": args.serve_saturate_rate,
                     "seconds": args.serve_saturate_seconds}
        for mode in ("batched", "per_frame"):
            reqs = [Request(r.rid, r.scene, r.camera, r.arrival) for r in strace]
            sampler.start()
            st = srv.serve(reqs, mode=mode, keep=lambda r: False,
                           progress=lambda d, n, t, m=mode: log(f"[bench]   saturated {m}: {d}/{n} done at {t:.1f} s"))
            b = sampler.stop()
            log(f"[bench] saturated {mode}: {json.dumps(st)}, GPU busy {b}%")
            saturated[mode] = {k: round(st[k], 3) for k in ("fps", "p50_ms", "p99_ms", "mean_batch")}
            saturated[mode]["gpu_busy_pct"] = b
        saturated["fps_ratio"] = round(saturated["batched"]["fps"] / saturated["per_frame"]["fps"], 3)
    return {"saturated": saturated, "fps_batched": round(sb["fps"], 2), "fps_per_frame": round(sp["fps"], 2),
            "fps_ratio": round(sb["fps"] / sp["fps"], 3),
            "p50_ms_batched": round(sb["p50_ms"], 2), "p99_ms_batched": round(sb["p99_ms"], 2),
            "p50_ms_per_frame": round(sp["p50_ms"], 2), "p99_ms_per_frame": round(sp["p99_ms"], 2),
            "gpu_busy_pct_batched": busy["batched"], "gpu_busy_pct_per_frame": busy["per_frame"],
            "mean_batch": round(sb["mean_batch"], 2), "overflows": sb["overflows"],
            "max_abs_err_batched_vs_per_frame": max_err, "frames_compared": n_cmp,
            "trace": {"requests": len(trace), "seconds": args.serve_seconds, "rate_hz": args.serve_rate,
                      "resolutions": [f"{w}x{h}" for h, w in resolutions], "points": [m[0] for m in meshes],
                      "tets": tets_of, "tile": args.tile,
                      "slots": {f"{b.W}x{b.H}": b.slots for b in srv.buckets.values()}}}


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one MI355X: this build runs the rasterizer on a single GPU "
              "(multi-GPU image-space sharding is a design note, not built); run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.delaunay_only:
        import torch
        from radiance_meshes_amd import scenes
        from radiance_meshes_amd.delaunay import delaunay_tets
        P = torch.from_numpy(scenes.gmm_points(args.points, 0).astype(np.float32)).cuda()
        delaunay_tets(P)
        torch.cuda.synchronize()
        wall = []
        for _ in range(3):
            t0 = time.perf_counter()
            t, info = delaunay_tets(P, return_info=True)
            torch.cuda.synchronize()
            wall.append(time.perf_counter() - t0)
        print(json.dumps({"delaunay_points": int(P.shape[0]), "tets": int(info["n_tets"]),
                          "ms_per_call": round(1e3 * sum(wall) / 3, 2), "info": info}), flush=True)
        return

    # the serving leg's other meshes build on host cores while the GPU runs the training leg
    pool, pending = None, []
    serve_points = [int(x) for x in args.serve_points.split(",") if x.strip()] if args.serve_seconds > 0 else []
    if serve_points and not args.profile_steps:
        import multiprocessing as mp
        others = [n for n in serve_points if n != args.points]
        if others:
            pool = mp.get_context("fork").Pool(len(others))
            pending = [pool.apply_async(_mesh, (n, not args.no_layout)) for n in others]

    pts, tets, cam = build_workload(args)
    res = None
    if not args.serve_only:
        for extra in [int(t) for t in args.sweep_tiles.split(",") if t.strip()]:
            r = run_config(args, pts, tets, cam, extra)
            log(json.dumps({"sweep_tile": extra, **{k: r[k] for k in ("value", "ms_per_step", "breakdown_ms", "config")}}))
        res = run_config(args, pts, tets, cam, args.tile)
    precision = None
    if not args.serve_only and not args.no_precision_report and not args.profile_steps:
        precision = precision_report(pts, tets, cam, args.tile)
    fps = None
    if res is not None and not args.no_forward_fps and not args.profile_steps:
        fps = forward_fps(args, pts, tets, (6.0, 30.0, 20.0), full_features(pts, tets))
    tri = None
    if not args.serve_only and not args.no_delaunay and not args.profile_steps:
        try:
            tri = delaunay_leg(args, pts, tets)
        except Exception as e:   # the training-step line is still printed
            log(f"[bench] delaunay leg failed: {type(e).__name__}: {e}")
            tri = {"error": f"{type(e).__name__}: {e}"}
    serving = None
    if serve_points and not args.profile_steps:
        import torch
        torch.cuda.empty_cache()
        built = {args.points: (pts, tets)}
        for job in pending:
            n, p_, t_, secs = job.get()
            log(f"[bench] serving mesh {n} points: {len(t_)} tets ({secs:.1f} s)")
            built[n] = (p_, t_)
        if pool is not None:
            pool.close()
        try:
            serving = serve_leg(args, [(n, *built[n]) for n in serve_points])
        except Exception as e:   # the training-step line is still printed
            log(f"[bench] serving leg failed: {type(e).__name__}: {e}")
            serving = {"error": f"{type(e).__name__}: {e}"}
    cpu = None
    if res is not None and not args.no_cpu_baseline and not args.profile_steps:
        cpu = cpu_baseline(args, pts, tets, cam, res["last"])
    if res is None:
        print(json.dumps({"metric": "batched-served vs per-frame fps", "serving": serving}), flush=True)
        return

    out = {
        "metric": "fwd+bwd Mrays/s @1M verts; batched-served vs per-frame fps + p99 ms @<=1e-4 err",
        "value": res["value"],
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": res["steps"],
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": res["config"],
        "roofline": res["roofline"],
        "hbm_frac_all_kernels": res["hbm_frac_all_kernels"],
        "breakdown_ms": res["breakdown_ms"],
        "tet_err": res.get("tet_err"),
        "image_err_vs_fp64": cpu.get("image_err_vs_fp64") if cpu else None,
        "forward_fps": fps,
        "precision": precision,
        "delaunay": tri,
        "cpu_baseline": cpu,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


def _cpu_model():
    cpu_model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return cpu_model


def cpu_baseline(args, pts, tets, cam, last):
    """The same algorithm on the host's cores, as the reference's own CPU path (tetra_quad) is the
    method the image is checked against but not a rasterizer: the CPU restatement (oracle/) of
    projection, binning and the per-tile compositor forward + backward (FP64 segments, OpenMP over
    tiles) on this step's features. Projection and binning run over the whole mesh; compositing
    over a bounded sample of tiles, extrapolated to the frame by tile count (labelled as such).
    The GPU image of the same step is compared with the FP64 image on those tiles
    (image_err_vs_fp64). The reference method itself -- midpoint quadrature testing every tet at
    every sample (jaxutil/tetra_quad.py), restated in C -- is timed on a few rays as well."""
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle import oracle as O
    H, W, tile = args.height, args.width, args.tile
    ocam = O.make_camera(cam, tile, 0.1)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    log(f"[bench] CPU baseline: oracle projection, binning and compositing on {cores} threads")
    pts = last["vertices"]
    t0 = time.perf_counter()
    ot, orr, ovs, occ = O.project(ocam, tets, pts)
    t_proj = time.perf_counter() - t0
    t0 = time.perf_counter()
    keys, vals, oranges = O.bin_tiles(ocam, ot, orr, ovs)
    t_bin = time.perf_counter() - t0
    same_bins = bool(np.array_equal(vals, last["sidx"]) and np.array_equal(oranges, last["ranges"]))
    G = ocam.gh * ocam.gw
    n_s = min(args.cpu_baseline_tiles, G)
    sample = np.sort(np.random.default_rng(7).choice(G, n_s, replace=False)).astype(np.int32)
    f64 = last["feats"].astype(np.float64)
    jit = np.full((H, W, 2), 0.5, np.float32)
    t0 = time.perf_counter()
    o_img, o_dist, o_nc, margin = O.composite_fwd(ocam, vals, oranges, tets, pts, f64, jit, tiles=sample)
    t_fwd = time.perf_counter() - t0
    gimg = np.zeros((H, W, 4))
    gimg[..., :3] = 2.0 * (o_img[..., :3] - last["target"]) / (3 * H * W)
    t0 = time.perf_counter()
    O.composite_bwd(ocam, vals, oranges, tets, pts, f64, jit, o_img, o_dist, o_nc, gimg, np.zeros((H, W, 5)),
                    tiles=sample)
    t_bwd = time.perf_counter() - t0
    frame_s = t_proj + t_bin + (t_fwd + t_bwd) * G / n_s
    pix = np.zeros((H, W), bool)
    for t in sample:
        ty, tx = divmod(int(t), ocam.gw)
        pix[ty * tile:(ty + 1) * tile, tx * tile:(tx + 1) * tile] = True
    near = pix & (margin < 1e-4)
    err = np.abs(last["img"][..., :3].astype(np.float64) - o_img[..., :3])
    image_err = {"max_abs_rgb": float(err[pix & ~near].max()), "mean_abs_rgb": float(err[pix & ~near].mean()),
                 "pixels": int(pix.sum()), "tiles": n_s,
                 "near_threshold_pixels": int(near.sum()),
                 "near_threshold_max_abs_rgb": float(err[near].max()) if near.any() else 0.0,
                 "reference": "oracle FP64 compositor on the same binning and features",
                 "binning_bit_exact_vs_oracle": same_bins}
    # the reference's CPU method (per-ray midpoint quadrature over every tet) on a few rays
    feats = last["feats"].astype(np.float64)
    pix_q = np.random.default_rng(0).choice(W * H, args.cpu_baseline_rays, replace=False)
    t0 = time.perf_counter()
    O.quadrature(ocam, jit, pix_q, 12.0, args.cpu_baseline_samples, tets, pts.astype(np.float64), feats)
    t_q = time.perf_counter() - t0
    log(f"[bench] CPU: projection {t_proj:.2f} s, binning {t_bin:.2f} s, compositing {n_s} tiles "
        f"{t_fwd:.2f} + {t_bwd:.2f} s; quadrature {t_q:.1f} s")
    return {"value": round(W * H / frame_s / 1e6, 6), "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": f"fwd+bwd of the {W}x{H} step's rasterizer: projection and binning of all "
                      f"{len(tets)} tets ({t_proj:.2f} + {t_bin:.2f} s), compositing forward + backward of "
                      f"{n_s} random {tile}x{tile} tiles of {G} ({t_fwd:.2f} + {t_bwd:.2f} s) extrapolated "
                      f"to the frame by tile count; the field is not included",
            "frame_s_extrapolated": round(frame_s, 3), "cpu": _cpu_model(),
            "reference_method": {"what": "midpoint quadrature testing every tet at every sample "
                                         "(jaxutil/tetra_quad.py) restated in C, forward only",
                                 "rays": int(len(pix_q)), "samples": args.cpu_baseline_samples,
                                 "seconds": round(t_q, 2), "mrays_per_s": round(len(pix_q) / t_q / 1e6, 9)},
            "image_err_vs_fp64": image_err}


def forward_fps(args, pts, tets, cam_pose, feats13):
    """Forward-only frame rate at 1920x1080, 2560x1440 and 3840x2160 (the resolutions of the
    reference's FPS table, tile_benchmark.py:80-88, 4x4 tiles as there) on the benchmark mesh:
    projection, binning (with its host read of the pair count) and compositing of fixed per-tet
    features, timed wall-clock per frame with the device synchronised."""
    import torch
    from radiance_meshes_amd import scenes
    from radiance_meshes_amd.rasterizer import AlphaBlendTiledRender, RenderGrid, vertex_and_tile_shader
    dev = "cuda"
    verts = torch.from_numpy(pts).to(dev)
    idx = torch.from_numpy(tets).to(dev)
    feats = torch.from_numpy(feats13).to(dev)
    out = {}
    for name, (H, W) in (("1080p", (1080, 1920)), ("2k", (1440, 2560)), ("4k", (2160, 3840))):
        cam = scenes.orbit_camera(*cam_pose, W, H, 50.0)
        tcam = {k: (v.to(dev) if hasattr(v, "to") else v) for k, v in cam.tcam(4, 0.1).items()}
        rg = RenderGrid(H, W, 4, 4)
        jit = torch.full((H, W, 2), 0.5, device=dev)

        def frame():
            with torch.no_grad():
                sidx, ranges, vs, cc, mask, rect = vertex_and_tile_shader(idx, verts, tcam, rg)
                img, dist, _ = AlphaBlendTiledRender.apply(sidx, ranges, idx, verts, feats, rg, tcam, jit)
            return img

        for _ in range(2):
            frame()
        torch.cuda.synchronize()
        t = []
        for _ in range(10):
            t0 = time.perf_counter()
            frame()
            torch.cuda.synchronize()
            t.append(time.perf_counter() - t0)
        ms = 1e3 * float(np.median(t))
        out[name] = {"width": W, "height": H, "ms_per_frame": round(ms, 3), "fps": round(1e3 / ms, 1),
                     "mrays_per_s": round(W * H / ms / 1e3, 1)}
        log(f"[bench] forward {name}: {ms:.2f} ms/frame")
    del verts, idx, feats
    torch.cuda.empty_cache()
    return {"tets": int(len(tets)), "tile": 4, "features": "13-channel field output of the training step",
            "timed": "projection + binning + compositing forward, wall clock per frame, device synchronised",
            **out}


if __name__ == "__main__":
    main()

"""Render serving on device-resident meshes: continuous batching of mixed-resolution views.

The reference renders one view per call (evaluate.py:50-63, tile_benchmark.py:59-94). Here:

* scenes (mesh + per-tet features) are registered once and stay in HBM; a loaded reference
  model (ingp.Model, e.g. from Model.load_ckpt) is registered with its field evaluated once, and
  each view of it gets its view-dependent features (activate_output at the view's camera centre,
  models/ingp_color.py:361-384) from one kernel on the batch's stream before the batched launch;
* requests are grouped by resolution bucket; each bucket owns a fixed number of slots with
  preallocated buffers, a pinned host plan, its own stream and (optionally) captured HIP graphs
  of the whole batched forward (`rm_batch_render`: projection, binning and compositing of every
  view of the batch in one launch sequence, no host synchronisation inside), one per capacity
  class (views, tets and pairs rounded up to powers of two) so that launches scale with the batch;
* a request is admitted when its bucket has a free batch and the HBM budget covers the bucket's
  reservation; requests that arrive while other batches run on the GPU are admitted as soon as a
  bucket frees up, so batches of different buckets overlap on the device (one stream each);
* `per_frame` serves the same trace one view at a time through the single-view path, the baseline.

Latency is completion time minus arrival time; p50/p99 and frames/s are reported per mode.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _capi
from ._capi import RmBatchInfo, RmCamera, RmScene, check, lib, ptr, stream_of


@dataclass
class Request:
    rid: int
    scene: int
    camera: object          # scenes.Camera
    arrival: float
    done: float = -1.0
    image: torch.Tensor = None


@dataclass
class _Field:
    """View-independent field outputs of a registered model, per tet."""
    density: torch.Tensor      # [T]
    rgb: torch.Tensor          # [T,3]
    grd: torch.Tensor          # [T,3]
    sh: torch.Tensor           # [T,K,3] float32 (values of the reference's FP16 SH)
    centroids: torch.Tensor    # [T,3]
    sh_degree: int


@dataclass
class _Scene:
    indices: torch.Tensor
    vertices: torch.Tensor
    features: torch.Tensor
    rec: RmScene
    field: _Field = None       # set for view-dependent (model) scenes


def _cam_struct(camera, tile, min_t):
    return _capi.camera_from_tcam(camera.tcam(tile, min_t))


def _pow2_at_least(x: int) -> int:
    return 1 << max(0, int(x - 1).bit_length())


class Bucket:
    """One resolution: up to `slots` views per batch. The workspace is sized for `slots` views of
    the largest registered scene. Each batch launches with the geometry of its capacity class --
    views, tets and (tile, tet) pairs each rounded up to a power of two -- so a batch of a few small
    views does not pay for the bucket's full capacity; one HIP graph is captured per class and all
    classes share the workspace (a bucket runs one batch at a time)."""

    def __init__(self, server, H, W, slots, pair_cap, use_graph, scene_pairs):
        self.server, self.H, self.W, self.slots = server, H, W, slots
        self.tile = server.tile
        dev = server.device
        big = max(range(len(server.scenes)), key=lambda i: server.scenes[i].rec.n_tets)
        cam = server.probe_camera(H, W)
        self.info = server.plan_info([(big, cam)] * slots)
        self.view_tiles = int(self.info.total_tiles) // slots
        self.view_blocks = int(self.info.total_workgroups) // slots
        self.view_bin_blocks = int(self.info.total_bin_blocks) // slots
        self.scene_pairs = list(scene_pairs)     # pairs of each scene seen from the probe pose
        self.plan_bytes = int(self.info.plan_bytes)
        self.plan_host = torch.empty(self.plan_bytes, dtype=torch.uint8, pin_memory=True)
        self.plan_dev = torch.empty(self.plan_bytes, dtype=torch.uint8, device=dev)
        npx = slots * H * W
        self.img = torch.empty((npx, 4), device=dev)
        self.dist = torch.empty((npx, 5), device=dev)
        self.nc = torch.empty((npx,), device=dev, dtype=torch.int32)
        self.counts = torch.zeros(2, device=dev, dtype=torch.int64)
        self.counts_host = torch.zeros(2, dtype=torch.int64, pin_memory=True)
        # per-slot features of view-dependent scenes (sized for the largest of them)
        t_vd = max([sc.rec.n_tets for sc in server.scenes if sc.field is not None], default=0)
        self.vfeat = torch.empty((slots, t_vd, server.n_features), device=dev) if t_vd else None
        self.stream = torch.cuda.Stream(device=dev)
        self.graphs = {}
        self.use_graph = use_graph
        self.busy = None   # (requests, outputs, event, keep_aux, pair capacity) of the batch in flight
        self.ws = None
        self._alloc(pair_cap)

    def _alloc(self, pair_cap):
        """(Re)sizes the workspace for `pair_cap` (tile, tet) pairs; captured graphs are dropped."""
        self.pair_cap = int(pair_cap)
        self.graphs = {}
        self.ws = None
        self.ws = torch.empty(lib().rm_batch_workspace_size(self.info, self.pair_cap), dtype=torch.uint8,
                              device=self.server.device)
        self.bytes = self.ws.numel() + self.img.numel() * 40 // 4 + 2 * self.plan_bytes + \
            (self.vfeat.numel() * 4 if self.vfeat is not None else 0)

    def _class_of(self, reqs, full=False):
        """(key, launch info, pair capacity) of the capacity class that holds the batch `reqs`."""
        if full:
            return ("full",), self.info, self.pair_cap
        scs = self.server.scenes
        nv = min(self.slots, _pow2_at_least(len(reqs)))
        # a plan starts every view's tets at a multiple of 256 (rm_batch_plan)
        T = min(int(self.info.total_tets), _pow2_at_least(sum(-(-scs[r.scene].rec.n_tets // 256) * 256 for r in reqs)))
        est = sum(int(self.server.pair_margin * self.scene_pairs[r.scene]) + 4096 for r in reqs)
        P = min(self.pair_cap, _pow2_at_least(est))
        info = RmBatchInfo(n_views=nv, total_tets=T, total_tiles=nv * self.view_tiles,
                           total_pixels=nv * self.H * self.W, total_workgroups=nv * self.view_blocks,
                           total_bin_blocks=nv * self.view_bin_blocks,
                           plan_bytes=self.plan_bytes, tile=self.info.tile, n_features=self.info.n_features)
        assert lib().rm_batch_workspace_size(info, P) <= self.ws.numel()
        return (nv, T, P), info, P

    def _render_call(self, info, cap):
        check(lib().rm_batch_render(ptr(self.plan_dev), info, cap, ptr(self.img), ptr(self.dist),
                                    ptr(self.nc), ptr(self.counts), ptr(self.ws), self.ws.numel(),
                                    stream_of(self.server.device)), "rm_batch_render")

    def launch(self, reqs, keep_aux=False, full=False):
        """Enqueues one batch on the bucket's stream and returns at once. The images (and with
        keep_aux the distortion images and contributor counts) are copied out on the same stream,
        so the next batch of this bucket cannot overwrite them. full=True launches with the
        bucket's whole capacity."""
        if len(reqs) > self.slots:
            raise RuntimeError("batch exceeds the bucket's capacity")
        key, cinfo, cap = self._class_of(reqs, full)
        hw = self.H * self.W
        cur = torch.cuda.current_stream(self.server.device)
        with torch.cuda.stream(self.stream):
            # The bucket's buffers come from torch's caching allocator on the current stream, which
            # hands out blocks freed there as soon as the host drops them, even while kernels queued
            # on that stream still use them (a single-view render's binning, say). Ordering this
            # stream after the current one keeps the batch from writing memory those kernels use.
            self.stream.wait_stream(cur)
            # views of model scenes read the features of their own camera, written into their
            # slot's buffer here (on this stream, after the previous batch that read it)
            views, extra = [], []
            n_sc = len(self.server.scenes)
            for i, r in enumerate(reqs):
                sc = self.server.scenes[r.scene]
                if sc.field is None:
                    views.append((r.scene, r.camera))
                    continue
                buf = self.vfeat[i, :sc.rec.n_tets]
                self.server.view_features(r.scene, r.camera, buf)
                rec = RmScene(indices_dev=sc.rec.indices_dev, vertices_dev=sc.rec.vertices_dev,
                              features_dev=buf.data_ptr(), n_tets=sc.rec.n_tets, n_verts=sc.rec.n_verts,
                              n_features=sc.rec.n_features)
                views.append((n_sc + len(extra), r.camera))
                extra.append(rec)
            info = self.server.plan_info(views, self.plan_host, extra)
            if info.total_tets > self.info.total_tets:
                raise RuntimeError("batch exceeds the bucket's capacity")
            self.plan_dev.copy_(self.plan_host, non_blocking=True)
            if self.use_graph:
                g = self.graphs.get(key)
                if g is None:
                    g = self.graphs[key] = self._capture(cinfo, cap)
                g.replay()
            else:
                self._render_call(cinfo, cap)
            self.counts_host.copy_(self.counts, non_blocking=True)
            n = len(reqs) * hw
            outs = (self.img[:n].clone(),) + ((self.dist[:n].clone(), self.nc[:n].clone()) if keep_aux else ())
            ev = torch.cuda.Event()
            ev.record(self.stream)
        self.busy = (reqs, outs, ev, keep_aux, cap)
        return ev

    def _capture(self, info, cap):
        # run once eagerly (first-launch set-up stays out of the graph), then capture
        self._render_call(info, cap)
        self.stream.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=self.stream):
            self._render_call(info, cap)
        return g

    def ready(self):
        return self.busy is not None and self.busy[2].query()

    def finish(self):
        """([(request, ([H,W,4] image, [H,W,5] distortion, [H,W] n_contrib) or (image,)), ...],
        overflowed) of the completed batch. A batch whose pairs exceeded its class's capacity is
        rendered again with the bucket's whole capacity, after the workspace grows when that is
        too small as well (or view by view through the single-view path when the HBM budget does
        not allow the growth)."""
        reqs, outs, ev, keep_aux, cap = self.busy
        ev.synchronize()
        self.busy = None
        cur = torch.cuda.current_stream(self.server.device)
        for o in outs:   # read on the caller's stream from here on: freed only after its work
            o.record_stream(cur)
        n_pairs = int(self.counts_host[1])
        if n_pairs > cap:
            if n_pairs <= self.pair_cap:
                self.launch(reqs, keep_aux, full=True)
                res, _ = self.finish()
                return res, True
            if self.server.on_overflow == "raise":
                raise RuntimeError(f"batch of {len(reqs)} views needs {n_pairs} pairs, capacity {self.pair_cap}")
            need = -(-n_pairs // len(reqs)) * self.slots * 5 // 4 + 4096
            grown = need < (1 << 31) - 1 and self.server.can_grow(self, need)
            if grown:
                self._alloc(need)
                self.launch(reqs, keep_aux, full=True)
                res, _ = self.finish()
            else:
                res = [(r, self.server.render_single(r.scene, r.camera, aux=True) if keep_aux
                        else (self.server.render_single(r.scene, r.camera),)) for r in reqs]
            return res, True
        hw = self.H * self.W
        shapes = ((self.H, self.W, 4), (self.H, self.W, 5), (self.H, self.W))
        res = [(r, tuple(o[i * hw:(i + 1) * hw].view(sh) for o, sh in zip(outs, shapes))) for i, r in enumerate(reqs)]
        return res, False


class RenderServer:
    def __init__(self, tile=4, n_features=13, min_t=0.1, hbm_budget_bytes=64 << 30, max_slots=8,
                 use_graphs=True, device="cuda", pair_margin=1.3, probe_radius=6.0, on_overflow="grow"):
        self.tile, self.n_features, self.min_t = tile, n_features, min_t
        self.pair_margin, self.probe_radius = pair_margin, probe_radius
        if on_overflow not in ("grow", "raise"):
            raise ValueError("on_overflow must be 'grow' or 'raise'")
        self.on_overflow = on_overflow
        self.budget = hbm_budget_bytes
        self.max_slots = max_slots
        self.use_graphs = use_graphs
        self.device = device
        self.scenes: list[_Scene] = []
        self.buckets: dict = {}

    # ------------------------------------------------------------------ scenes
    def add_model(self, model, chunk=1 << 20):
        """Registers a reference model (ingp.Model: vertices, indices, backbone) as a scene whose
        features depend on the view: the field (density, rgb, grd, SH) is evaluated once here, in
        chunks of `chunk` tets; each view gets activate_output at its camera centre (7 features,
        so the server must have n_features 7). Returns the scene id."""
        if self.n_features != 7:
            raise ValueError("model scenes have 7 features: create the server with n_features=7")
        with torch.no_grad():
            V = model.vertices.detach().to(self.device, torch.float32).contiguous()
            idx = model.indices.to(self.device, torch.int32).contiguous()
            T = idx.shape[0]
            outs = [model.compute_batch_features(V, idx, s, min(s + chunk, T)) for s in range(0, T, chunk)]
            density = torch.cat([o[1].reshape(-1) for o in outs])
            rgb = torch.cat([o[2].reshape(-1, 3) for o in outs])
            grd = torch.cat([o[3].reshape(-1, 3) for o in outs])
            K = model.sh_dim
            sh = torch.cat([o[4].reshape(-1, K, 3).float() for o in outs]) if K else None
            centroids = V[idx.long()].mean(dim=1)
        sid = self.add_scene(V, idx, torch.zeros((T, 7), device=self.device))
        self.scenes[sid].field = _Field(density.contiguous(), rgb.contiguous(), grd.contiguous(),
                                        sh.contiguous() if sh is not None else None, centroids.contiguous(),
                                        int(model.current_sh_deg))
        return sid

    def view_features(self, sid, camera, out=None):
        """[T,7] features of model scene `sid` seen from `camera` (rm_activate_fwd at its centre) on
        the current stream, into `out` when given."""
        sc = self.scenes[sid]
        f = sc.field
        T = sc.rec.n_tets
        out = torch.empty((T, 7), device=self.device) if out is None else out
        center = torch.as_tensor(np.asarray(camera.cam_pos, np.float32).reshape(3)).to(self.device, non_blocking=True)
        K = 0 if f.sh is None else f.sh.shape[1]
        check(lib().rm_activate_fwd(ptr(center), ptr(f.density), ptr(f.rgb), ptr(f.grd), ptr(f.sh) if K else None,
                                    K, f.sh_degree, ptr(sc.indices), ptr(f.centroids), ptr(sc.vertices), T, ptr(out),
                                    stream_of(self.device)), "rm_activate_fwd")
        return out

    def add_scene(self, vertices, indices, features):
        """Registers a device-resident scene and returns its id. Buckets sized for the scenes
        registered before are dropped (their workspaces and per-scene pair estimates no longer
        fit) and are created again on their next use; a batch must not be in flight."""
        if any(b.busy is not None for b in self.buckets.values()):
            raise RuntimeError("add_scene while a batch is in flight: finish the batches first")
        if self.buckets:
            torch.cuda.synchronize(self.device)
            self.buckets = {}
        v = vertices.to(self.device, torch.float32).contiguous()
        i = indices.to(self.device, torch.int32).contiguous()
        f = features.to(self.device, torch.float32).contiguous()
        if f.shape[1] != self.n_features:
            raise ValueError(f"features must have {self.n_features} channels")
        rec = RmScene(indices_dev=i.data_ptr(), vertices_dev=v.data_ptr(), features_dev=f.data_ptr(),
                      n_tets=i.shape[0], n_verts=v.shape[0], n_features=f.shape[1])
        self.scenes.append(_Scene(i, v, f, rec))
        return len(self.scenes) - 1

    def _pairs_of(self, sid, camera):
        from .rasterizer import RenderGrid, vertex_and_tile_shader
        sc = self.scenes[sid]
        rg = RenderGrid(camera.height, camera.width, self.tile, self.tile)
        with torch.no_grad():
            sidx, *_ = vertex_and_tile_shader(sc.indices, sc.vertices, self._tcam(camera), rg)
        return sidx.numel()

    def _tcam(self, camera):
        return camera.tcam(self.tile, self.min_t)

    # ------------------------------------------------------------------ single view (baseline)
    def render_single(self, sid, camera, aux=False):
        """One view through the single-view path (host sync for the pair count, as the reference):
        the [H,W,4] image, or with aux the tuple (image, [H,W,5] distortion, [H,W] n_contrib) that
        batched views return."""
        from .rasterizer import RenderGrid, _cam, tile_order, vertex_and_tile_shader
        sc = self.scenes[sid]
        H, W = camera.height, camera.width
        rg = RenderGrid(H, W, self.tile, self.tile)
        tcam = self._tcam(camera)
        with torch.no_grad():
            feats = sc.features if sc.field is None else self.view_features(sid, camera)
            sidx, ranges, vs, cc, mask, rect = vertex_and_tile_shader(sc.indices, sc.vertices, tcam, rg)
            jit = torch.full((H, W, 2), 0.5, device=self.device)
            img = torch.empty((H, W, 4), device=self.device)
            dist = torch.empty((H, W, 5), device=self.device)
            nc = torch.empty((H, W), device=self.device, dtype=torch.int32)
            # 4x4 tiles: the longest lists first, as the batched path does (same images in any order)
            order = tile_order(ranges) if self.tile == 4 else None
            check(lib().rm_composite_fwd_flags(_cam(tcam, rg), ptr(sidx), sidx.numel(), ptr(ranges), ptr(sc.indices),
                                               ptr(sc.vertices), ptr(feats), feats.shape[1], ptr(jit), ptr(img),
                                               ptr(dist), ptr(nc), None, ptr(order), stream_of(self.device)),
                  "rm_composite_fwd_flags")
        return (img, dist, nc) if aux else img

    # ------------------------------------------------------------------ buckets
    def probe_camera(self, H, W):
        from . import scenes
        return scenes.orbit_camera(self.probe_radius, 0.0, 0.0, W, H, 50.0)

    def plan_info(self, views, plan_host=None, extra=()):
        """rm_batch_plan over [(scene, camera), ...]; packs into `plan_host` (pinned) when given.
        `extra` scene records follow the registered ones (ids len(scenes) + i): views of model
        scenes, each with its own features."""
        n = len(views)
        cams = (RmCamera * n)(*[_cam_struct(c, self.tile, self.min_t) for _, c in views])
        sid = (C.c_int32 * n)(*[s for s, _ in views])
        recs = [sc.rec for sc in self.scenes] + list(extra)
        scs = (RmScene * len(recs))(*recs)
        info = RmBatchInfo()
        buf = None if plan_host is None else C.c_void_p(plan_host.data_ptr())
        nbytes = 0 if plan_host is None else plan_host.numel()
        check(lib().rm_batch_plan(scs, len(recs), cams, sid, n, buf, nbytes, C.byref(info)),
              "rm_batch_plan")
        return info

    def can_grow(self, bucket, pair_cap):
        """Whether `bucket` may grow its workspace to `pair_cap` pairs within the HBM budget."""
        extra = lib().rm_batch_workspace_size(bucket.info, pair_cap) - bucket.ws.numel()
        return sum(b.bytes for b in self.buckets.values()) + extra <= self.budget

    def bucket(self, H, W):
        """The bucket of resolution HxW, created on first use. Its pair capacity per view is
        pair_margin times the largest pair count of any registered scene seen from a probe pose at this resolution (each
        visible tet costs at least one pair, so small images of large meshes need far more pairs
        per pixel than large ones); slots (up to max_slots) are what the HBM budget left over by
        the other buckets allows."""
        key = (H, W)
        if key not in self.buckets:
            probe = self.probe_camera(H, W)
            scene_pairs = [self._pairs_of(s, probe) for s in range(len(self.scenes))]
            per_view_pairs = int(self.pair_margin * max(scene_pairs)) + 4096
            big = max(range(len(self.scenes)), key=lambda i: self.scenes[i].rec.n_tets)
            remaining = self.budget - sum(b.bytes for b in self.buckets.values())
            slots = self.max_slots
            while slots > 1:
                info = self.plan_info([(big, probe)] * slots)
                need = lib().rm_batch_workspace_size(info, per_view_pairs * slots) + slots * H * W * 40
                if need <= remaining and per_view_pairs * slots < (1 << 31) - 1:
                    break
                slots //= 2
            self.buckets[key] = Bucket(self, H, W, slots, per_view_pairs * slots, self.use_graphs, scene_pairs)
        return self.buckets[key]

    def render_views(self, views, keep_aux=False):
        """Synchronous batched render of [(scene, camera), ...] of one resolution (tests, warm-up):
        a list of (image,) or, with keep_aux, (image, distortion, n_contrib) per view."""
        cam0 = views[0][1]
        b = self.bucket(cam0.height, cam0.width)
        out = []
        for k in range(0, len(views), b.slots):
            chunk = [Request(i, s, c, 0.0) for i, (s, c) in enumerate(views[k:k + b.slots])]
            b.launch(chunk, keep_aux)
            res, _ = b.finish()
            out += [o for _, o in res]
        return out

    # ------------------------------------------------------------------ trace replay
    def serve(self, trace, mode="batched", max_batch=None, keep=None, poll_s=2e-4, progress=None,
              progress_every_s=5.0):
        """Replays `trace` (Requests with arrival times in seconds from now) in real time.
        mode "batched": continuous batching (at most max_batch views per batch, default the
        bucket's slots); "per_frame": one view at a time through the single-view path.
        Results are stored in Request.image for the requests `keep(request)` accepts (all when
        keep is None). progress(done, total, seconds) is called about every progress_every_s.
        Returns stats: frames/s over the makespan, latency p50/p99 (ms), batches, overflows."""
        keep = keep or (lambda r: True)
        last_report = [0.0]

        def report(n_done, now):
            if progress is not None and now - last_report[0] >= progress_every_s:
                last_report[0] = now
                progress(n_done, len(trace), now)
        torch.cuda.synchronize()
        reqs = sorted(trace, key=lambda r: r.arrival)
        pending = deque(reqs)
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0  # noqa: E731
        done = 0
        n_batches = 0
        overflows = 0
        sizes = []
        if mode == "per_frame":
            while pending:
                r = pending.popleft()
                wait = r.arrival - clock()
                if wait > 0:
                    time.sleep(wait)
                img = self.render_single(r.scene, r.camera)
                torch.cuda.current_stream().synchronize()
                r.done = clock()
                if keep(r):
                    r.image = img
                n_batches += 1
                sizes.append(1)
                report(n_batches, r.done)
        else:
            queue: dict = {}
            in_use = 0
            while done < len(reqs):
                now = clock()
                report(done, now)
                while pending and pending[0].arrival <= now:
                    r = pending.popleft()
                    queue.setdefault((r.camera.height, r.camera.width), deque()).append(r)
                progressed = False
                for b in self.buckets.values():
                    if b.ready():
                        held = b.bytes
                        res, ov = b.finish()
                        t = clock()
                        for r, o in res:
                            r.done = t
                            if keep(r):
                                r.image = o[0]
                        done += len(res)
                        in_use -= held
                        overflows += int(ov)
                        progressed = True
                # admission: oldest waiting request first, if its bucket is free and the budget allows
                order = sorted((q[0].arrival, key) for key, q in queue.items() if q)
                for _, key in order:
                    b = self.bucket(*key)
                    if b.busy is not None or in_use + b.bytes > self.budget:
                        continue
                    q = queue[key]
                    batch = [q.popleft() for _ in range(min(b.slots, len(q), max_batch or b.slots))]
                    b.launch(batch)
                    in_use += b.bytes
                    n_batches += 1
                    sizes.append(len(batch))
                    progressed = True
                if not progressed:
                    nxt = pending[0].arrival - clock() if pending else poll_s
                    time.sleep(max(0.0, min(poll_s, nxt)))
        torch.cuda.synchronize()
        lat = np.array([(r.done - r.arrival) * 1e3 for r in reqs])
        makespan = max(r.done for r in reqs) - min(r.arrival for r in reqs)
        return {"mode": mode, "frames": len(reqs), "fps": len(reqs) / makespan,
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "mean_ms": float(lat.mean()), "batches": n_batches, "mean_batch": float(np.mean(sizes)),
                "overflows": overflows, "makespan_s": makespan}


def poisson_trace(n_scenes, resolutions, rate_hz, duration_s, seed=0, radius=6.0):
    """Requests with exponential inter-arrival times (rate_hz), a uniform scene, a random orbit pose
    (azimuth U[0,360), elevation U[-20,40) degrees, 50 degree vertical FOV) and a uniform resolution."""
    from . import scenes
    rng = np.random.default_rng(seed)
    t, out, rid = 0.0, [], 0
    while True:
        t += rng.exponential(1.0 / rate_hz)
        if t > duration_s:
            break
        H, W = resolutions[rng.integers(len(resolutions))]
        cam = scenes.orbit_camera(radius, rng.uniform(0, 360), rng.uniform(-20, 40), W, H, 50.0)
        out.append(Request(rid, int(rng.integers(n_scenes)), cam, t))
        rid += 1
    return out

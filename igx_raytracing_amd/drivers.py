"""Workload drivers: the multi-scene batch and the serving loop, measured
against plain launches on the same GPU.  Used by bench.py; each returns a
dict of measurements.  Sizes follow the workload definitions:

* multi-scene batch: scenes from the cluster generator with triangle counts
  log-uniform over [10k, 2M], one ray batch per scene log-uniform over
  [1k, 1M] rays, half closest hit and half occlusion, BVHs built on the GPU
  by LBVH, launch shape and node width chosen by the autotuner;
* serving: resident scenes from the cluster generator, Poisson arrivals,
  request sizes log-uniform over [256, 4M] rays, half closest hit and half
  occlusion, offered load as a fraction of the measured peak, p50/p99
  latency beside Mrays/s, served (continuous batching) and plain (one
  launch per request).
"""
import threading
import time

import numpy as np

from . import workload as wl
from .rt import Batch, Scene, rays_to_device
from .serve import DONE, Server


def log_uniform_int(rng, lo, hi, size=None):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size)).astype(np.int64)


def cluster_scene(seed, ntris):
    """About ntris triangles from the cluster generator (4,096 per cluster)."""
    tpc = 4096 if ntris >= 4096 else int(ntris)
    return wl.clusters(seed, max(1, int(round(ntris / tpc))), tpc)


def _events_ms(fn, iters):
    import torch
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(iters):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / iters


# ---- multi-scene batch ------------------------------------------------------

def batch_workload(ctx, nscenes=64, seed=4, iters=5, widths=(4, 8), log=print):
    import torch
    rng = np.random.default_rng(seed)
    tri_counts = log_uniform_int(rng, 10_000, 2_000_000, nscenes)
    ray_counts = log_uniform_int(rng, 1_000, 1_000_000, nscenes)
    kinds = ["closest" if k % 2 == 0 else "occluded" for k in range(nscenes)]
    log(f"batch: {nscenes} scenes, {int(tri_counts.sum())} triangles, {int(ray_counts.sum())} rays")
    tri_sets = [cluster_scene(wl.seed(1000 + k), int(tri_counts[k])) for k in range(nscenes)]
    d_tris = [torch.from_numpy(t).cuda() for t in tri_sets]
    d_rays, outs = [], []
    for k in range(nscenes):
        lo, hi = wl.bounds(tri_sets[k])
        r = wl.rays_uniform(wl.seed(2000 + k), lo, hi, int(ray_counts[k]),
                            tmax=None if kinds[k] == "closest" else (0.0, 20.0))
        d_rays.append(rays_to_device(r))
        n = r.shape[0]
        outs.append(torch.empty((n, 4), dtype=torch.int32, device="cuda") if kinds[k] == "closest"
                    else torch.empty(n, dtype=torch.uint8, device="cuda"))
    total_rays = int(ray_counts.sum())
    result = {"scenes": nscenes, "triangles": int(tri_counts.sum()), "rays": total_rays, "per_width": {}}
    best = None
    for w in widths:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        scenes = [Scene(ctx, device_tris=d, width=w, keep_source=False) for d in d_tris]
        torch.cuda.synchronize()
        build_ms = (time.perf_counter() - t0) * 1e3
        queries = [(k, kinds[k], d_rays[k], outs[k]) for k in range(nscenes)]
        batch = Batch(ctx, scenes)
        opts, _ = batch.autotune(queries, iters=2)
        table = batch._table(queries)
        ms_batch = _events_ms(lambda: batch.trace(queries, opts=opts, table=table), iters)
        ref = [o.clone() for o in outs]

        def plain():
            for k in range(nscenes):
                if kinds[k] == "closest":
                    scenes[k].trace_closest(d_rays[k], outs[k])
                else:
                    scenes[k].trace_occluded(d_rays[k], outs[k])
        ms_plain = _events_ms(plain, iters)
        torch.cuda.synchronize()
        same = all(torch.equal(a, b) for a, b in zip(ref, outs))
        # Algorithmic bytes of the batch: every ray's record and output plus
        # the node and triangle bytes of its walk, counted by the counting
        # variant on each scene (the same walk the batch kernel makes).
        alg_bytes = 0.0
        for k in range(nscenes):
            cnt = torch.empty((d_rays[k].shape[0], 2), dtype=torch.int32, device="cuda")
            if kinds[k] == "closest":
                scenes[k].trace_closest(d_rays[k], outs[k], counts=cnt)
            else:
                scenes[k].trace_occluded(d_rays[k], outs[k], counts=cnt)
            c = cnt.cpu().numpy().view(np.uint32).astype(np.float64)
            alg_bytes += c.shape[0] * (32 + (16 if kinds[k] == "closest" else 1)) + \
                scenes[k].info["node_bytes"] * c[:, 0].sum() + 36 * c[:, 1].sum()
        achieved = alg_bytes / (ms_batch * 1e-3) / 1e9
        entry = {"lbvh_build_ms_all_scenes": round(build_ms, 1), "batched_ms": round(ms_batch, 3),
                 "plain_ms": round(ms_plain, 3), "batched_mrays_s": round(total_rays / ms_batch / 1e3, 1),
                 "plain_mrays_s": round(total_rays / ms_plain / 1e3, 1), "speedup": round(ms_plain / ms_batch, 2),
                 "bitexact_batched_vs_plain": bool(same),
                 "tuned": {"block": opts.block_size, "lds_stack": opts.lds_stack, "chunk": opts.chunk,
                           "refill": opts.refill},
                 "nodes_mb": round(sum(s.info["nodes_total_bytes"] for s in scenes) / 1e6, 1),
                 "roofline": {"kernel": "k_trace_batch", "bound": "hbm", "achieved": round(achieved, 1),
                              "peak": 8000.0, "unit": "GB/s", "frac": round(achieved / 8000.0, 4),
                              "algorithmic_bytes_per_launch": round(alg_bytes)}}
        log(f"batch BVH{w}: {entry}")
        result["per_width"][w] = entry
        if best is None or ms_batch < result["per_width"][best]["batched_ms"]:
            best = w
        batch.close()
        for s in scenes:
            s.close()
    result["tuned_width"] = best
    result.update({k: result["per_width"][best][k] for k in
                   ("batched_mrays_s", "plain_mrays_s", "speedup", "bitexact_batched_vs_plain", "roofline")})
    return result


# ---- serving -------------------------------------------------------------

class _Pools:
    """Pinned host buffers every request slices from (rays) and writes to."""

    def __init__(self, max_rays, seed):
        import torch
        box = (np.full(3, -55.0, np.float32), np.full(3, 55.0, np.float32))
        closest = wl.rays_uniform(seed, box[0], box[1], max_rays)
        occl = wl.rays_uniform(seed, box[0], box[1], max_rays, tmax=(0.0, 20.0))
        self.rays = {}
        for kind, r in (("closest", closest), ("occluded", occl)):
            t = torch.empty((max_rays, 8), dtype=torch.float32).pin_memory()
            t.copy_(torch.from_numpy(r.view(np.float32).reshape(-1, 8)))
            self.rays[kind] = t
        # Outputs of concurrent requests share these buffers: the benchmark
        # measures time; per-request results are checked by the tests.
        self.out = {"closest": torch.empty((max_rays, 4), dtype=torch.int32).pin_memory(),
                    "occluded": torch.empty(max_rays, dtype=torch.uint8).pin_memory()}
        self.max_rays = max_rays
        self.device = False

    def slice(self, kind, off, n):
        return self.rays[kind][off:off + n], self.out[kind][:n]


class _DevPools(_Pools):
    """The same rays, resident in HBM: requests whose rays a GPU stage made,
    with no host copies (RT_REQ_DEVICE)."""

    def __init__(self, host):
        self.rays = {k: v.cuda() for k, v in host.rays.items()}
        self.out = {k: v.cuda() for k, v in host.out.items()}
        self.max_rays = host.max_rays
        self.device = True


def _draw_requests(rng, count, nscenes, max_rays):
    sizes = log_uniform_int(rng, 256, max_rays, count)
    kinds = rng.integers(0, 2, count)
    scenes = rng.integers(0, nscenes, count)
    offs = rng.integers(0, max_rays, count)
    offs = np.minimum(offs, max_rays - sizes)
    return sizes, kinds, scenes, offs


def _run_arrivals(srv, pools, ids, rate, duration_s, rng, max_rays, uploads=None):
    """Poisson arrivals at `rate` requests/s for duration_s; returns (latencies ms, rays, wall s)."""
    n_req = max(1, int(rate * duration_s))
    gaps = rng.exponential(1.0 / rate, n_req)
    sizes, kinds, scs, offs = _draw_requests(rng, n_req, len(ids), max_rays)
    tickets = []
    t0 = time.perf_counter()
    at = t0
    upload_threads = []
    for k in range(n_req):
        at += gaps[k]
        while time.perf_counter() < at:
            pass
        kind = "closest" if kinds[k] == 0 else "occluded"
        src, out = pools.slice(kind, int(offs[k]), int(sizes[k]))
        tickets.append(srv.submit(ids[int(scs[k])], kind, int(sizes[k]), src, out, device=pools.device))
        if uploads and k in uploads:
            th = threading.Thread(target=uploads[k], args=(srv,))
            th.start()
            upload_threads.append(th)
    lat = []
    for t in tickets:
        st, status, ms = srv.wait(t, 120_000)
        if st == DONE and status == 0:
            lat.append(ms)
        srv.forget(t)
    wall = time.perf_counter() - t0
    for th in upload_threads:
        th.join()
    return np.array(lat), int(sizes.sum()), wall


def _peak(srv, pools, ids, rng, max_rays, count=300):
    sizes, kinds, scs, offs = _draw_requests(rng, count, len(ids), max_rays)
    t0 = time.perf_counter()
    tickets = []
    for k in range(count):
        kind = "closest" if kinds[k] == 0 else "occluded"
        src, out = pools.slice(kind, int(offs[k]), int(sizes[k]))
        tickets.append(srv.submit(ids[int(scs[k])], kind, int(sizes[k]), src, out, device=pools.device))
    for t in tickets:
        srv.wait(t, 120_000)
        srv.forget(t)
    wall = time.perf_counter() - t0
    return int(sizes.sum()) / wall, count / wall


def _serve_modes(ctx, scenes, pools, rng, max_rays, mean_size, duration_s, loads, width, log, make_upload=None,
                 upload_tris=(), own_peak=True):
    """Peak and offered-load runs of the served and plain servers over one
    request pool (host or device buffers)."""
    modes = {}
    for mode in ("served", "plain"):
        srv = Server(ctx, mem_budget=8 << 30, streams=4, plain=(mode == "plain"), node_width=width)
        ids = [srv.add_scene(s) for s in scenes]
        _peak(srv, pools, ids, rng, max_rays, count=30)  # warm-up (allocations)
        peak_rays, peak_req = _peak(srv, pools, ids, rng, max_rays)
        entry = {"peak_mrays_s": round(peak_rays / 1e6, 1), "peak_requests_s": round(peak_req, 1), "loads": {}}
        # Offered load is a fraction of the SERVED peak for both modes, so
        # the plain server sees the same arrival process.
        base = modes["served"]["peak_mrays_s"] * 1e6 if mode == "plain" else peak_rays
        for load in loads:
            rate = load * base / mean_size
            uploads = None
            if make_upload and mode == "served" and load == 0.8:
                n = max(1, int(rate * duration_s))
                uploads = {n // 3: make_upload(upload_tris[0]), (2 * n) // 3: make_upload(upload_tris[1])}
            lat, rays, wall = _run_arrivals(srv, pools, ids, rate, duration_s, rng, max_rays, uploads)
            e = {"offered_requests_s": round(float(rate), 1), "requests": int(lat.size),
                 "p50_ms": round(float(np.percentile(lat, 50)), 3) if lat.size else None,
                 "p99_ms": round(float(np.percentile(lat, 99)), 3) if lat.size else None,
                 "mrays_s": round(rays / wall / 1e6, 1)}
            entry["loads"][str(load)] = e
            log(f"serve {mode} load {load}: {e}")
        if own_peak and mode == "plain":
            # The same loads as fractions of the plain server's own peak.
            entry["loads_of_own_peak"] = {}
            for load in loads:
                rate = load * peak_rays / mean_size
                lat, rays, wall = _run_arrivals(srv, pools, ids, rate, duration_s, rng, max_rays)
                e = {"offered_requests_s": round(float(rate), 1), "requests": int(lat.size),
                     "p50_ms": round(float(np.percentile(lat, 50)), 3) if lat.size else None,
                     "p99_ms": round(float(np.percentile(lat, 99)), 3) if lat.size else None,
                     "mrays_s": round(rays / wall / 1e6, 1)}
                entry["loads_of_own_peak"][str(load)] = e
                log(f"serve plain load {load} of its own peak: {e}")
        st = srv.stats()
        entry["mean_requests_per_tick"] = round(st["mean_requests_per_tick"], 2)
        entry["mem_peak_bytes"] = int(st["mem_peak"])
        modes[mode] = entry
        srv.close()
    return modes


def serve_workload(ctx, total_tris=32_000_000, nscenes=32, max_rays=4 << 20, duration_s=2.0,
                   loads=(0.5, 0.8, 1.0), seed=5, width=4, log=print):
    import torch
    rng = np.random.default_rng(seed)
    counts = log_uniform_int(rng, 100_000, 2_000_000, nscenes).astype(np.float64)
    counts = np.maximum(4096, np.round(counts * total_tris / counts.sum())).astype(np.int64)
    log(f"serve: {nscenes} resident scenes, {int(counts.sum())} triangles")
    t0 = time.perf_counter()
    scenes = []
    for k in range(nscenes):
        d = torch.from_numpy(cluster_scene(wl.seed(3000 + k), int(counts[k]))).cuda()
        scenes.append(Scene(ctx, device_tris=d, width=width, keep_source=False))
        del d
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    resident = sum(s.info["nodes_total_bytes"] + s.info["tris_total_bytes"] for s in scenes)
    pools = _Pools(max_rays, wl.seed(3999))
    mean_size = (max_rays - 256) / np.log(max_rays / 256)
    result = {"scenes": nscenes, "triangles": int(sum(s.info["num_tris"] for s in scenes)),
              "resident_bytes": int(resident), "pool_build_s": round(build_s, 2),
              "request_rays": f"log-uniform [256, {max_rays}]", "mean_request_rays": int(mean_size),
              "duration_s": duration_s, "modes": {}}
    upload_tris = [cluster_scene(wl.seed(3900 + k), n) for k, n in enumerate((300_000, 700_000))]
    upload_ms = []

    def make_upload(tris):
        def go(srv):
            a = time.perf_counter()
            srv.upload_scene(tris)
            upload_ms.append((time.perf_counter() - a) * 1e3)
        return go

    result["modes"] = _serve_modes(ctx, scenes, pools, rng, max_rays, mean_size, duration_s, loads, width, log,
                                   make_upload=make_upload, upload_tris=upload_tris)
    # The same arrival process with the rays already in HBM (RT_REQ_DEVICE):
    # no PCIe copies, so the served loop's gain is the batched launch's.
    dev = _DevPools(pools)
    result["device_requests"] = _serve_modes(ctx, scenes, dev, rng, max_rays, mean_size, duration_s,
                                             tuple(x for x in loads if x <= 0.8), width,
                                             lambda m: log("device requests: " + m), own_peak=False)
    del dev
    result["uploads_during_serving_ms"] = [round(x, 1) for x in upload_ms]
    for s in scenes:
        s.close()
    return result


# ---- wavefront path tracing ------------------------------------------------

def pathtrace_workload(ctx, width=1920, height=1080, spp=4, max_depth=3, frames=5, node_width=4, presplit=16,
                       log=print):
    """Diffuse path tracing of the displaced icosphere + clusters (2,097,152
    triangles), one hipGraph replay per frame, beside eager launches.  The
    scene's BVH is built with reference splitting (`presplit`): its
    icosphere is a shell of large overlapping slivers.  The same frames on
    the unsplit BVH are timed beside it, and the two images must be equal."""
    import torch
    from .rt import PathTracer, screen_camera
    tris = wl.c3_scene()
    cam = screen_camera((0.0, 4.0, -32.0), (0.0, 0.0, 0.0), 60.0, aspect=width / height)
    out = {"triangles": int(tris.shape[0]), "resolution": [width, height], "spp": spp, "bounces": max_depth,
           "presplit": presplit}
    images = {}
    for label, ps in (("split", presplit), ("no_split", 1)):
        t0 = time.perf_counter()
        sc = Scene(ctx, tris, width=node_width, keep_source=False, presplit=ps)
        build_s = time.perf_counter() - t0
        pt = PathTracer(sc, cam, width, height, spp, max_depth=max_depth)
        img = torch.empty((height, width), dtype=torch.float32, device="cuda")
        pt.render(0, image=img, use_graph=True)  # capture
        pt.render(0, image=img, use_graph=False)
        counts = pt.counts()
        rays_per_frame = int(counts.sum())
        res = {"sah_build_s": round(build_s, 2), "records": int(sc.info["tris_total_bytes"] // 48)}
        modes = (("graph", True), ("eager", False)) if label == "split" else (("graph", True),)
        for mode, graph in modes:
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for f in range(frames):
                pt.render(f + 1, image=img, use_graph=graph)
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t1) * 1e3 / frames
            res[f"{mode}_ms_per_frame"] = round(ms, 3)
            res[f"{mode}_mrays_s"] = round(rays_per_frame / ms / 1e3, 1)
        images[label] = img.cpu().numpy()
        if label == "split":
            out.update(res)
            out["rays_per_depth"] = [int(c) for c in counts]
            out["rays_per_frame"] = rays_per_frame
            out["mean_radiance"] = round(float(img.mean().item()), 5)
            out["roofline"] = _bounce_roofline(sc, pt, cam, width, height, spp, max_depth, frames, img)
        else:
            out["no_split"] = res
        pt.close()
        sc.close()
    out["same_image_as_no_split"] = bool(np.array_equal(images["split"].view(np.uint32),
                                                        images["no_split"].view(np.uint32)))
    log(f"pathtrace: {out}")
    return out


def _bounce_roofline(sc, pt, cam, width, height, spp, max_depth, frames, img):
    """Algorithmic bytes of one frame's bounce launches over their time.

    Each depth's rays are read back from the ray queues (a path tracer keeps
    its last two depths, so one with max_depth - 2 supplies the earlier ones
    of the same frame seed) and traced by the counting variant of the plain
    closest-hit walk, which visits the nodes the bounce walk visits.  Per ray
    of depth d: its record (32 B), queue entry (4 B), path state read and
    written (64 B), 64 B per node and 36 B per triangle test of its walk, the
    hit triangle's two edge words (32 B) when it hits, and, for the rays of
    depth d + 1, the new ray and queue entry (36 B).  The frame time comes
    from HIP events around graph replays on the caller's stream, and it also
    holds the camera and resolve kernels (about 1% of a frame)."""
    import torch
    from .rt import PathTracer
    seed = frames + 100
    depth_rays = {}
    tracers = [pt]
    md = max_depth
    while md >= 0:
        p = pt if md == max_depth else PathTracer(sc, cam, width, height, spp, max_depth=md)
        if p is not pt:
            tracers.append(p)
        p.render(seed, image=img, use_graph=False)
        for d in (md, md - 1):
            if d >= 0 and d not in depth_rays:
                depth_rays[d] = p.queue_rays(d)
        md -= 2
    per_depth = []
    total = 0.0
    for d in range(max_depth + 1):
        rays = depth_rays[d]
        n = rays.shape[0]
        hits = torch.empty((n, 4), dtype=torch.int32, device="cuda")
        cnt = torch.empty((n, 2), dtype=torch.int32, device="cuda")
        if n:
            sc.trace_closest(rays, hits, counts=cnt)
        torch.cuda.synchronize()
        c = cnt.cpu().numpy().view(np.uint32).astype(np.float64)
        nhit = int((hits[:, 0].cpu().numpy().view(np.uint32) != 0xFFFFFFFF).sum()) if n else 0
        nnext = depth_rays[d + 1].shape[0] if d + 1 in depth_rays else 0
        b = n * (32 + 4 + 64) + sc.info["node_bytes"] * c[:, 0].sum() + 36 * c[:, 1].sum() + 32 * nhit + 36 * nnext
        total += b
        per_depth.append({"rays": n, "nodes_per_ray": round(float(c[:, 0].mean()), 3) if n else 0.0,
                          "tris_per_ray": round(float(c[:, 1].mean()), 3) if n else 0.0, "hits": nhit,
                          "algorithmic_bytes": round(b)})
    for p in tracers[1:]:
        p.close()
    ms = _events_ms(lambda: pt.render(seed, image=img, use_graph=True), frames)
    achieved = total / (ms * 1e-3) / 1e9
    return {"kernel": "k_pt_bounce", "bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
            "frac": round(achieved / 8000.0, 4), "algorithmic_bytes_per_frame": round(total),
            "frame_ms_events": round(ms, 3), "per_depth": per_depth}


# ---- the reference's own case ----------------------------------------------

def c1_workload(ctx, iters=20, log=print):
    """The reference's CPU-runnable case: 512x512 pinhole primary rays (eye
    (0,5,-25) looking at the origin, 60 deg) on the 4,096-triangle scene of
    cubes, a ground grid and loose triangles; binary SAH with at most 4
    triangles per leaf (BVH2), and the same build collapsed to BVH4.
    Returns (figures, the BVH2 hits as an (n, 4) uint32 array)."""
    import torch
    tris = wl.c1_scene()
    rays = wl.c1_rays()
    d = rays_to_device(rays)
    n = rays.shape[0]
    hits = torch.empty((n, 4), dtype=torch.int32, device="cuda")
    out = {"rays": n, "triangles": int(tris.shape[0]), "bvh": "binary SAH, at most 4 triangles per leaf"}
    first = None
    for w in (2, 4):
        sc = Scene(ctx, tris, width=w, max_leaf=4, keep_source=False)
        ms = _events_ms(lambda: sc.trace_closest(d, hits), iters)
        out[f"bvh{w}_mrays_s"] = round(n / ms / 1e3, 1)
        out[f"bvh{w}_ms"] = round(ms, 4)
        torch.cuda.synchronize()
        h = hits.cpu().numpy().view(np.uint32).copy()
        if first is None:
            first = h
        else:
            out["bvh4_equals_bvh2"] = bool(np.array_equal(h, first))
        sc.close()
    out["hit_fraction"] = round(float(np.mean(first[:, 0] != 0xFFFFFFFF)), 4)
    log(f"c1: {out}")
    return out, first


# ---- analytic primitives ------------------------------------------------------

def analytic_workload(ctx, width=1920, height=1080, iters=10, node_width=4, log=print):
    """Scenes with analytic objects after the triangles (trace.glsl:22-46):

    * the reference's demo scene (3 triangles, 7 spheres, 2 cubes, 1 plane):
      one primary ray per pixel at width x height through the query API and
      through the raygen caller, and one shadow ray per hit toward a point
      light;
    * 4,096 triangles plus 64 spheres, 32 cubes and 2 planes: 1,048,576
      random rays, closest hit, next to the same rays on the triangles alone
      (the difference is the analytic pass)."""
    import torch
    from .rt import screen_camera
    out = {}
    tris, ana = wl.niels_scene()
    sc = Scene(ctx, tris, width=node_width)
    sc.set_analytic(**ana)
    rays = wl.pinhole(width, height, eye=(0.0, 5.0, -9.0), target=(0.0, 4.0, 0.0), vfov_deg=60.0)
    n = rays.shape[0]
    d = rays_to_device(rays)
    hits = torch.empty((n, 4), dtype=torch.int32, device="cuda")
    sc.trace_closest(d, hits)
    ms = _events_ms(lambda: sc.trace_closest(d, hits), iters)
    h = hits.cpu().numpy().view(np.uint32)
    hit = h[:, 0] != 0xFFFFFFFF
    t = h[:, 1].view(np.float32)
    p = (rays["org"][hit] + t[hit, None] * rays["dir"][hit]).astype(np.float32)
    lv = (np.float32([0.0, 12.0, 0.0])[None, :] - p).astype(np.float32)
    dist = np.sqrt((lv.astype(np.float64) ** 2).sum(1)).astype(np.float32)
    sh = np.zeros(p.shape[0], rays.dtype)
    sh["org"] = p
    sh["dir"] = lv / dist[:, None]
    sh["tmax"] = dist
    sh["prev_id"] = h[hit, 0]
    ds = rays_to_device(sh)
    flags = torch.empty(sh.shape[0], dtype=torch.uint8, device="cuda")
    ms_sh = _events_ms(lambda: sc.trace_occluded(ds, flags), iters)
    cam = screen_camera((0.0, 5.0, -9.0), (0.0, 4.0, 0.0), 60.0, aspect=width / height)
    dob = torch.empty((height, width, 4), dtype=torch.float32, device="cuda")
    uvn = torch.empty((height, width, 4), dtype=torch.float32, device="cuda")
    ms_rg = _events_ms(lambda: sc.raygen_closest(cam, width, height, dir_object=dob, uv_normal=uvn), iters)
    out["niels_scene"] = {
        "rays": n, "objects": {"triangles": 3, "spheres": 7, "cubes": 2, "planes": 1},
        "closest_mrays_s": round(n / (ms * 1e-3) / 1e6, 1),
        "shadow_rays": int(sh.shape[0]), "shadow_mrays_s": round(sh.shape[0] / (ms_sh * 1e-3) / 1e6, 1),
        "raygen_mrays_s": round(n / (ms_rg * 1e-3) / 1e6, 1),
        "hit_fraction": round(float(hit.mean()), 4),
    }
    log(f"analytic, demo scene: {out['niels_scene']}")
    sc.close()

    tris = wl.c1_scene()
    ana = wl.analytic_mix(5)
    lo = np.float32([-15, -15, -15])
    hi = np.float32([15, 15, 15])
    rays = wl.rays_uniform(wl.seed(77), lo, hi, 1 << 20)
    d = rays_to_device(rays)
    hits = torch.empty((rays.shape[0], 4), dtype=torch.int32, device="cuda")
    sc = Scene(ctx, tris, width=node_width)
    sc.trace_closest(d, hits)
    ms_tri = _events_ms(lambda: sc.trace_closest(d, hits), iters)
    sc.set_analytic(**ana)
    sc.trace_closest(d, hits)
    ms_mix = _events_ms(lambda: sc.trace_closest(d, hits), iters)
    nobj = sum(a.shape[0] for a in ana.values())
    out["mixed_4k"] = {
        "rays": int(rays.shape[0]), "triangles": int(tris.shape[0]), "analytic_objects": int(nobj),
        "closest_mrays_s": round(rays.shape[0] / (ms_mix * 1e-3) / 1e6, 1),
        "triangles_only_mrays_s": round(rays.shape[0] / (ms_tri * 1e-3) / 1e6, 1),
        "analytic_pass_ms": round(ms_mix - ms_tri, 4),
    }
    log(f"analytic, 4k triangles + {nobj} objects: {out['mixed_4k']}")
    sc.close()
    return out

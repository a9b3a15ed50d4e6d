"""The batched, path-tracing and serving workloads at their full sizes.

* Multi-scene batch: 64 scenes from the cluster generator with triangle
  counts log-uniform over [10k, 2M] (33M triangles), one ray batch per scene
  log-uniform over [1k, 1M] rays (6.3M rays), half closest hit and half
  occlusion, BVHs built on the GPU by LBVH, launch shape from the autotuner:
  every query equals its plain launch, and every 64th ray equals the
  oracle's walk of the scene's downloaded BVH.
* Path tracing: 1920x1080 at 4 spp, 3 bounces, on the 2,097,152-triangle
  displaced icosphere plus clusters: the graph replay equals the eager
  frame, the split BVH's image equals the unsplit one's, and 1,024 paths
  (256 pixels x 4 samples) equal the oracle's path tracer, whose hits come
  from the reference's linear scan over all 2M triangles.
* Serving: 32 resident scenes totalling ~32M triangles, Poisson arrivals at
  0.8 of the measured peak for 1 s, requests log-uniform over [256, 4M]
  rays, half closest hit and half occlusion, with a scene uploaded mid-run:
  every request equals the plain launch of its rays, and 1 in 16 requests is
  checked against the oracle's walk on a sample of its rays.
"""
import queue
import threading
import time

import numpy as np
import pytest

import oracle
from igx_raytracing_amd import Batch, PathTracer, Scene, screen_camera
from igx_raytracing_amd import drivers
from igx_raytracing_amd import workload as wl
from igx_raytracing_amd.serve import DONE, Server

from helpers import assert_hits_equal, hits_u32, to_dev

pytestmark = pytest.mark.gpu

NT = 16


def test_batch_64_scenes_full_size(gpu_ctx):
    import torch
    rng = np.random.default_rng(4)
    nscenes = 64
    tri_counts = drivers.log_uniform_int(rng, 10_000, 2_000_000, nscenes)
    ray_counts = drivers.log_uniform_int(rng, 1_000, 1_000_000, nscenes)
    kinds = ["closest" if k % 2 == 0 else "occluded" for k in range(nscenes)]
    print(f"batch: {int(tri_counts.sum())} triangles, {int(ray_counts.sum())} rays", flush=True)
    tri_sets = [drivers.cluster_scene(wl.seed(1000 + k), int(tri_counts[k])) for k in range(nscenes)]
    scenes, rays, queries, plain = [], [], [], []
    for k in range(nscenes):
        sc = Scene(gpu_ctx, device_tris=torch.from_numpy(tri_sets[k]).cuda(), width=4, keep_source=False)
        lo, hi = wl.bounds(tri_sets[k])
        r = wl.rays_uniform(wl.seed(2000 + k), lo, hi, int(ray_counts[k]),
                            tmax=None if kinds[k] == "closest" else (0.0, 20.0))
        d = to_dev(r)
        n = r.shape[0]
        if kinds[k] == "closest":
            out = torch.full((n, 4), 0x55, dtype=torch.int32, device="cuda")
            ref = sc.trace_closest(d)
        else:
            out = torch.full((n,), 0x55, dtype=torch.uint8, device="cuda")
            ref = sc.trace_occluded(d)
        scenes.append(sc)
        rays.append(r)
        queries.append((k, kinds[k], d, out))
        plain.append(ref)
    batch = Batch(gpu_ctx, scenes)
    opts, ms = batch.autotune(queries, iters=1)
    print(f"batch: tuned block {opts.block_size} ring {opts.lds_stack} chunk {opts.chunk} refill {opts.refill}, "
          f"{ms:.3f} ms", flush=True)
    for q in queries:
        q[3].fill_(0x55)
    batch.trace(queries, opts=opts)
    torch.cuda.synchronize()
    for (k, kind, d, out), ref in zip(queries, plain):
        if kind == "closest":
            assert_hits_equal(out, ref, f"batch scene {k}")
        else:
            assert torch.equal(out, ref), f"batch scene {k} occlusion"
    assert gpu_ctx.check_device_error() == 0
    print("batch: every query equals its plain launch; checking every 64th ray against the oracle", flush=True)
    checked = 0
    for k in range(nscenes):
        nodes, recs = scenes[k].download()
        sub = rays[k][::64]
        got = queries[k][3].cpu().numpy()[::64]
        want = oracle.walk(4, nodes, recs, scenes[k].info["root_ref"], sub, occlusion=kinds[k] == "occluded",
                           nthreads=NT)
        if kinds[k] == "closest":
            assert_hits_equal(got, want, f"batch scene {k} vs walk")
        else:
            np.testing.assert_array_equal(got, want, err_msg=f"batch scene {k} vs walk")
        checked += sub.shape[0]
    print(f"batch: {checked} rays equal the oracle's walk", flush=True)
    batch.close()
    for s in scenes:
        s.close()


def test_pathtrace_full_frame(gpu_ctx):
    import torch
    W, H, spp, depth = 1920, 1080, 4, 3
    tris = wl.c3_scene()
    cam = screen_camera((0.0, 4.0, -32.0), (0.0, 0.0, 0.0), 60.0, aspect=W / H)
    cam_arr = np.concatenate([np.float32(x) for x in cam] + [np.float32([0.5, 0.5])])
    print(f"pathtrace: {tris.shape[0]} triangles, {W}x{H}x{spp}", flush=True)
    images = {}
    for label, presplit in (("split", 16), ("unsplit", 1)):
        sc = Scene(gpu_ctx, tris, width=4, keep_source=False, presplit=presplit)
        pt = PathTracer(sc, cam, W, H, spp, max_depth=depth)
        img = torch.empty((H, W), dtype=torch.float32, device="cuda")
        eager = pt.render(3, use_graph=False).cpu().numpy()
        counts = pt.counts()
        for rep in range(2):
            pt.render(3, image=img, use_graph=True)
            np.testing.assert_array_equal(img.cpu().numpy().view(np.uint32), eager.view(np.uint32),
                                          err_msg=f"{label} graph replay {rep}")
            pt.render(4, use_graph=False)  # eager work between replays
        assert counts[0] == W * H * spp and counts[depth] > 0
        images[label] = eager
        print(f"pathtrace {label}: rays per depth {counts.tolist()}, mean {eager.mean():.5f}", flush=True)
        assert gpu_ctx.check_device_error() == 0
        pt.close()
        sc.close()
    np.testing.assert_array_equal(images["split"].view(np.uint32), images["unsplit"].view(np.uint32))
    px = np.random.default_rng(23).choice(W * H, 256, replace=False).astype(np.uint32)
    t0 = time.perf_counter()
    want = oracle.pathtrace_pixels(tris, cam_arr, W, H, spp, px, max_depth=depth, seed=3, nthreads=NT)
    print(f"pathtrace: 1,024 oracle paths over the linear scan in {time.perf_counter() - t0:.1f} s", flush=True)
    np.testing.assert_array_equal(images["split"].reshape(-1)[px].view(np.uint32), want.view(np.uint32))


def test_serving_32m_triangles_at_0_8_load(gpu_ctx):
    import torch
    rng = np.random.default_rng(5)
    nscenes, total, max_rays = 32, 32_000_000, 4 << 20
    counts = drivers.log_uniform_int(rng, 100_000, 2_000_000, nscenes).astype(np.float64)
    counts = np.maximum(4096, np.round(counts * total / counts.sum())).astype(np.int64)
    scenes = []
    for k in range(nscenes):
        d = torch.from_numpy(drivers.cluster_scene(wl.seed(3000 + k), int(counts[k]))).cuda()
        scenes.append(Scene(gpu_ctx, device_tris=d, width=4, keep_source=False))
        del d
    ntris = sum(s.info["num_tris"] for s in scenes)
    print(f"serve: {nscenes} scenes, {ntris} triangles", flush=True)
    box = (np.full(3, -55.0, np.float32), np.full(3, 55.0, np.float32))
    pool = {"closest": wl.rays_uniform(wl.seed(3999), box[0], box[1], max_rays),
            "occluded": wl.rays_uniform(wl.seed(3999), box[0], box[1], max_rays, tmax=(0.0, 20.0))}
    dpool = {k: to_dev(v) for k, v in pool.items()}
    # Plain launches of the whole pool on every scene: a request's rays are a
    # slice of the pool, so its plain result is the same slice of these.
    plain = {}
    for s in range(nscenes):
        plain[s, "closest"] = scenes[s].trace_closest(dpool["closest"])
        plain[s, "occluded"] = scenes[s].trace_occluded(dpool["occluded"])
    torch.cuda.synchronize()
    srv = Server(gpu_ctx, mem_budget=8 << 30, streams=4)
    ids = [srv.add_scene(s) for s in scenes]

    class Dev:
        device = True
        rays = dpool
        out = {"closest": torch.empty((max_rays, 4), dtype=torch.int32, device="cuda"),
               "occluded": torch.empty(max_rays, dtype=torch.uint8, device="cuda")}

        def slice(self, kind, off, n):
            return self.rays[kind][off:off + n], self.out[kind][:n]

    drivers._peak(srv, Dev(), ids, rng, max_rays, count=30)
    peak_rays, _ = drivers._peak(srv, Dev(), ids, rng, max_rays, count=200)
    del Dev.out
    mean_size = (max_rays - 256) / np.log(max_rays / 256)
    rate = 0.8 * peak_rays / mean_size
    duration = 1.0
    n_req = max(1, int(rate * duration))
    sizes, kinds, scs, offs = drivers._draw_requests(rng, n_req, nscenes, max_rays)
    gaps = rng.exponential(1.0 / rate, n_req)
    print(f"serve: peak {peak_rays / 1e6:.0f} Mrays/s; {n_req} requests at 0.8 load over {duration} s",
          flush=True)
    work = queue.Queue()
    failures, lat, sampled, mismatches = [], [], [], []

    def checker():
        while True:
            item = work.get()
            if item is None:
                return
            k, ticket, s, kind, off, n, out = item
            st, status, ms = srv.wait(ticket, 120_000)
            if st != DONE or status != 0:
                failures.append(f"request {k}: state {st} status {status}")
                continue
            lat.append(ms)
            if not torch.equal(out, plain[s, kind][off:off + n]):
                failures.append(f"request {k} (scene {s}, {kind}, {n} rays) differs from its plain launch")
                mismatches.append((k, s, kind, off, n, out.cpu().numpy(), plain[s, kind][off:off + n].cpu().numpy(),
                                   out))
            if k % 16 == 0:
                step = max(1, n // 2048)
                sampled.append((s, kind, off, step, out[::step].cpu().numpy()))
            srv.forget(ticket)

    th = threading.Thread(target=checker)
    th.start()
    upload_tris = drivers.cluster_scene(wl.seed(3900), 300_000)
    upload = {}
    t0 = time.perf_counter()
    at = t0
    for k in range(n_req):
        at += gaps[k]
        while time.perf_counter() < at:
            pass
        kind = "closest" if kinds[k] == 0 else "occluded"
        s, off, n = int(scs[k]), int(offs[k]), int(sizes[k])
        out = torch.empty((n, 4), dtype=torch.int32, device="cuda") if kind == "closest" else \
            torch.empty(n, dtype=torch.uint8, device="cuda")
        ticket = srv.submit(ids[s], kind, n, dpool[kind][off:off + n], out, device=True)
        work.put((k, ticket, s, kind, off, n, out))
        if k == n_req // 2:
            up = threading.Thread(target=lambda: upload.setdefault("id", srv.upload_scene(upload_tris)))
            up.start()
    work.put(None)
    th.join()
    up.join()
    wall = time.perf_counter() - t0
    stats = srv.stats()
    lat = np.array(lat)
    print(f"serve: {len(lat)} requests in {wall:.2f} s, {int(sizes.sum()) / wall / 1e6:.0f} Mrays/s, "
          f"p50 {np.percentile(lat, 50):.2f} ms p99 {np.percentile(lat, 99):.2f} ms, "
          f"{stats['mean_requests_per_tick']:.2f} requests per tick", flush=True)
    torch.cuda.synchronize()
    for k, s_, kind, off, n, got, ref, out in mismatches:
        # Which side is wrong: the served words, the plain launch, a plain
        # launch of only these rays, the oracle's walk and the linear scan;
        # and whether the served buffer still differs once the device is idle.
        rows = np.nonzero((got != ref).reshape(n, -1).any(axis=1))[0]
        later = np.nonzero((out.cpu().numpy() != ref).reshape(n, -1).any(axis=1))[0]
        print(f"mismatch: request {k} ({n} rays from {off}): {rows.size} rays differ, rows {rows[:3]}..{rows[-3:]}; "
              f"{later.size} differ when read again", flush=True)
        sub = pool[kind][off:off + n][rows]
        again = (scenes[s_].trace_closest if kind == "closest" else scenes[s_].trace_occluded)(to_dev(sub))
        nd, recs = scenes[s_].download()
        walk = oracle.walk(4, nd, recs, scenes[s_].info["root_ref"], sub, occlusion=kind == "occluded", nthreads=NT)
        tri_s = drivers.cluster_scene(wl.seed(3000 + s_), int(counts[s_]))
        scan = (oracle.closest_brute if kind == "closest" else oracle.occluded_brute)(tri_s, sub, nthreads=NT)
        for j, r in enumerate(rows[:4]):
            print(f"mismatch: request {k} scene {s_} {kind} ray {off + r}: served {got[r]} plain {ref[r]} "
                  f"alone {again.cpu().numpy()[j]} walk {walk[j]} scan {scan[j]} ray {sub[j]}", flush=True)
    assert not failures, failures[:5]
    assert len(lat) == n_req and stats["failed"] == 0
    # The scene uploaded mid-run serves like a directly built one.
    assert "id" in upload
    lo, hi = wl.bounds(upload_tris)
    r = wl.rays_uniform(wl.seed(3901), lo, hi, 65536)
    out = torch.empty((65536, 4), dtype=torch.int32, device="cuda")
    t = srv.submit(upload["id"], "closest", 65536, to_dev(r), out, device=True)
    assert srv.wait(t, 60_000)[0] == DONE
    direct = Scene(gpu_ctx, device_tris=torch.from_numpy(upload_tris.copy()).cuda(), width=4, keep_source=False)
    assert_hits_equal(out, direct.trace_closest(to_dev(r)), "uploaded scene")
    direct.close()
    srv.close()
    assert gpu_ctx.check_device_error() == 0
    # 1 in 16 requests against the oracle's walk of its scene's BVH.
    nodes = {}
    checked = 0
    for s, kind, off, step, got in sampled:
        if s not in nodes:
            nodes[s] = scenes[s].download()
        nd, recs = nodes[s]
        sub = pool[kind][off:off + got.shape[0] * step:step][:got.shape[0]]
        want = oracle.walk(4, nd, recs, scenes[s].info["root_ref"], sub, occlusion=kind == "occluded", nthreads=NT)
        if kind == "closest":
            assert_hits_equal(got, want, f"served scene {s} vs walk")
        else:
            np.testing.assert_array_equal(got, want)
        checked += sub.shape[0]
    print(f"serve: {len(sampled)} requests, {checked} rays equal the oracle's walk", flush=True)
    for s in scenes:
        s.close()

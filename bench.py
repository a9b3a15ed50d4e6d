"""Benchmark: closest-hit Mrays/s on 1,048,576 clustered triangles with
1,048,576 incoherent rays, on one MI355X.

A step is one closest-hit query over the whole ray batch (one kernel launch);
the BVH and the rays are resident in HBM before the timed region.  Prints one
JSON line (rank 0) with the roofline of the traversal kernel and a CPU
baseline measured on this host.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--width 2|4|8]
                       [--no-cpu] [--quick]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
RAY_BYTES, HIT_BYTES, TRI_BYTES = 32, 16, 36


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(counts, node_bytes, out_bytes_per_ray):
    """Sum over rays of 32 (ray) + output + node_bytes * nodes + 36 * triangle tests."""
    c = counts.astype(np.float64)
    n = counts.shape[0]
    return n * (RAY_BYTES + out_bytes_per_ray) + node_bytes * c[:, 0].sum() + TRI_BYTES * c[:, 1].sum()


def measured_profile(width):
    """(HBM bytes per launch, cache and occupancy counters) of the BVH4
    traversal kernel on this workload, as the copy-calibrated rocprofv3
    FETCH_SIZE + WRITE_SIZE runs and the L2/L1/occupancy counter runs
    measured them (profiles/k_trace_dyn_bvh4_traffic.json; PMC counters
    cannot be read from inside this process).  (None, None) for other node
    widths."""
    path = os.path.join(ROOT, "profiles", "k_trace_dyn_bvh4_traffic.json")
    if width != 4 or not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    return d["traffic_bytes_per_launch"], d.get("counters")


def time_launches(fn, steps, warmup, stream, min_warm_s=0.3):
    """Mean ms per launch over `steps` launches on `stream`, after `warmup`
    untimed launches and at least min_warm_s of untimed work: the launches
    run after seconds of host-side scene building, and the first ~20 of a
    run came out a few percent slower as the device came back to speed."""
    import torch
    t0 = time.perf_counter()
    k = 0
    while k < warmup or time.perf_counter() - t0 < min_warm_s:
        fn()
        k += 1
        if k % 16 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        fn()
    end.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps  # ms per launch


def cpu_baseline(tris, rays, min_seconds=10.0, c1_gpu_hits=None):
    """The oracle's binary-SAH traversal (the reference's triangle test and
    tie rule) on all host cores, repeated over the full ray batch until
    min_seconds have passed.  Also the reference's linear scan on its own
    512x512 case, checked against the GPU's hits for it when given."""
    import oracle
    import igx_raytracing_amd as rt
    nthreads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    bvh = rt.HostBvh(tris, width=2)
    root = bvh.info["root_ref"]
    reps, total = 0, 0.0
    while total < min_seconds or reps == 0:
        t0 = time.perf_counter()
        oracle.walk(2, bvh.nodes, bvh.tris, root, rays, nthreads=nthreads)
        total += time.perf_counter() - t0
        reps += 1
    value = reps * rays.shape[0] / total / 1e6
    # The reference's literal linear scan, timed on a small subsample.
    sub = rays[:256]
    t0 = time.perf_counter()
    oracle.closest_brute(tris, sub, nthreads=nthreads)
    brute = sub.shape[0] / (time.perf_counter() - t0) / 1e6
    # The reference's own scan on its own demo scene (3 triangles and 10
    # analytic objects), the workload of the analytic section.
    from igx_raytracing_amd import workload as wl
    ntris, nana = wl.niels_scene()
    nrays = wl.pinhole(1920, 1080, eye=(0.0, 5.0, -9.0), target=(0.0, 4.0, 0.0), vfov_deg=60.0)
    t0 = time.perf_counter()
    oracle.closest_brute(ntris, nrays, nthreads=nthreads, analytic=nana)
    niels = nrays.shape[0] / (time.perf_counter() - t0) / 1e6
    # The reference's own case: its literal scan over all 262,144 primary
    # rays, the answer the GPU's binary-SAH hits must equal bit for bit.
    c1_tris, c1_rays = wl.c1_scene(), wl.c1_rays()
    t0 = time.perf_counter()
    c1_want = oracle.closest_brute(c1_tris, c1_rays, nthreads=nthreads)
    c1_scan = c1_rays.shape[0] / (time.perf_counter() - t0) / 1e6
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    bvh.close()
    return {
        "value": round(value, 3), "unit": "Mrays/s", "cores": nthreads, "kind": "port",
        "sample": f"{reps} x {rays.shape[0]} rays, binary SAH walk of the oracle, {nthreads} threads, {model}",
        "bruteforce_mrays_s_extrapolated": round(brute, 6),
        "bruteforce_sample": f"{sub.shape[0]} rays x {tris.shape[0]} tris linear scan (the reference's loop)",
        "niels_scene_mrays_s": round(niels, 2),
        "niels_scene_sample": "1920x1080 primary rays, the reference's scan over its demo scene's 13 objects",
        "c1_scan_mrays_s": round(c1_scan, 3),
        "c1_sample": "512x512 primary rays x 4,096 triangles, the reference's linear scan",
        "c1_gpu_bitexact_vs_scan": None if c1_gpu_hits is None else bool(np.array_equal(c1_gpu_hits, c1_want)),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--width", type=int, default=4, choices=(2, 4, 8))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="skip the other node widths")
    ap.add_argument("--skip-batch", action="store_true", help="skip the 64-scene batch workload")
    ap.add_argument("--skip-serve", action="store_true", help="skip the serving workload")
    ap.add_argument("--skip-pathtrace", action="store_true", help="skip the wavefront path-tracing workload")
    ap.add_argument("--skip-analytic", action="store_true", help="skip the analytic-primitive workloads")
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="arrival window per offered load")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run a plain 256 MiB device copy (a kernel of known traffic, to scale PMC counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import torch
    import igx_raytracing_amd as rt
    from igx_raytracing_amd import workload as wl

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    log("generating 1,048,576 clustered triangles and 1,048,576 rays")
    tris = wl.c2_scene()
    rays = wl.c2_rays(tris, 1 << 20)
    occ = wl.c2_rays(tris, 1 << 20, occlusion=True)
    n = rays.shape[0]
    ctx = rt.Context(0)
    stream = torch.cuda.current_stream()
    d_rays = rt.rays_to_device(rays)
    d_occ = rt.rays_to_device(occ)
    hits = torch.empty((n, 4), dtype=torch.int32, device="cuda")
    flags = torch.empty(n, dtype=torch.uint8, device="cuda")

    if args.copy_probe:
        src = torch.empty(64 << 20, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    # The box's measured device-copy bandwidth (bytes read + written), the
    # practical ceiling the roofline's 8 TB/s nominal peak is also set against.
    src = torch.empty(128 << 20, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    copy_ms = time_launches(lambda: dst.copy_(src), 10, 2, torch.cuda.current_stream())
    copy_gbs = 2 * src.numel() * 4 / (copy_ms * 1e-3) / 1e9
    del src, dst

    per_width = {}
    widths = [args.width] if args.quick else [args.width] + [w for w in (2, 4, 8) if w != args.width]
    main_scene = None
    for w in widths:
        log(f"building binary SAH -> BVH{w}")
        sc = rt.Scene(ctx, tris, width=w, keep_source=False)
        ms = time_launches(lambda: sc.trace_closest(d_rays, hits, stream=stream), args.steps if w == args.width else
                           max(5, args.steps // 2), args.warmup, stream)
        ms_occ = time_launches(lambda: sc.trace_occluded(d_occ, flags, stream=stream),
                               max(5, args.steps // 2), args.warmup, stream)
        counts = torch.empty((n, 2), dtype=torch.int32, device="cuda")
        sc.trace_closest(d_rays, hits, stream=stream, opts=rt.trace_opts(), counts=counts)
        ocounts = torch.empty((n, 2), dtype=torch.int32, device="cuda")
        sc.trace_occluded(d_occ, flags, stream=stream, opts=rt.trace_opts(), counts=ocounts)
        torch.cuda.synchronize()
        cnt = counts.cpu().numpy().view(np.uint32)
        ocnt = ocounts.cpu().numpy().view(np.uint32)
        nb = sc.info["node_bytes"]
        b_closest = algorithmic_bytes(cnt, nb, HIT_BYTES)
        b_occ = algorithmic_bytes(ocnt, nb, 1)
        per_width[w] = {
            "closest_mrays_s": round(n / (ms * 1e-3) / 1e6, 2),
            "closest_ms": round(ms, 4),
            "occlusion_mrays_s": round(n / (ms_occ * 1e-3) / 1e6, 2),
            "occlusion_ms": round(ms_occ, 4),
            "nodes_per_ray": round(float(cnt[:, 0].mean()), 3),
            "tris_per_ray": round(float(cnt[:, 1].mean()), 3),
            "bytes_per_ray": round(b_closest / n, 1),
            "closest_gbs": round(b_closest / (ms * 1e-3) / 1e9, 1),
            "occlusion_gbs": round(b_occ / (ms_occ * 1e-3) / 1e9, 1),
            "node_bytes": nb,
            "num_nodes": sc.info["num_nodes"],
            "build_ms_host": round(sc.info["build_ms"], 1),
        }
        log(f"BVH{w}: {per_width[w]}")
        if w == args.width:
            main_scene = sc
        else:
            sc.close()
    err = ctx.check_device_error()
    hit_frac = float((rt.hits_to_numpy(hits)["id"] != rt.RT_MISS).mean())

    m = per_width[args.width]
    ms = m["closest_ms"]
    value = m["closest_mrays_s"]
    achieved = m["closest_gbs"]
    traffic, counters = measured_profile(args.width)
    result = {
        "metric": "Mrays/s closest-hit @1M tris incoherent; p50/p99 latency served vs plain",
        "value": value,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded PCG32 generators; no dataset)",
        "config": {
            "workload": "1,048,576 FP32 triangles in 256 Gaussian clusters; 1,048,576 incoherent rays "
                        "(origins uniform in the scene box, directions uniform on the sphere), closest hit",
            "node": f"BVH{args.width} (binary SAH collapsed, {m['node_bytes']} B nodes)",
            "rays": n,
            "triangles": int(tris.shape[0]),
            "hit_fraction": round(hit_frac, 4),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "counters": counters,
        "measured_copy_gbs": round(copy_gbs, 1),
        "achieved_frac_of_copy": round(achieved / copy_gbs, 4),
        "occlusion_mrays_s": m["occlusion_mrays_s"],
        "per_node_width": per_width,
        "device_error": err,
    }
    # LBVH build of the same 1M-triangle scene on the GPU.
    d_tris = torch.from_numpy(tris).cuda()
    rt.Scene(ctx, device_tris=d_tris, width=args.width, keep_source=False).close()  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lb = rt.Scene(ctx, device_tris=d_tris, width=args.width, keep_source=False)
    torch.cuda.synchronize()
    lbvh_ms = (time.perf_counter() - t0) * 1e3
    ms_lb = time_launches(lambda: lb.trace_closest(d_rays, hits, stream=stream), max(5, args.steps // 2),
                          args.warmup, stream)
    result["lbvh"] = {"build_ms": round(lbvh_ms, 2), "closest_mrays_s": round(n / (ms_lb * 1e-3) / 1e6, 2),
                      "num_nodes": lb.info["num_nodes"], "sah_host_build_ms": m["build_ms_host"]}
    lb.close()
    del d_tris

    from igx_raytracing_amd import drivers
    if not args.skip_batch and not args.quick:
        log("multi-scene batch workload")
        result["batch"] = drivers.batch_workload(ctx, log=log)
    if not args.skip_pathtrace and not args.quick:
        log("wavefront path-tracing workload")
        result["pathtrace"] = drivers.pathtrace_workload(ctx, log=log)
    if not args.skip_analytic and not args.quick:
        log("analytic-primitive workloads")
        result["analytic"] = drivers.analytic_workload(ctx, log=log)
    c1_hits = None
    if not args.quick:
        log("the reference's 512x512 case")
        result["c1"], c1_hits = drivers.c1_workload(ctx, log=log)
    if not args.skip_serve and not args.quick:
        log("serving workload")
        sv = drivers.serve_workload(ctx, duration_s=args.serve_seconds, log=log)
        result["serve"] = sv
        for mode in ("served", "plain"):
            for load, e in sv["modes"][mode]["loads"].items():
                result[f"{mode}_p50_ms@{load}"] = e["p50_ms"]
                result[f"{mode}_p99_ms@{load}"] = e["p99_ms"]
                result[f"{mode}_mrays_s@{load}"] = e["mrays_s"]
        for load, e in sv["modes"]["plain"].get("loads_of_own_peak", {}).items():
            result[f"plain_p50_ms@{load}_of_own_peak"] = e["p50_ms"]
            result[f"plain_p99_ms@{load}_of_own_peak"] = e["p99_ms"]
        # Requests whose rays are already in HBM (no PCIe copies).
        dv = sv["device_requests"]
        for mode in ("served", "plain"):
            result[f"{mode}_device_peak_mrays_s"] = dv[mode]["peak_mrays_s"]
            for load, e in dv[mode]["loads"].items():
                result[f"{mode}_device_p50_ms@{load}"] = e["p50_ms"]
                result[f"{mode}_device_p99_ms@{load}"] = e["p99_ms"]
        result["served_vs_plain_device_peak"] = round(dv["served"]["peak_mrays_s"] / dv["plain"]["peak_mrays_s"], 2)

    # Any stack overflow or other device-side error in the sections above.
    result["device_error_end"] = ctx.check_device_error()
    if not args.no_cpu:
        log("CPU baseline")
        result["cpu_baseline"] = cpu_baseline(tris, rays, c1_gpu_hits=c1_hits)
    else:
        result["cpu_baseline"] = None
    main_scene.close()
    ctx.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

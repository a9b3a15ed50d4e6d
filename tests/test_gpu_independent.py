"""GPU parity that does not rest on the GPU's own triangle test.

The GPU linear scan (k_brute) calls the same Moller-Trumbore routine as
every traversal kernel, so a defect there would cancel out of a GPU-vs-GPU
comparison.  These tests compare the kernels with the CPU oracle instead:

* all 1,048,576 incoherent rays of the 1M-triangle clustered scene, closest
  hit and occlusion, for BVH2/BVH4/BVH8 built by the host SAH builder and by
  the GPU LBVH builder, against the oracle's walk of the same node bytes (its
  triangle test, orc_mt_f32, is pinned by known answers and FP64
  reconstruction, and its walk by the linear scan);
* 4,096 of those rays against the oracle's linear scan over all 1,048,576
  triangles (the reference's algorithm itself);
* adversarial rays (grazing, axis-aligned from exact vertex / edge / face
  positions, far origins) on every builder and width against the oracle's
  linear scan, with FP32 artifacts of the scan adjudicated as in
  tests/helpers.py::assert_hits_adjudicated.
"""
import numpy as np
import pytest

import oracle
from igx_raytracing_amd import Scene
from igx_raytracing_amd import workload as wl

import adversarial as adv
from helpers import (assert_flags_adjudicated, assert_hits_adjudicated, assert_hits_equal, hits_u32, scene_pad,
                     to_dev)

pytestmark = pytest.mark.gpu

WIDTHS = (2, 4, 8)
NT = 16


def _differing_rays_adjudicated(tris, rays, got, want, what, occlusion=False):
    """Rays where the kernel and the oracle's walk differ (their visiting
    orders differ, so an FP32 artifact of one leaf can be reached by one and
    culled by the other) are checked against the linear scan instead."""
    if occlusion:
        bad = np.nonzero(got != want)[0]
    else:
        bad = np.nonzero((hits_u32(got) != hits_u32(want)).any(axis=1))[0]
    assert bad.size <= 16, f"{what}: {bad.size} rays differ from the oracle's walk"
    if bad.size == 0:
        return 0
    sub = rays[bad]
    guard = scene_pad(tris) / 2
    if occlusion:
        assert_flags_adjudicated(got[bad], oracle.occluded_brute(tris, sub, nthreads=NT),
                                 oracle.occluded_brute_guard(tris, sub, guard, NT), what)
    else:
        assert_hits_adjudicated(hits_u32(got)[bad], tris, sub, oracle.closest_brute(tris, sub, nthreads=NT),
                                oracle.closest_brute_guard(tris, sub, guard, NT),
                                oracle.closest_brute_guard(tris, sub, float("inf"), NT), what)
    return int(bad.size)


@pytest.mark.parametrize("builder", ["sah", "lbvh"])
def test_c2_all_rays_vs_oracle_walk(gpu_ctx, c2_data, builder):
    tris, rays, occ = c2_data
    d, do = to_dev(rays), to_dev(occ)
    for width in WIDTHS:
        print(f"c2 all rays vs walk: {builder} BVH{width}", flush=True)
        sc = Scene(gpu_ctx, tris, width=width, builder=builder, keep_source=False)
        nodes, recs = sc.download()
        got = hits_u32(sc.trace_closest(d))
        want = oracle.walk(width, nodes, recs, sc.info["root_ref"], rays, nthreads=NT)
        n1 = _differing_rays_adjudicated(tris, rays, got, want, f"C2 {builder} BVH{width} closest")
        gflags = sc.trace_occluded(do).cpu().numpy()
        wflags = oracle.walk(width, nodes, recs, sc.info["root_ref"], occ, occlusion=True, nthreads=NT)
        n2 = _differing_rays_adjudicated(tris, occ, gflags, wflags, f"C2 {builder} BVH{width} occlusion", True)
        print(f"  hits {np.mean(got[:, 0] != 0xFFFFFFFF):.4f}, occluded {gflags.mean():.4f}, "
              f"adjudicated {n1} + {n2}", flush=True)
        assert gpu_ctx.check_device_error() == 0
        sc.close()


def test_c2_sample_vs_linear_scan(gpu_ctx, c2_data):
    """4,096 rays against the reference's own algorithm on the CPU: every ray
    scans all 1,048,576 triangles."""
    tris, rays, occ = c2_data
    idx = np.random.default_rng(17).choice(rays.shape[0], 4096, replace=False)
    sub, osub = rays[idx], occ[idx]
    print("c2 4096-ray linear scan on the CPU", flush=True)
    want = oracle.closest_brute(tris, sub, nthreads=NT)
    wflags = oracle.occluded_brute(tris, osub, nthreads=NT)
    print(f"  hits {np.mean(want[:, 0] != 0xFFFFFFFF):.4f}, occluded {wflags.mean():.4f}", flush=True)
    for builder in ("sah", "lbvh"):
        for width in WIDTHS:
            sc = Scene(gpu_ctx, tris, width=width, builder=builder, keep_source=False)
            assert_hits_equal(sc.trace_closest(to_dev(sub)), want, f"C2 sample {builder} BVH{width}")
            np.testing.assert_array_equal(sc.trace_occluded(to_dev(osub)).cpu().numpy(), wflags)
            sc.close()


def _adv_scenes():
    return {"mixed4k": (wl.c1_scene(), 20000), "clusters64k": (wl.clusters(wl.seed(7), 16, 4096), 6000)}


@pytest.fixture(scope="module", params=sorted(_adv_scenes()))
def adversarial_case(request):
    tris, n = _adv_scenes()[request.param]
    guard = scene_pad(tris) / 2
    out = {}
    for name, (cr, orays) in adv.all_sets(tris, n, 99).items():
        out[name] = (cr, orays, oracle.closest_brute(tris, cr, nthreads=NT),
                     oracle.closest_brute_guard(tris, cr, guard, NT),
                     oracle.closest_brute_guard(tris, cr, float("inf"), NT),
                     oracle.occluded_brute(tris, orays, nthreads=NT),
                     oracle.occluded_brute_guard(tris, orays, guard, NT))
    return request.param, tris, out


@pytest.mark.parametrize("builder", ["sah", "lbvh"])
def test_adversarial_rays_vs_linear_scan(gpu_ctx, adversarial_case, builder):
    name, tris, sets = adversarial_case
    for width in WIDTHS:
        sc = Scene(gpu_ctx, tris, width=width, builder=builder, keep_source=False)
        total = 0
        for sname, (cr, orays, full, guarded, lower, ofull, oguarded) in sets.items():
            what = f"{name} {sname} {builder} BVH{width}"
            total += assert_hits_adjudicated(sc.trace_closest(to_dev(cr)), tris, cr, full, guarded, lower, what)
            total += assert_flags_adjudicated(sc.trace_occluded(to_dev(orays)).cpu().numpy(), ofull, oguarded,
                                              what + " occlusion")
        print(f"{name} {builder} BVH{width}: {total} adjudicated rays", flush=True)
        assert gpu_ctx.check_device_error() == 0
        sc.close()

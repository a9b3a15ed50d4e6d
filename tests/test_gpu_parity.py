"""GPU parity: the HIP query kernels, called through the C ABI, against the
oracle (CPU restatement of trace.glsl / primitive.glsl) and against the
reference's own linear scan run on the GPU.

Bar: hit ids, miss flags and occlusion flags bit-exact; t, u, v bit-exact
(kernels and oracle evaluate the same FP32 operations in the same order with
contraction off, so the 2-ulp allowance is not needed and not used)."""
import numpy as np
import pytest

import oracle
from igx_raytracing_amd import RAY_DTYPE, RT_MISS, Scene, hits_to_numpy, trace_opts
from igx_raytracing_amd import workload as wl

from helpers import NIELS_TRIS, assert_hits_equal, hits_u32, shadow_rays, to_dev

pytestmark = pytest.mark.gpu

WIDTHS = (2, 4, 8)


@pytest.fixture(scope="module")
def c1():
    tris = wl.c1_scene()
    rays = wl.c1_rays()
    want = oracle.closest_brute(tris, rays, nthreads=16)
    return tris, rays, want


@pytest.mark.parametrize("width", WIDTHS)
def test_c1_closest_bitexact_vs_oracle(gpu_ctx, c1, width):
    tris, rays, want = c1
    print(f"c1 closest width={width}", flush=True)
    sc = Scene(gpu_ctx, tris, width=width)
    got = sc.trace_closest(to_dev(rays))
    assert_hits_equal(got, want, f"C1 BVH{width}")
    assert gpu_ctx.check_device_error() == 0
    sc.close()


@pytest.mark.parametrize("width", WIDTHS)
def test_c1_shadow_rays_vs_oracle(gpu_ctx, c1, width):
    tris, rays, prim = c1
    srays = shadow_rays(rays, prim, light=(6.0, 15.0, -12.0))
    want = oracle.occluded_brute(tris, srays, nthreads=16)
    print(f"c1 occlusion width={width} n={srays.shape[0]} occluded={want.mean():.3f}", flush=True)
    sc = Scene(gpu_ctx, tris, width=width)
    d = to_dev(srays)
    got = sc.trace_occluded(d).cpu().numpy()
    np.testing.assert_array_equal(got, want)
    # wave64 ballot layout: bit (i & 63) of word i >> 6
    masks = sc.trace_occluded_mask(d).cpu().numpy().view(np.uint64)
    bits = ((masks[np.arange(got.size) >> 6] >> (np.arange(got.size) & 63).astype(np.uint64)) & 1).astype(np.uint8)
    np.testing.assert_array_equal(bits, want)
    sc.close()


def test_gpu_linear_scan_vs_oracle(gpu_ctx, c1):
    """The reference's own algorithm on the GPU equals its CPU restatement."""
    tris, rays, want = c1
    sub = rays[::7]
    sc = Scene(gpu_ctx, tris, width=4)
    got = sc.trace_closest_brute(to_dev(sub))
    assert_hits_equal(got, want[::7], "GPU linear scan")
    srays = shadow_rays(rays, want, light=(6.0, 15.0, -12.0))[::3]
    np.testing.assert_array_equal(sc.trace_occluded_brute(to_dev(srays)).cpu().numpy(),
                                  oracle.occluded_brute(tris, srays, nthreads=16))
    sc.close()


@pytest.mark.parametrize("width", WIDTHS)
def test_c2_walk_and_counts_bitexact(gpu_ctx, c2_data, width):
    """Same BVH bytes, same visiting order: results and per-ray node/triangle
    counts equal the oracle's traversal restatement; a brute-force subsample
    pins the traversal to the reference's scan."""
    tris, rays, occ = c2_data
    n = 1 << 16
    print(f"c2 walk width={width}", flush=True)
    sc = Scene(gpu_ctx, tris, width=width)
    nodes, recs = sc.download()
    import torch
    d = to_dev(rays[:n])
    counts = torch.empty((n, 2), dtype=torch.int32, device="cuda")
    got = sc.trace_closest(d, opts=trace_opts(), counts=counts)
    want, wcnt = oracle.walk(width, nodes, recs, sc.info["root_ref"], rays[:n], counts=True, nthreads=16)
    assert_hits_equal(got, want, f"C2 BVH{width} vs walk")
    np.testing.assert_array_equal(counts.cpu().numpy().view(np.uint32), wcnt)
    brute = oracle.closest_brute(tris, rays[:256], nthreads=16)
    assert_hits_equal(hits_u32(got)[:256], brute, f"C2 BVH{width} vs brute")
    # occlusion, with counts
    do = to_dev(occ[:n])
    ocnt = torch.empty((n, 2), dtype=torch.int32, device="cuda")
    gflags = sc.trace_occluded(do, opts=trace_opts(), counts=ocnt).cpu().numpy()
    wflags, wocnt = oracle.walk(width, nodes, recs, sc.info["root_ref"], occ[:n], occlusion=True, counts=True,
                                nthreads=16)
    np.testing.assert_array_equal(gflags, wflags)
    np.testing.assert_array_equal(ocnt.cpu().numpy().view(np.uint32), wocnt)
    np.testing.assert_array_equal(gflags[:256], oracle.occluded_brute(tris, occ[:256], nthreads=16))
    assert gpu_ctx.check_device_error() == 0
    sc.close()


def test_c2_full_vs_gpu_linear_scan(gpu_ctx, c2_data):
    """All 1,048,576 incoherent rays: BVH2/BVH4/BVH8 closest and occlusion
    results equal the reference's linear scan over all 1,048,576 triangles."""
    tris, rays, occ = c2_data
    import torch
    sc = Scene(gpu_ctx, tris, width=4)
    d, do = to_dev(rays), to_dev(occ)
    print("c2 full linear scan (closest)", flush=True)
    ref_h = sc.trace_closest_brute(d)
    print("c2 full linear scan (occlusion)", flush=True)
    ref_o = sc.trace_occluded_brute(do)
    torch.cuda.synchronize()
    ref_h = hits_u32(ref_h)
    ref_o = ref_o.cpu().numpy()
    print(f"c2 hits {np.mean(ref_h[:, 0] != RT_MISS):.4f} occluded {ref_o.mean():.4f}", flush=True)
    for width in WIDTHS:
        s2 = sc if width == 4 else Scene(gpu_ctx, tris, width=width, keep_source=False)
        assert_hits_equal(s2.trace_closest(d), ref_h, f"C2 full BVH{width}")
        np.testing.assert_array_equal(s2.trace_occluded(do).cpu().numpy(), ref_o)
        if s2 is not sc:
            s2.close()
    # occluded == (closest.t < tmax) for the same origins and directions
    t = ref_h[:, 1].view(np.float32)
    np.testing.assert_array_equal(ref_o, (t < occ["tmax"]).astype(np.uint8))
    assert gpu_ctx.check_device_error() == 0
    sc.close()


@pytest.mark.parametrize("lds_stack", (1, 4, 8, 16, 32))
def test_stack_ring_size_does_not_change_results(gpu_ctx, c2_data, lds_stack):
    """An LDS ring of one entry spills everything to the global stack; no
    entry is ever dropped, so results stay bit-exact."""
    tris, rays, occ = c2_data
    n = 1 << 15
    sc = Scene(gpu_ctx, tris, width=8, keep_source=False)
    d = to_dev(rays[:n])
    base = hits_u32(sc.trace_closest(d))
    got = sc.trace_closest(d, opts=trace_opts(lds_stack=lds_stack))
    assert_hits_equal(got, base, f"lds_stack={lds_stack}")
    do = to_dev(occ[:n])
    np.testing.assert_array_equal(sc.trace_occluded(do, opts=trace_opts(lds_stack=lds_stack)).cpu().numpy(),
                                  sc.trace_occluded(do).cpu().numpy())
    assert gpu_ctx.check_device_error() == 0
    sc.close()


@pytest.mark.parametrize("width", (2, 4))
def test_two_node_steps_stay_within_the_stack_bound(gpu_ctx, width):
    """Closest-hit steps that test two nodes push more entries than the
    one-node walk the builder's max_stack bounds; the spill depth and the
    size limit for such steps must keep every stack within capacity.  Rays
    that start on a displaced sphere and leave it at random angles (diffuse
    bounces) graze many boxes at every level and build deep stacks; a
    one-entry LDS ring makes the capacity exactly what the bound provides.
    Results must equal the oracle's one-node walk, and no stack may
    overflow (the device error word stays 0)."""
    tris = wl.icosphere(wl.seed(2), 7, 10.0, 0.5)
    rng = np.random.default_rng(17)
    n = 1 << 16
    pick = rng.integers(0, tris.shape[0], n)
    bary = rng.dirichlet((1.0, 1.0, 1.0), n).astype(np.float32)
    t9 = tris[pick].reshape(n, 3, 3)
    org = (bary[:, :, None] * t9).sum(1).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros(n, RAY_DTYPE)
    rays["org"] = org
    rays["dir"] = d.astype(np.float32)
    rays["tmax"] = np.float32(3.4028235e38)
    rays["prev_id"] = pick.astype(np.uint32)
    sc = Scene(gpu_ctx, tris, width=width, keep_source=False)
    nodes, recs = sc.download()
    want = oracle.walk(width, nodes, recs, sc.info["root_ref"], rays)
    dr = to_dev(rays)
    for lds_stack in (1, 4, 16):
        for sched in ("dynamic", "static"):
            got = sc.trace_closest(dr, opts=trace_opts(lds_stack=lds_stack, schedule=sched))
            assert_hits_equal(got, want, f"lds_stack={lds_stack} {sched}")
            assert gpu_ctx.check_device_error() == 0, f"stack overflow at lds_stack={lds_stack} {sched}"
    sc.close()


def test_persistent_grid_size_does_not_change_results(gpu_ctx, c2_data):
    tris, rays, _ = c2_data
    n = 100_003
    sc = Scene(gpu_ctx, tris, width=4, keep_source=False)
    d = to_dev(rays[:n])
    base = hits_u32(sc.trace_closest(d))
    for wpc in (4, 8, 32):
        assert_hits_equal(sc.trace_closest(d, opts=trace_opts(waves_per_cu=wpc)), base, f"waves_per_cu={wpc}")
    sc.close()


def test_niels_scene_triangles(gpu_ctx):
    """The reference demo scene's three triangles (niels_scene.cpp:26-42)."""
    tris = NIELS_TRIS
    eye = np.float32([0.3, 2.0, -6.0])
    rng = np.random.default_rng(5)
    n = 4096
    rays = np.empty(n, RAY_DTYPE)
    tgt = rng.uniform([-1.2, -0.2, -0.2], [1.2, 7.2, 1.2], size=(n, 3)).astype(np.float32)
    dvec = tgt - eye
    rays["org"] = eye
    rays["dir"] = (dvec / np.linalg.norm(dvec, axis=1, keepdims=True)).astype(np.float32)
    rays["tmax"] = np.float32(3.4028235e38)
    rays["prev_id"] = np.where(rng.random(n) < 0.2, rng.integers(0, 3, n), RT_MISS).astype(np.uint32)
    want = oracle.closest_brute(tris, rays)
    assert (want[:, 0] != RT_MISS).mean() > 0.2
    for width in WIDTHS:
        sc = Scene(gpu_ctx, tris, width=width)
        assert_hits_equal(sc.trace_closest(to_dev(rays)), want, f"niels BVH{width}")
        sc.close()


def rays_at(tris, n, seed):
    """Rays toward random points on the given triangles from random origins
    1-10 units away, tmax 0.5-1.5 times the distance: about half are
    occluded."""
    rng = np.random.default_rng(seed)
    k = rng.integers(0, tris.shape[0], n)
    b = rng.random((n, 2))
    flip = b.sum(1) > 1
    b[flip] = 1 - b[flip]
    v = tris[k].reshape(n, 3, 3).astype(np.float64)
    p = v[:, 0] + b[:, :1] * (v[:, 1] - v[:, 0]) + b[:, 1:] * (v[:, 2] - v[:, 0])
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    dist = rng.uniform(1, 10, n)
    r = np.zeros(n, RAY_DTYPE)
    r["org"] = p - d * dist[:, None]
    r["dir"] = d
    r["tmax"] = dist * rng.uniform(0.5, 1.5, n)
    r["prev_id"] = RT_MISS
    return r


def test_edge_cases(gpu_ctx):
    import torch
    tris = wl.c1_scene()
    sc = Scene(gpu_ctx, tris, width=4)
    base = wl.c1_rays(64, 64)[::17].copy()
    cases = []
    for tmax in (0.0, -1.0, np.nan, np.inf, 1e-3, 20.0):
        r = base.copy()
        r["tmax"] = np.float32(tmax)
        cases.append(r)
    rays = np.concatenate(cases)
    want = oracle.closest_brute(tris, rays)
    assert_hits_equal(sc.trace_closest(to_dev(rays)), want, "tmax edge cases")
    wocc = oracle.occluded_brute(tris, rays)
    np.testing.assert_array_equal(sc.trace_occluded(to_dev(rays)).cpu().numpy(), wocc)
    # an infinite tmax is always occluded (trace.glsl:97: FLT_MAX < inf)
    assert wocc[len(base) * 3:len(base) * 4].all()
    # single ray, zero rays, odd counts
    one = to_dev(base[:1])
    assert_hits_equal(sc.trace_closest(one), oracle.closest_brute(tris, base[:1]), "one ray")
    empty = torch.empty((0, 8), dtype=torch.float32, device="cuda")
    assert sc.trace_closest(empty).shape == (0, 4)
    assert sc.trace_occluded(empty).shape == (0,)
    sc.close()
    # one triangle and three (the root is a leaf: no root node for the
    # any-hit kernels to copy), five (an interior root), and an empty scene;
    # closest hit and occlusion on both schedules
    for t in (tris[:1], tris[:3], tris[:5], np.zeros((0, 9), np.float32)):
        for width in (2, 4, 8):
            s1 = Scene(gpu_ctx, t, width=width)
            r = wl.c1_rays(32, 32)
            got = s1.trace_closest(to_dev(r))
            if t.shape[0]:
                assert_hits_equal(got, oracle.closest_brute(t, r), f"{t.shape[0]} triangles")
            else:
                assert (hits_u32(got)[:, 0] == RT_MISS).all()
            occ = rays_at(t if t.shape[0] else tris[:1], 1024, seed=3)
            wocc = oracle.occluded_brute(t, occ)
            if t.shape[0]:
                assert 0.1 < wocc.mean() < 0.9, f"{t.shape[0]} triangles: occluded fraction {wocc.mean()}"
            for sched in ("static", "dynamic"):
                np.testing.assert_array_equal(
                    s1.trace_occluded(to_dev(occ), opts=trace_opts(schedule=sched)).cpu().numpy(), wocc,
                    err_msg=f"occlusion {t.shape[0]} triangles BVH{width} {sched}")
            s1.close()


def test_ray_order_independence(gpu_ctx, c2_data):
    tris, rays, _ = c2_data
    n = 1 << 16
    sc = Scene(gpu_ctx, tris, width=4, keep_source=False)
    a = hits_u32(sc.trace_closest(to_dev(rays[:n])))
    perm = np.random.default_rng(1).permutation(n)
    b = hits_u32(sc.trace_closest(to_dev(rays[:n][perm])))
    np.testing.assert_array_equal(b, a[perm])
    sc.close()


@pytest.mark.parametrize("width", WIDTHS)
def test_dynamic_schedule_equals_static(gpu_ctx, c2_data, width):
    """Idle lanes refetching rays from the work counter change who traces a
    ray, never its result; repeated launches check the counter is re-zeroed
    by the last wave of each launch."""
    tris, rays, occ = c2_data
    n = 300_001
    sc = Scene(gpu_ctx, tris, width=width, keep_source=False)
    d, do = to_dev(rays[:n]), to_dev(occ[:n])
    ref = hits_u32(sc.trace_closest(d, opts=trace_opts(schedule="static")))
    ref_o = sc.trace_occluded(do, opts=trace_opts(schedule="static")).cpu().numpy()
    for refill in (1, 16, 64):
        for lds in (1, 8):
            o = trace_opts(schedule="dynamic", refill=refill, lds_stack=lds)
            for _ in range(2):
                assert_hits_equal(sc.trace_closest(d, opts=o), ref, f"dynamic refill={refill} lds={lds}")
                np.testing.assert_array_equal(sc.trace_occluded(do, opts=o).cpu().numpy(), ref_o)
    # small launches: fewer rays than lanes in the grid
    for m in (1, 63, 64, 65, 1000):
        assert_hits_equal(sc.trace_closest(to_dev(rays[:m])), ref[:m], f"n={m}")
    assert gpu_ctx.check_device_error() == 0
    sc.close()


def test_two_streams_at_once(gpu_ctx, c2_data):
    """Each stream has its own spill stack and work counter."""
    import torch
    tris, rays, _ = c2_data
    n = 1 << 18
    sc = Scene(gpu_ctx, tris, width=4, keep_source=False)
    d = to_dev(rays[:n])
    ref = hits_u32(sc.trace_closest(d))
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    outs = []
    for _ in range(3):
        h1 = sc.trace_closest(d, stream=s1)
        h2 = sc.trace_closest(d, stream=s2)
        outs += [(s1, h1), (s2, h2)]
    torch.cuda.synchronize()
    for _, h in outs:
        assert_hits_equal(h, ref, "concurrent streams")
    sc.close()

"""Analytic primitives on the GPU (spheres, cubes, planes after the triangles
in the object-index space, trace.glsl:22-46), bit-exact against the oracle's
literal scan: closest hit and occlusion through every query entry point
(BVH2/4/8 traversal then the analytic pass, the GPU linear scan, ballot
masks), the raygen packing with analytic geometry normals, and the shadow
tiles.  Scenes: the reference's own demo scene (NielsScene: 3 triangles,
7 spheres, 2 cubes, 1 plane) and a 4,096-triangle scene with 64 spheres,
32 cubes and 2 planes."""
import numpy as np
import pytest

import oracle
from igx_raytracing_amd import RAY_DTYPE, RT_MISS, Scene, screen_camera, trace_opts
from igx_raytracing_amd import workload as wl

from helpers import NIELS_ANALYTIC, NIELS_TRIS, assert_hits_equal, shadow_rays, to_dev

pytestmark = pytest.mark.gpu
F32_MAX = np.float32(3.4028235e38)


def random_rays(n, seed, lo=-8.0, hi=8.0, tmax=None, prev_max=None):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, RAY_DTYPE)
    r["org"] = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    r["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r["tmax"] = F32_MAX if tmax is None else rng.uniform(tmax[0], tmax[1], n)
    r["prev_id"] = RT_MISS if prev_max is None else np.where(rng.random(n) < 0.3, rng.integers(0, prev_max, n),
                                                             RT_MISS)
    return r


def big_analytic(seed):
    return wl.analytic_mix(seed)


def camera_rays(W, H, eye=(0.0, 5.0, -9.0), target=(0.0, 4.0, 0.0)):
    return wl.pinhole(W, H, eye=eye, target=target, vfov_deg=60.0)


@pytest.mark.parametrize("width", (2, 4, 8))
def test_niels_scene_closest_and_occlusion(gpu_ctx, width):
    import torch
    rays = camera_rays(96, 64)
    want = oracle.closest_brute(NIELS_TRIS, rays, analytic=NIELS_ANALYTIC)
    ids = want[:, 0]
    nt, ns, nc = 3, 7, 2
    # every object class is seen
    assert ((ids < nt).sum() > 10 and ((ids >= nt) & (ids < nt + ns)).sum() > 10
            and ((ids >= nt + ns) & (ids < nt + ns + nc)).sum() > 10 and (ids == nt + ns + nc).sum() > 10)
    sc = Scene(gpu_ctx, NIELS_TRIS, width=width)
    sc.set_analytic(**NIELS_ANALYTIC)
    assert (sc.info["num_spheres"], sc.info["num_cubes"], sc.info["num_planes"]) == (7, 2, 1)
    d = to_dev(rays)
    for opts in (None, trace_opts(schedule="static"), trace_opts(schedule="dynamic", lds_stack=8)):
        assert_hits_equal(sc.trace_closest(d, opts=opts), want, f"closest {opts}")
    assert_hits_equal(sc.trace_closest_brute(d), want, "linear scan")
    # shadow rays from every hit toward a point light, skipping the hit object
    sh = shadow_rays(rays, want, (0.0, 12.0, 0.0))
    want_o = oracle.occluded_brute(NIELS_TRIS, sh, analytic=NIELS_ANALYTIC)
    assert 0.05 < want_o.mean() < 0.95
    ds = to_dev(sh)
    for opts in (None, trace_opts(schedule="dynamic"), trace_opts(schedule="static")):
        np.testing.assert_array_equal(sc.trace_occluded(ds, opts=opts).cpu().numpy(), want_o)
    np.testing.assert_array_equal(sc.trace_occluded_brute(ds).cpu().numpy(), want_o)
    masks = sc.trace_occluded_mask(ds).cpu().numpy().view(np.uint64)
    bits = ((masks[np.arange(sh.shape[0]) >> 6] >> (np.arange(sh.shape[0]) & 63).astype(np.uint64)) & 1)
    np.testing.assert_array_equal(bits.astype(np.uint8), want_o)
    # ragged tail: the last mask word's unused bits are zero
    tail = sh.shape[0] & 63
    if tail:
        assert int(masks[-1]) >> tail == 0
    # flags and ballot words requested together
    fl = torch.empty(sh.shape[0], dtype=torch.uint8, device="cuda")
    mk = torch.empty((sh.shape[0] + 63) // 64, dtype=torch.int64, device="cuda")
    sc.trace_occluded(ds, flags=fl, masks=mk)
    np.testing.assert_array_equal(fl.cpu().numpy(), want_o)
    np.testing.assert_array_equal(mk.cpu().numpy().view(np.uint64), masks)
    assert gpu_ctx.check_device_error() == 0
    sc.close()
    torch.cuda.synchronize()


@pytest.mark.parametrize("width", (4, 8))
def test_many_analytic_random_rays(gpu_ctx, width):
    tris = wl.c1_scene()
    ana = big_analytic(5)
    ntot = tris.shape[0] + 64 + 32 + 2
    rays = random_rays(20000, 9, lo=-15, hi=15, prev_max=ntot)
    want = oracle.closest_brute(tris, rays, analytic=ana)
    ids = want[:, 0].astype(np.int64)
    assert ((ids >= 4096) & (ids < 4096 + 64)).sum() > 200 and ((ids >= 4160) & (ids < 4192)).sum() > 200
    assert (ids < 4096).sum() > 200 and (ids == RT_MISS).sum() > 50
    sc = Scene(gpu_ctx, tris, width=width)
    sc.set_analytic(**ana)
    d = to_dev(rays)
    assert_hits_equal(sc.trace_closest(d), want, "closest")
    assert_hits_equal(sc.trace_closest_brute(d), want, "linear scan")
    # finite tmax on closest queries filters the scan's result
    rt = rays.copy()
    rt["tmax"] = np.random.default_rng(2).uniform(0, 25, rt.shape[0])
    assert_hits_equal(sc.trace_closest(to_dev(rt)), oracle.closest_brute(tris, rt, analytic=ana), "closest tmax")
    occ = random_rays(20000, 10, lo=-15, hi=15, tmax=(0, 20), prev_max=ntot)
    want_o = oracle.occluded_brute(tris, occ, analytic=ana)
    do = to_dev(occ)
    np.testing.assert_array_equal(sc.trace_occluded(do).cpu().numpy(), want_o)
    np.testing.assert_array_equal(sc.trace_occluded(do, opts=trace_opts(schedule="dynamic")).cpu().numpy(), want_o)
    # replacing the set: no analytic primitives gives the triangle-only results
    sc.set_analytic()
    assert_hits_equal(sc.trace_closest(d), oracle.closest_brute(tris, rays), "after clearing")
    sc.close()


def test_analytic_only_scene(gpu_ctx):
    empty = np.zeros((0, 9), np.float32)
    rays = camera_rays(64, 48)
    want = oracle.closest_brute(empty, rays, analytic=NIELS_ANALYTIC)
    assert (want[:, 0] != RT_MISS).mean() > 0.3
    sc = Scene(gpu_ctx, empty, width=4)
    sc.set_analytic(**NIELS_ANALYTIC)
    d = to_dev(rays)
    assert_hits_equal(sc.trace_closest(d), want, "analytic only")
    assert_hits_equal(sc.trace_closest_brute(d), want, "analytic only, linear scan")
    sc.close()


def cam_array(cam, jitter):
    return np.concatenate([np.float32(x) for x in cam] + [np.float32(jitter)])


@pytest.mark.parametrize("width", (2, 4))
def test_raygen_with_analytic_normals(gpu_ctx, width):
    import torch
    W, H = 80, 52
    cam = screen_camera((0.0, 5.0, -9.0), (0.0, 4.0, 0.0), 60.0, aspect=W / H)
    sc = Scene(gpu_ctx, NIELS_TRIS, width=width)
    sc.set_analytic(**NIELS_ANALYTIC)
    fill = np.float32(-7.0)
    uvn = torch.full((H, W, 4), float(fill), dtype=torch.float32, device="cuda")
    dob, uvn = sc.raygen_closest(cam, W, H, uv_normal=uvn, jitter=(0.3, 0.6))
    want_d, want_u = oracle.raygen(NIELS_TRIS, cam_array(cam, (0.3, 0.6)), W, H, fill=fill, analytic=NIELS_ANALYTIC)
    np.testing.assert_array_equal(dob.cpu().numpy().view(np.uint32), want_d.view(np.uint32))
    np.testing.assert_array_equal(uvn.cpu().numpy().view(np.uint32), want_u.view(np.uint32))
    ids = want_d[..., 3].view(np.uint32)
    assert (ids >= 3).sum() > 200 and ((ids != RT_MISS) & (ids >= 3)).sum() > 200
    sc.close()


def test_shadow_tiles_with_analytic(gpu_ctx):
    W, H, S = 40, 24, 2
    rays = camera_rays(W, H)
    prim = oracle.closest_brute(NIELS_TRIS, rays, analytic=NIELS_ANALYTIC)
    all_rays = np.zeros((S, H * W), RAY_DTYPE)
    for s, light in enumerate(((0.0, 12.0, 0.0), (-3.0, 8.0, -6.0))):
        hit = prim[:, 0] != RT_MISS
        t = prim[:, 1].view(np.float32)
        p = (rays["org"] + np.where(hit, t, 0)[:, None] * rays["dir"]).astype(np.float32)
        lv = (np.float32(light)[None, :] - p).astype(np.float32)
        dist = np.sqrt((lv.astype(np.float64) ** 2).sum(1)).astype(np.float32)
        r = all_rays[s]
        r["org"] = p
        r["dir"] = (lv / dist[:, None]).astype(np.float32)
        r["tmax"] = np.where(hit, dist, np.float32(0.0))
        r["prev_id"] = np.where(hit, prim[:, 0], RT_MISS)
    flat = all_rays.reshape(-1)
    want = oracle.occluded_brute(NIELS_TRIS, flat, analytic=NIELS_ANALYTIC).reshape(S, H, W)
    assert 0.02 < want.mean() < 0.98
    sc = Scene(gpu_ctx, NIELS_TRIS, width=4)
    sc.set_analytic(**NIELS_ANALYTIC)
    words = sc.shadow_tiles(to_dev(flat), W, H, S).cpu().numpy().view(np.uint64)
    tiles_x, tiles_y = (W + 15) // 16, (H + 3) // 4
    for s in range(S):
        for y in range(H):
            for x in range(W):
                idx = (x >> 4) + (y >> 2) * tiles_x + s * tiles_x * tiles_y
                assert (int(words[idx]) >> ((x & 15) | ((y & 3) << 4))) & 1 == want[s, y, x], (s, x, y)
    sc.close()


def _analytic_batch_queries(gpu_ctx):
    """Three scenes (the reference's demo scene, a 4,096-triangle scene with
    98 analytic objects, the same triangles alone) and a closest-hit and an
    occlusion query on each, with the oracle's answers."""
    tris = wl.c1_scene()
    ana = big_analytic(5)
    niels = Scene(gpu_ctx, NIELS_TRIS, width=4)
    niels.set_analytic(**NIELS_ANALYTIC)
    mixed = Scene(gpu_ctx, tris, width=4)
    mixed.set_analytic(**ana)
    plain = Scene(gpu_ctx, tris, width=4)
    ntot = tris.shape[0] + 98
    cam = camera_rays(64, 40)
    sh = shadow_rays(cam, oracle.closest_brute(NIELS_TRIS, cam, analytic=NIELS_ANALYTIC), (0.0, 12.0, 0.0))
    r = random_rays(6000, 21, lo=-15, hi=15, prev_max=ntot)
    o = random_rays(6000, 22, lo=-15, hi=15, tmax=(0, 20), prev_max=ntot)
    spec = [(0, "closest", cam, oracle.closest_brute(NIELS_TRIS, cam, analytic=NIELS_ANALYTIC)),
            (0, "occluded", sh, oracle.occluded_brute(NIELS_TRIS, sh, analytic=NIELS_ANALYTIC)),
            (1, "closest", r, oracle.closest_brute(tris, r, analytic=ana)),
            (1, "occluded", o, oracle.occluded_brute(tris, o, analytic=ana)),
            (2, "closest", r, oracle.closest_brute(tris, r)),
            (2, "occluded", o, oracle.occluded_brute(tris, o))]
    # the analytic objects change results (else the test would pin nothing)
    assert (spec[2][3][:, 0] != spec[4][3][:, 0]).sum() > 500
    assert (spec[3][3] != spec[5][3]).sum() > 200
    return [niels, mixed, plain], spec


def _outs(spec):
    import torch
    return [torch.empty((rays.shape[0], 4), dtype=torch.int32, device="cuda") if kind == "closest" else
            torch.empty(rays.shape[0], dtype=torch.uint8, device="cuda") for _, kind, rays, _ in spec]


def _check(spec, outs, what):
    for (si, kind, rays, want), out in zip(spec, outs):
        if kind == "closest":
            assert_hits_equal(out, want, f"{what}: scene {si} closest")
        else:
            np.testing.assert_array_equal(out.cpu().numpy(), want, err_msg=f"{what}: scene {si} occlusion")


def test_batch_with_analytic_scenes(gpu_ctx):
    """One batch launch over scenes with and without analytic objects, then
    the batch's analytic pass: every query equals the oracle's scan."""
    import igx_raytracing_amd as rt
    scenes, spec = _analytic_batch_queries(gpu_ctx)
    outs = _outs(spec)
    queries = [(si, kind, to_dev(rays), out) for (si, kind, rays, _), out in zip(spec, outs)]
    batch = rt.Batch(gpu_ctx, scenes)
    for opts in (None, trace_opts(lds_stack=8, chunk=64, refill=4)):
        for out in outs:
            out.fill_(-1)  # results must not depend on what the buffers held
        batch.trace(queries, opts=opts)
        _check(spec, outs, f"batch {opts}")
    # a scene that replaces its analytic set after the batch was made: the
    # batch's table holds the old buffers, so the trace is refused
    scenes[0].set_analytic(**NIELS_ANALYTIC)
    with pytest.raises(rt.RtError, match="changed after the batch"):
        batch.trace(queries)
    batch.close()
    batch = rt.Batch(gpu_ctx, scenes)
    batch.trace(queries)
    _check(spec, outs, "new batch")
    batch.close()
    assert gpu_ctx.check_device_error() == 0
    for sc in scenes:
        sc.close()


def test_server_with_analytic_scenes(gpu_ctx):
    """Requests against scenes with analytic objects, served in a mix over
    several streams, equal the oracle; a scene that replaces its set after
    registration fails its requests instead of being traced with freed
    buffers."""
    import torch
    from igx_raytracing_amd import Server
    from igx_raytracing_amd.serve import DONE, FAILED
    scenes, spec = _analytic_batch_queries(gpu_ctx)
    outs = _outs(spec)
    srv = Server(gpu_ctx, mem_budget=64 << 20, streams=4)
    ids = [srv.add_scene(sc) for sc in scenes]
    devs = [to_dev(rays) for _, _, rays, _ in spec]  # held until the requests are done
    tickets = [srv.submit(ids[si], kind, rays.shape[0], d, out, device=True)
               for (si, kind, rays, _), d, out in zip(spec, devs, outs)]
    for t in tickets:
        st, status, _ = srv.wait(t, 60_000)
        assert st == DONE and status == 0
    _check(spec, outs, "served")
    scenes[1].set_analytic(**big_analytic(6))
    si, kind, rays, _ = spec[2]
    t = srv.submit(ids[si], kind, rays.shape[0], devs[2], outs[2], device=True)
    st, status, _ = srv.wait(t, 60_000)
    assert st == FAILED and status != 0
    srv.close()
    torch.cuda.synchronize()
    assert gpu_ctx.check_device_error() == 0
    for sc in scenes:
        sc.close()


def test_pathtracer_rejects_analytic_scenes(gpu_ctx):
    import igx_raytracing_amd as rt
    sc = Scene(gpu_ctx, NIELS_TRIS, width=4)
    sc.set_analytic(**NIELS_ANALYTIC)
    cam = screen_camera((0.0, 4.0, -14.0), (0.0, 2.0, 0.0), 60.0)
    with pytest.raises(rt.RtError, match="triangle-only"):
        rt.PathTracer(sc, cam, 8, 8, 1)
    sc.close()


def test_set_analytic_is_stream_ordered(gpu_ctx):
    """The reference moves its spheres every frame (NielsScene::update,
    test/scene/niels_scene.cpp:61-71).  rt_scene_set_analytic replaces the
    set without a device-wide synchronisation: queries queued before the
    call, on another stream and still running when it returns, read the old
    set; queries after it read the new one."""
    import torch
    tris = wl.c1_scene()
    base = big_analytic(21)
    sets = []
    for f in range(7):
        a = {k: v.copy() for k, v in base.items()}
        a["spheres"][:, 0] += np.float32(0.75 * f)
        a["cubes"][:, [1, 4]] -= np.float32(0.5 * f)
        sets.append(a)
    sc = Scene(gpu_ctx, tris, width=4)
    long_rays = wl.c1_rays(1024, 1024)  # the linear scan over these takes milliseconds
    rays = wl.c1_rays(128, 128)
    d_long, d = to_dev(long_rays), to_dev(rays)
    sA, sB = torch.cuda.Stream(), torch.cuda.Stream()
    sc.set_analytic(**sets[0])
    torch.cuda.synchronize()
    outs, busy = [], []
    for f in range(6):
        with torch.cuda.stream(sA):
            big = sc.trace_closest_brute(d_long, stream=sA)
            a_out = sc.trace_closest(d, stream=sA)
        sc.set_analytic(**sets[f + 1])
        busy.append(not sA.query())
        with torch.cuda.stream(sB):
            b_out = sc.trace_closest(d, stream=sB)
        outs.append((big, a_out, b_out))
    torch.cuda.synchronize()
    print(f"set_analytic returned while the earlier stream still ran: {sum(busy)} of {len(busy)} frames",
          flush=True)
    assert any(busy), "set_analytic waited for work queued on another stream"
    sample = np.arange(0, long_rays.shape[0], 61)
    for f, (big, a_out, b_out) in enumerate(outs):
        assert_hits_equal(a_out, oracle.closest_brute(tris, rays, analytic=sets[f], nthreads=16), f"frame {f} before")
        assert_hits_equal(b_out, oracle.closest_brute(tris, rays, analytic=sets[f + 1], nthreads=16),
                          f"frame {f} after")
        assert_hits_equal(big.cpu().numpy()[sample],
                          oracle.closest_brute(tris, long_rays[sample], analytic=sets[f], nthreads=16),
                          f"frame {f} long scan")
    assert gpu_ctx.check_device_error() == 0
    sc.close()

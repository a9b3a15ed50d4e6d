"""Reference splitting in the host SAH builder (rt_build_opts.presplit):
large triangles are cut into parts with tight boxes before the build, and a
triangle's record appears in every leaf holding one of its parts.  Results
must not change: the oracle's traversal of split trees equals the
reference's linear scan; every point of every triangle lies in a leaf box
that holds its record (so no ray can miss it); each leaf holds a triangle
once; the splitting lowers the triangle tests on the scene it is for."""
import numpy as np
import pytest

import oracle
import igx_raytracing_amd as rt
from igx_raytracing_amd import RT_MISS
from igx_raytracing_amd import workload as wl

from test_lbvh import emitted_leaves


def spiky(subdiv=4):
    """A displaced icosphere: large random slivers in a thick shell."""
    return wl.icosphere(wl.seed(2), subdiv, 10.0, 0.5)


def scenes():
    dup = np.concatenate([wl.c1_scene()[:500]] * 3)
    return {"spiky": spiky(), "mixed": wl.c1_scene(), "dup": dup}


@pytest.mark.parametrize("name", ["spiky", "mixed", "dup"])
@pytest.mark.parametrize("width", (2, 4, 8))
def test_split_trees_equal_the_linear_scan(name, width):
    tris = scenes()[name]
    rays = wl.pinhole(64, 40, eye=(0.0, 4.0, -32.0), target=(0.0, 0.0, 0.0), vfov_deg=60.0) if name == "spiky" \
        else wl.c1_rays(64, 40)
    want = oracle.closest_brute(tris, rays)
    occ = rays.copy()
    occ["tmax"] = np.random.default_rng(1).uniform(0, 40, occ.shape[0]).astype(np.float32)
    want_o = oracle.occluded_brute(tris, occ)
    for ps in (2, 4, 16):
        h = rt.HostBvh(tris, width=width, presplit=ps)
        ids = h.tris[:, 3].view(np.uint32)
        assert h.tris.shape[0] >= tris.shape[0]
        assert np.array_equal(np.unique(ids), np.arange(tris.shape[0])), "every triangle keeps a record"
        got = oracle.walk(width, h.nodes, h.tris, h.info["root_ref"], rays)
        np.testing.assert_array_equal(got, want, err_msg=f"{name} BVH{width} presplit {ps}")
        np.testing.assert_array_equal(
            oracle.walk(width, h.nodes, h.tris, h.info["root_ref"], occ, occlusion=True), want_o)
        # a leaf holds each triangle at most once
        for first, count in emitted_leaves(h.nodes, width, h.info["root_ref"]):
            leaf_ids = ids[first:first + count]
            assert np.unique(leaf_ids).size == count


def test_splitting_lowers_triangle_tests_on_slivers():
    tris = spiky(6)
    rays = wl.pinhole(64, 40, eye=(0.0, 4.0, -32.0), target=(0.0, 0.0, 0.0), vfov_deg=30.0)
    tests = {}
    for ps in (1, 16):
        h = rt.HostBvh(tris, width=4, presplit=ps)
        _, cnt = oracle.walk(4, h.nodes, h.tris, h.info["root_ref"], rays, counts=True)
        tests[ps] = cnt[:, 1].mean()
    assert tests[16] < 0.8 * tests[1], tests
    # the budget bounds the records: at most presplit per triangle
    for ps in (2, 4):
        h = rt.HostBvh(tris, width=4, presplit=ps)
        ids = h.tris[:, 3].view(np.uint32)
        assert np.bincount(ids, minlength=tris.shape[0]).max() <= ps


def test_parts_cover_their_triangles():
    """BVH2 nodes hold float child boxes: every sampled point of every
    triangle lies in the box of a leaf that holds the triangle's record."""
    tris = spiky(3)
    h = rt.HostBvh(tris, width=2, presplit=16)
    assert h.tris.shape[0] > 1.3 * tris.shape[0], "the scene is split"
    nodes = h.nodes.reshape(-1, 64)
    boxes = nodes[:, :48].copy().view(np.float32).reshape(-1, 2, 6)  # lo0 hi0 | lo1 hi1
    refs = nodes[:, 48:56].copy().view(np.uint32).reshape(-1, 2)
    ids = h.tris[:, 3].view(np.uint32)
    boxes_of = {}
    for i in range(refs.shape[0]):
        for c in range(2):
            r = int(refs[i, c])
            if r & 0x80000000:
                first, count = r & 0x0FFFFFFF, ((r >> 28) & 7) + 1
                for t in ids[first:first + count]:
                    boxes_of.setdefault(int(t), []).append(boxes[i, c])
    assert len(boxes_of) == tris.shape[0]
    rng = np.random.default_rng(7)
    for t in range(tris.shape[0]):
        b = rng.random((64, 2))
        b[b.sum(1) > 1] = 1 - b[b.sum(1) > 1]
        v = tris[t].reshape(3, 3).astype(np.float64)
        pts = v[0] + b[:, :1] * (v[1] - v[0]) + b[:, 1:] * (v[2] - v[0])
        bx = np.array(boxes_of[t], np.float64)
        inside = ((pts[:, None, :] >= bx[None, :, :3]) & (pts[:, None, :] <= bx[None, :, 3:])).all(2).any(1)
        assert inside.all(), f"triangle {t}: a point lies in none of its leaf boxes"


def test_presplit_option_checks():
    tris = wl.c1_scene()[:100]
    with pytest.raises(rt.RtError, match="presplit"):
        rt.HostBvh(tris, width=4, presplit=65)
    h = rt.HostBvh(tris, width=4, presplit=0)  # 0 means no splitting
    assert h.tris.shape[0] == 100


@pytest.mark.gpu
@pytest.mark.parametrize("width", (2, 4, 8))
def test_gpu_split_scene_queries(gpu_ctx, width):
    """Every query entry point on a split scene equals the linear scan; the
    kernels' node and triangle counts equal the oracle's walk of the same
    nodes; analytic objects keep their ids after the input triangles."""
    import torch
    from helpers import assert_hits_equal, to_dev
    tris = spiky(6)
    rays = wl.pinhole(96, 64, eye=(0.0, 4.0, -32.0), target=(0.0, 0.0, 0.0), vfov_deg=30.0)
    occ = rays.copy()
    occ["tmax"] = np.random.default_rng(2).uniform(0, 40, occ.shape[0]).astype(np.float32)
    want = oracle.closest_brute(tris, rays)
    want_o = oracle.occluded_brute(tris, occ)
    sc = rt.Scene(gpu_ctx, tris, width=width, presplit=16)
    assert sc.info["tris_total_bytes"] // 48 > tris.shape[0]
    assert sc.info["num_tris"] == tris.shape[0]
    d, do = to_dev(rays), to_dev(occ)
    for opts in (None, rt.trace_opts(schedule="static"), rt.trace_opts(lds_stack=1)):
        assert_hits_equal(sc.trace_closest(d, opts=opts), want, f"split BVH{width} {opts}")
        np.testing.assert_array_equal(sc.trace_occluded(do, opts=opts).cpu().numpy(), want_o)
    assert_hits_equal(sc.trace_closest_brute(d), want, "linear scan")
    nodes, recs = sc.download()
    counts = torch.empty((rays.shape[0], 2), dtype=torch.int32, device="cuda")
    got = sc.trace_closest(d, counts=counts)
    wwalk, wcnt = oracle.walk(width, nodes, recs, sc.info["root_ref"], rays, counts=True)
    assert_hits_equal(got, wwalk, "walk")
    np.testing.assert_array_equal(counts.cpu().numpy().view(np.uint32), wcnt)
    ana = wl.analytic_mix(5)
    sc.set_analytic(**ana)
    assert_hits_equal(sc.trace_closest(d), oracle.closest_brute(tris, rays, analytic=ana), "split + analytic")
    assert gpu_ctx.check_device_error() == 0
    sc.close()


@pytest.mark.gpu
def test_gpu_path_tracing_same_image_with_split_scene(gpu_ctx):
    """Splitting changes only which leaves hold a triangle: the path
    tracer's image and ray counts are identical."""
    import torch
    tris = spiky(6)
    cam = rt.screen_camera((0.0, 4.0, -32.0), (0.0, 0.0, 0.0), 40.0, aspect=64 / 48)
    out = []
    for ps in (1, 16):
        sc = rt.Scene(gpu_ctx, tris, width=4, presplit=ps)
        pt = rt.PathTracer(sc, cam, 64, 48, 2, max_depth=3)
        img = torch.empty((48, 64), dtype=torch.float32, device="cuda")
        pt.render(3, image=img, use_graph=False)
        torch.cuda.synchronize()
        out.append((img.cpu().numpy().copy(), pt.counts().copy()))
        pt.close()
        sc.close()
    assert out[0][1][1] > 100, "some rays bounce"
    np.testing.assert_array_equal(out[0][0].view(np.uint32), out[1][0].view(np.uint32))
    np.testing.assert_array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("seed", range(6))
def test_random_soups_split_equal_the_scan(seed):
    """Random soups mixing tiny, huge and thin triangles with exact
    duplicates: the oracle's walk of every split tree equals the linear scan,
    closest hit and occlusion.  With points and segments added (whose
    Moller-Trumbore test yields NaN t, the one documented divergence between
    the scan and any tree), split trees still equal the unsplit tree."""
    rng = np.random.default_rng(100 + seed)
    n = 300
    c = rng.uniform(-5, 5, (n, 1, 3))
    scale = np.exp(rng.uniform(np.log(0.01), np.log(8.0), (n, 1, 1)))
    tris = (c + scale * rng.normal(size=(n, 3, 3))).astype(np.float32)
    tris[:20, 2] = tris[:20, 0] + np.float32(1e-3) * (tris[:20, 1] - tris[:20, 0])  # slivers
    tris[20:30] = tris[30:40]                                                      # duplicates
    tris = tris.reshape(n, 9)
    rays = np.zeros(2000, rt.RAY_DTYPE)
    rays["org"] = rng.uniform(-12, 12, (2000, 3))
    d = rng.normal(size=(2000, 3))
    rays["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays["tmax"] = np.float32(3.4028235e38)
    rays["prev_id"] = np.where(rng.random(2000) < 0.2, rng.integers(0, n, 2000), RT_MISS)
    occ = rays.copy()
    occ["tmax"] = rng.uniform(0, 20, 2000).astype(np.float32)
    want = oracle.closest_brute(tris, rays)
    want_o = oracle.occluded_brute(tris, occ)
    assert (want[:, 0] != RT_MISS).sum() > 100
    degenerate = tris.copy().reshape(n, 3, 3)
    degenerate[40:50, 1] = degenerate[40:50, 0]   # segments
    degenerate[50:55, 1:] = degenerate[50:55, :1]  # points
    degenerate = degenerate.reshape(n, 9)
    for width in (2, 4, 8):
        base = rt.HostBvh(degenerate, width=width)
        want_d = oracle.walk(width, base.nodes, base.tris, base.info["root_ref"], rays)
        for ps in (2, 8, 64):
            ml = int(rng.integers(1, 9))
            h = rt.HostBvh(tris, width=width, presplit=ps, max_leaf=ml)
            got = oracle.walk(width, h.nodes, h.tris, h.info["root_ref"], rays)
            np.testing.assert_array_equal(got, want, err_msg=f"seed {seed} BVH{width} presplit {ps}")
            np.testing.assert_array_equal(oracle.walk(width, h.nodes, h.tris, h.info["root_ref"], occ,
                                                      occlusion=True), want_o)
            hd = rt.HostBvh(degenerate, width=width, presplit=ps, max_leaf=ml)
            np.testing.assert_array_equal(oracle.walk(width, hd.nodes, hd.tris, hd.info["root_ref"], rays), want_d,
                                          err_msg=f"degenerate seed {seed} BVH{width} presplit {ps}")


@pytest.mark.parametrize("width", (2, 4, 8))
def test_infinite_coordinates_are_refused_and_nan_ones_built(width):
    """A box with an infinite side has no quantised form: the build refuses
    the scene (it once looped on the exponent).  NaN coordinates build, and
    one such triangle leaves splitting on for the others (the size
    threshold is a mean over finite boxes)."""
    base = spiky(3)
    for bad in (np.inf, -np.inf):
        tris = base.copy()
        tris[9, 4] = bad
        with pytest.raises(rt.RtError, match="infinite coordinate"):
            rt.HostBvh(tris, width=width, presplit=16)
    tris = base.copy()
    tris[5, 0] = np.nan
    split = rt.HostBvh(tris, width=width, presplit=16)
    assert split.tris.shape[0] > 1.3 * tris.shape[0], "splitting stays on"
    clean = rt.HostBvh(base, width=width, presplit=16)
    assert split.tris.shape[0] >= clean.tris.shape[0] - 16

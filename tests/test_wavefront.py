"""Wavefront path tracing: the image and the rays traced per depth equal the
oracle's path-by-path restatement bit for bit, eagerly and by hipGraph
replay, for any queue order (replays interleaved with other work)."""
import numpy as np
import pytest

import oracle
from igx_raytracing_amd import PathTracer, Scene, screen_camera
from igx_raytracing_amd import workload as wl

pytestmark = pytest.mark.gpu


def small_scene():
    return wl.c3_scene(subdiv=3, cluster_tris=512)


@pytest.mark.parametrize("width", (2, 4, 8))
def test_pathtrace_matches_oracle(gpu_ctx, width):
    import torch
    tris = small_scene()
    W, H, spp = 48, 32, 2
    cam = screen_camera((0.0, 3.0, -30.0), (0.0, 0.0, 0.0), 50.0, aspect=W / H)
    cam_arr = np.concatenate([np.float32(x) for x in cam] + [np.float32([0.5, 0.5])])
    sc = Scene(gpu_ctx, tris, width=width)
    pt = PathTracer(sc, cam, W, H, spp)
    for seed in (0, 12345):
        want, wcounts = oracle.pathtrace(tris, cam_arr, W, H, spp, seed=seed, nthreads=16)
        assert wcounts[0] == W * H * spp and wcounts[1] > 0 and wcounts[3] > 0
        eager = pt.render(seed, use_graph=False).cpu().numpy()
        np.testing.assert_array_equal(eager.view(np.uint32), want.view(np.uint32))
        np.testing.assert_array_equal(pt.counts(), wcounts)
        img = torch.empty((H, W), dtype=torch.float32, device="cuda")
        for rep in range(3):
            pt.render(seed, image=img, use_graph=True)
            got = img.cpu().numpy()  # synchronizes; eager work in between replays
            np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32), err_msg=f"graph replay {rep}")
            np.testing.assert_array_equal(pt.counts(), wcounts)
            pt.render(seed + 1, use_graph=False)
    assert gpu_ctx.check_device_error() == 0
    pt.close()
    sc.close()


@pytest.mark.parametrize("chunk,refill", [(16, 1), (64, 16), (256, 64), (4096, 8)])
def test_pathtrace_feed_shape_never_changes_the_image(gpu_ctx, chunk, refill):
    """The bounce feed's chunk and refill (rt_pt_opts) move work between
    waves only: every shape renders the oracle's image and counts."""
    tris = small_scene()
    W, H, spp = 48, 32, 2
    cam = screen_camera((0.0, 3.0, -30.0), (0.0, 0.0, 0.0), 50.0, aspect=W / H)
    cam_arr = np.concatenate([np.float32(x) for x in cam] + [np.float32([0.5, 0.5])])
    want, wcounts = oracle.pathtrace(tris, cam_arr, W, H, spp, seed=7, nthreads=16)
    sc = Scene(gpu_ctx, tris, width=4)
    pt = PathTracer(sc, cam, W, H, spp, chunk=chunk, refill=refill)
    for graph in (False, True):
        got = pt.render(7, use_graph=graph).cpu().numpy()
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32), err_msg=f"graph={graph}")
        np.testing.assert_array_equal(pt.counts(), wcounts)
    assert gpu_ctx.check_device_error() == 0
    pt.close()
    sc.close()


def test_pathtrace_feed_shape_is_checked(gpu_ctx):
    from igx_raytracing_amd import RtError
    sc = Scene(gpu_ctx, small_scene(), width=4)
    cam = screen_camera((0.0, 3.0, -30.0), (0.0, 0.0, 0.0), 50.0, aspect=1.5)
    for chunk, refill in ((8, 8), (8192, 8), (64, 65)):
        with pytest.raises(RtError, match="chunk"):
            PathTracer(sc, cam, 48, 32, 1, chunk=chunk, refill=refill)
    sc.close()


def test_frames_queue_without_host_synchronisation(gpu_ctx):
    """rt_pt_render stages each frame's seed through a pinned ring instead of
    waiting for the stream: frames with different seeds queue back to back
    (graph replays and eager frames mixed), each renders the oracle's image,
    and the call returns while its frame is still on the device."""
    import torch
    tris = small_scene()
    W, H, spp = 160, 120, 4
    cam = screen_camera((0.0, 3.0, -30.0), (0.0, 0.0, 0.0), 50.0, aspect=W / H)
    cam_arr = np.concatenate([np.float32(x) for x in cam] + [np.float32([0.5, 0.5])])
    sc = Scene(gpu_ctx, tris, width=4)
    pt = PathTracer(sc, cam, W, H, spp)
    stream = torch.cuda.Stream()
    seeds = [3, 11, 12, 40, 41, 97, 5, 6, 7, 8, 9, 10]  # more frames than ring slots
    images = [torch.empty((H, W), dtype=torch.float32, device="cuda") for _ in seeds]
    gimg = torch.empty((H, W), dtype=torch.float32, device="cuda")
    pt.render(1, image=gimg, use_graph=True, stream=stream)  # capture once, on gimg
    torch.cuda.synchronize()
    busy = []
    with torch.cuda.stream(stream):
        for k, seed in enumerate(seeds):
            if k % 2:  # graph replay into gimg, copied out in stream order
                pt.render(seed, image=gimg, use_graph=True, stream=stream)
                busy.append(not stream.query())
                images[k].copy_(gimg)
            else:
                pt.render(seed, image=images[k], use_graph=False, stream=stream)
                busy.append(not stream.query())
    torch.cuda.synchronize()
    print(f"render returned before its frame finished: {sum(busy)} of {len(busy)}", flush=True)
    assert any(busy), "rt_pt_render waited for its frame"
    for k, seed in enumerate(seeds[:6]):
        want, _ = oracle.pathtrace(tris, cam_arr, W, H, spp, seed=seed, nthreads=16)
        np.testing.assert_array_equal(images[k].cpu().numpy().view(np.uint32), want.view(np.uint32),
                                      err_msg=f"frame {k} seed {seed}")
    assert gpu_ctx.check_device_error() == 0
    pt.close()
    sc.close()


def test_queue_rays_follow_the_plain_query(gpu_ctx):
    """rt_pt_queue_rays returns the rays a frame traced at its last two depths.
    Every camera ray that hits continues (no roulette before the second
    bounce) from o + t*d with the hit triangle as prevHit, so the multiset of
    (prevHit, origin) at depth 1 is what the plain closest-hit query over the
    depth-0 rays predicts.  A tracer with a deeper max_depth traces the same
    depth-1 rays for the same frame seed."""
    tris = small_scene()
    W, H, spp = 48, 32, 2
    cam = screen_camera((0.0, 3.0, -30.0), (0.0, 0.0, 0.0), 50.0, aspect=W / H)
    sc = Scene(gpu_ctx, tris, width=4)
    p1 = PathTracer(sc, cam, W, H, spp, max_depth=1)
    p2 = PathTracer(sc, cam, W, H, spp, max_depth=2)
    p1.render(5, use_graph=False)
    p2.render(5, use_graph=False)
    r0, r1 = p1.queue_rays(0), p1.queue_rays(1)
    r1b, r2 = p2.queue_rays(1), p2.queue_rays(2)
    assert r0.shape[0] == W * H * spp
    assert r2.shape[0] == p2.counts()[2] > 0
    h = sc.trace_closest(r0).cpu().numpy().view(np.uint32)
    v0 = r0.cpu().numpy()
    hit = h[:, 0] != 0xFFFFFFFF
    t = h[hit, 1].view(np.float32)
    org = v0[hit, 0:3] + t[:, None] * v0[hit, 4:7]  # float32, one rounding per operation as in the shader
    want = np.concatenate([h[hit, 0:1], org.view(np.uint32)], axis=1)
    v1 = r1.cpu().numpy().view(np.uint32)
    got = np.concatenate([v1[:, 7:8], v1[:, 0:3]], axis=1)

    def rows(a):
        return a[np.lexsort(a.T[::-1])]

    assert got.shape == want.shape
    np.testing.assert_array_equal(rows(got), rows(want))
    np.testing.assert_array_equal(rows(r1b.cpu().numpy().view(np.uint32)), rows(v1))
    with pytest.raises(Exception):
        p2.queue_rays(0)  # overwritten by depth 2
    with pytest.raises(Exception):
        p1.queue_rays(2)
    assert gpu_ctx.check_device_error() == 0
    p1.close()
    p2.close()
    sc.close()

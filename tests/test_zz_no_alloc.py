"""rt_trace_closest_noalloc against rt_trace_closest, through the C ABI.

The allocation-free entry point keeps the traversal's spill stack and work
counter in the library's fixed per-device scratch; its hits must equal the
allocating entry point's bit for bit, call after call, whatever the size of
the previous call.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(lib, fn, ctx, sc, rays_d, n, torch):
    hits = torch.full((n, 4), -1, dtype=torch.int32, device="cuda")
    rc = fn(ctx.handle, sc.handle, ctypes.c_void_p(rays_d.data_ptr()), n, ctypes.c_void_p(hits.data_ptr()), None)
    assert rc == 0, lib.rt_last_error(ctx.handle)
    torch.cuda.synchronize()
    return hits.cpu().numpy()


def test_noalloc_equals_allocating_entry_point_twice():
    import torch
    import igx_raytracing_amd as rt
    from igx_raytracing_amd import _lib
    from igx_raytracing_amd import workload as wl
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    lib = _lib.load()
    tris = wl.c2_scene()
    rays = wl.c2_rays(tris, 1 << 20)
    with rt.Context(0) as ctx:
        for width in (4, 8):
            sc = rt.Scene(ctx, tris, width=width)
            rays_d = rt.rays_to_device(rays)
            print(f"noalloc: width {width}, calls of {rays.shape[0]} then 65536 rays", flush=True)
            # Two calls in a row, the second smaller: the scratch keeps the
            # first call's contents, and the second must not see them.
            for n in (rays.shape[0], 65536):
                want = _run(lib, lib.rt_trace_closest, ctx, sc, rays_d, n, torch)
                got = _run(lib, lib.rt_trace_closest_noalloc, ctx, sc, rays_d, n, torch)
                assert np.array_equal(got, want), f"width {width}, n {n}: {int((got != want).any(1).sum())} rays differ"
                assert (want[:, 0] != -1).any()
            assert ctx.check_device_error() == 0
            sc.close()


def test_noalloc_refuses_launches_beyond_its_limits():
    import torch
    import igx_raytracing_amd as rt
    from igx_raytracing_amd import _lib
    from igx_raytracing_amd import workload as wl
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    lib = _lib.load()
    tris = wl.c1_scene()
    with rt.Context(0) as ctx:
        sc = rt.Scene(ctx, tris, width=4)
        rays_d = rt.rays_to_device(wl.c1_rays(8, 8))
        hits = torch.empty((64, 4), dtype=torch.int32, device="cuda")
        # More rays than one launch takes: refused on the host, before any launch.
        rc = lib.rt_trace_closest_noalloc(ctx.handle, sc.handle, ctypes.c_void_p(rays_d.data_ptr()), 0x80000000,
                                          ctypes.c_void_p(hits.data_ptr()), None)
        assert rc == -1  # RT_ERR_INVALID
        rc = lib.rt_trace_closest_noalloc(ctx.handle, sc.handle, None, 64, ctypes.c_void_p(hits.data_ptr()), None)
        assert rc == -1  # RT_ERR_INVALID
        # Zero rays: nothing to do.
        assert lib.rt_trace_closest_noalloc(ctx.handle, sc.handle, None, 0, None, None) == 0
        sc.close()

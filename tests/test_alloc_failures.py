"""Allocation failures: every device or pinned allocation of the library,
made to fail in turn (RT_FAIL_DEVICE_ALLOC=k fails the k-th one), makes the
call that needed it return RT_ERR_OOM (or, for a served request, finish
FAILED with RT_ERR_OOM) without crashing, leaking a half-made handle or
poisoning later calls: the same call made again afterwards succeeds, and
query results are then the oracle's.

The reference's only failure handling is log()->fatal
(src/rt/raytracing_interface.cpp:75-76); these are the C ABI's own error
paths.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
from igx_raytracing_amd import workload as wl

from helpers import hits_u32

pytestmark = pytest.mark.gpu

RT_ERR_OOM = -2
ENV = "RT_FAIL_DEVICE_ALLOC"


class Sequence:
    """A fixed sequence of entry points through the raw C ABI, each step
    recording its return code; a failed step's dependants are skipped."""

    def __init__(self, lib, data):
        self.lib, self.d = lib, data
        self.rc = {}
        self.h = {}

    def step(self, name, fn):
        try:
            self.rc[name] = fn()
        except _Skip:
            self.rc[name] = None

    def need(self, *names):
        for n in names:
            if self.rc.get(n) != 0:
                raise _Skip()

    def run(self):
        import torch
        lib, d, h = self.lib, self.d, self.h
        from igx_raytracing_amd._lib import RtBuildOpts, RtPtOpts, RtQuery, RtRequest, RtServerOpts
        from igx_raytracing_amd.rt import make_camera
        vp = ctypes.c_void_p
        h["ctx"] = vp()
        self.step("ctx", lambda: lib.rt_ctx_create(0, ctypes.byref(h["ctx"])))

        def scene():
            self.need("ctx")
            o = RtBuildOpts()
            lib.rt_build_opts_default(ctypes.byref(o))
            h["scene"] = vp()
            return lib.rt_scene_create(h["ctx"], d["tris"].ctypes.data, d["tris"].shape[0], ctypes.byref(o),
                                       ctypes.byref(h["scene"]))
        self.step("scene", scene)

        def closest():
            self.need("scene")
            d["hits"].fill_(-1)
            rc = lib.rt_trace_closest(h["ctx"], h["scene"], vp(d["rays"].data_ptr()), d["n"],
                                      vp(d["hits"].data_ptr()), None)
            torch.cuda.synchronize()
            return rc
        self.step("closest", closest)

        def occluded():
            self.need("scene")
            rc = lib.rt_trace_occluded(h["ctx"], h["scene"], vp(d["orays"].data_ptr()), d["n"],
                                       vp(d["flags"].data_ptr()), None)
            torch.cuda.synchronize()
            return rc
        self.step("occluded", occluded)

        def lbvh():
            self.need("ctx")
            o = RtBuildOpts()
            lib.rt_build_opts_default(ctypes.byref(o))
            o.builder = 1
            h["lscene"] = vp()
            return lib.rt_scene_create(h["ctx"], d["tris"].ctypes.data, d["tris"].shape[0], ctypes.byref(o),
                                       ctypes.byref(h["lscene"]))
        self.step("lbvh", lbvh)

        def batch():
            self.need("scene")
            arr = (vp * 1)(h["scene"])
            h["batch"] = vp()
            return lib.rt_batch_create(h["ctx"], arr, 1, ctypes.byref(h["batch"]))
        self.step("batch", batch)

        def batch_trace():
            self.need("batch")
            q = (RtQuery * 1)()
            q[0].scene, q[0].kind, q[0].n = 0, 0, d["n"]
            q[0].rays, q[0].out = d["rays"].data_ptr(), d["bhits"].data_ptr()
            d["bhits"].fill_(-1)
            rc = lib.rt_batch_trace(h["ctx"], h["batch"], q, 1, None, None)
            torch.cuda.synchronize()
            return rc
        self.step("batch_trace", batch_trace)

        def analytic():
            self.need("scene")
            sp = d["spheres"]
            return lib.rt_scene_set_analytic(h["ctx"], h["scene"], sp.ctypes.data, sp.shape[0], None, 0, None, 0)
        self.step("analytic", analytic)

        def server():
            self.need("lbvh")
            o = RtServerOpts()
            lib.rt_server_opts_default(ctypes.byref(o))
            o.mem_budget = 16 << 20
            o.streams = 2
            h["srv"] = vp()
            return lib.rt_server_create(h["ctx"], ctypes.byref(o), ctypes.byref(h["srv"]))
        self.step("server", server)

        def serve():
            self.need("server")
            sid = ctypes.c_uint32()
            rc = lib.rt_server_add_scene(h["srv"], h["lscene"], ctypes.byref(sid))
            if rc:
                return rc
            r = RtRequest()
            r.scene, r.kind, r.n, r.memory = sid.value, 0, d["n"], 1
            r.rays, r.out = d["rays"].data_ptr(), d["shits"].data_ptr()
            d["shits"].fill_(-1)
            t = ctypes.c_uint64()
            rc = lib.rt_server_submit(h["srv"], ctypes.byref(r), ctypes.byref(t))
            if rc:
                return rc
            st, status, lat = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
            rc = lib.rt_server_wait(h["srv"], t.value, 60_000.0, ctypes.byref(st), ctypes.byref(status),
                                    ctypes.byref(lat))
            return rc if rc else status.value
        self.step("serve", serve)

        def pathtracer():
            self.need("lbvh")
            o = RtPtOpts()
            lib.rt_pt_opts_default(ctypes.byref(o))
            o.width, o.height, o.spp = 32, 16, 1
            cam = make_camera((0.0, 5.0, -25.0), (-1.0, 4.0, -24.0), (1.0, 4.0, -24.0), (-1.0, 6.0, -24.0))
            h["pt"] = vp()
            rc = lib.rt_pt_create(h["ctx"], h["lscene"], ctypes.byref(cam), ctypes.byref(o), ctypes.byref(h["pt"]))
            if rc:
                return rc
            rc = lib.rt_pt_render(h["pt"], 1, vp(d["img"].data_ptr()), 0, None)
            torch.cuda.synchronize()
            return rc
        self.step("pathtracer", pathtracer)

    def close(self):
        lib, h = self.lib, self.h
        for key, fn in (("pt", lib.rt_pt_destroy), ("srv", lib.rt_server_destroy), ("batch", lib.rt_batch_destroy),
                        ("lscene", lib.rt_scene_destroy), ("scene", lib.rt_scene_destroy),
                        ("ctx", lib.rt_ctx_destroy)):
            if h.get(key):
                fn(h[key])


class _Skip(Exception):
    pass


def test_every_allocation_failure_is_reported_and_recoverable():
    import torch
    from igx_raytracing_amd import _lib
    from igx_raytracing_amd import rays_to_device
    lib = _lib.load()
    tris = wl.c1_scene()
    rays = wl.c1_rays(64, 64)
    orays = rays.copy()
    orays["tmax"] = np.float32(20.0)
    n = rays.shape[0]
    data = {"tris": tris, "n": n, "rays": rays_to_device(rays), "orays": rays_to_device(orays),
            "hits": torch.empty((n, 4), dtype=torch.int32, device="cuda"),
            "bhits": torch.empty((n, 4), dtype=torch.int32, device="cuda"),
            "shits": torch.empty((n, 4), dtype=torch.int32, device="cuda"),
            "flags": torch.empty(n, dtype=torch.uint8, device="cuda"),
            "img": torch.empty((16, 32), dtype=torch.float32, device="cuda"),
            "spheres": np.float32([[0, 2, 0, 1.5], [4, 1, 3, 1.0]])}
    want = oracle.closest_brute(tris, rays, nthreads=16)
    want_occ = oracle.occluded_brute(tris, orays, nthreads=16)
    os.environ.pop(ENV, None)
    clean = Sequence(lib, data)
    clean.run()
    bad = {k: v for k, v in clean.rc.items() if v != 0}
    assert not bad, f"steps failed without an injected failure: {bad}: {lib.rt_last_error_global()}"
    clean.close()
    failed_steps = []
    try:
        for k in range(1, 200):
            os.environ[ENV] = str(k)
            seq = Sequence(lib, data)
            seq.run()
            os.environ.pop(ENV, None)
            bad = {name: rc for name, rc in seq.rc.items() if rc not in (0, None)}
            if not bad:
                seq.close()
                print(f"{k - 1} allocations; failures reported at {failed_steps}", flush=True)
                break
            assert list(bad.values()) == [RT_ERR_OOM], (k, seq.rc, lib.rt_last_error_global())
            name = next(iter(bad))
            failed_steps.append(name)
            # Steps after the failure ran normally: queries that ran give the
            # oracle's answers.
            if seq.rc.get("closest") == 0:
                np.testing.assert_array_equal(hits_u32(data["hits"]), want)
            if seq.rc.get("occluded") == 0:
                np.testing.assert_array_equal(data["flags"].cpu().numpy(), want_occ)
            if seq.rc.get("batch_trace") == 0:
                np.testing.assert_array_equal(hits_u32(data["bhits"]), want)
            seq.close()
        else:
            pytest.fail("the sequence never ran without an injected failure")
    finally:
        os.environ.pop(ENV, None)
    # every kind of call that allocates was made to fail at least once
    for name in ("ctx", "scene", "closest", "lbvh", "batch", "batch_trace", "analytic", "server", "serve",
                 "pathtracer"):
        assert name in failed_steps, (name, failed_steps)
    # and the library still works normally afterwards
    again = Sequence(lib, data)
    again.run()
    bad = {k: v for k, v in again.rc.items() if v != 0}
    assert not bad, f"steps failed after the injections: {bad}"
    np.testing.assert_array_equal(hits_u32(data["hits"]), want)
    again.close()

"""Host mirror of the serving API (rt_server_* in include/rt/rt_c.h)."""
import ctypes

from ._lib import RtRequest, RtServerOpts, RtServerStats, check, load, ptr

PENDING, RUNNING, DONE, FAILED = 0, 1, 2, 3
KINDS = {"closest": 0, "occluded": 1}


class Server:
    """Continuous-batching ray-query server over resident scenes."""

    def __init__(self, ctx=None, mem_budget=0, streams=4, max_tick_rays=8 << 20, plain=False, node_width=4,
                 max_scenes=1024, simulate=False, launch=None):
        lib = load()
        o = RtServerOpts()
        lib.rt_server_opts_default(ctypes.byref(o))
        o.mem_budget = mem_budget
        o.streams = streams
        o.max_tick_rays = max_tick_rays
        o.plain = 1 if plain else 0
        o.node_width = node_width
        o.max_scenes = max_scenes
        # simulate: True models the device's service time; "fail" also makes
        # every tick complete with a HIP error (failure-path tests).
        o.simulate = 2 if simulate == "fail" else (1 if simulate else 0)
        if launch is not None:
            o.launch = launch
        self.ctx = ctx
        self._ctxh = ctx.handle if ctx is not None else None
        h = ctypes.c_void_p()
        check(lib.rt_server_create(self._ctxh, ctypes.byref(o), ctypes.byref(h)), self._ctxh)
        self.handle = h

    def add_scene(self, scene=None):
        i = ctypes.c_uint32()
        check(load().rt_server_add_scene(self.handle, scene.handle if scene is not None else None, ctypes.byref(i)),
              self._ctxh)
        return i.value

    def upload_scene(self, tris):
        import numpy as np
        t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        i = ctypes.c_uint32()
        check(load().rt_server_upload_scene(self.handle, ptr(t), t.shape[0], ctypes.byref(i)), self._ctxh)
        return i.value

    def submit(self, scene, kind, n, rays=None, out=None, device=False, deadline_ms=0.0, stream=None):
        """Queues a request.  Device buffers are handed over from `stream`
        (default: torch's current stream): the tick waits for the work queued
        there at submit, so a buffer the caching allocator recycled from a
        still-pending kernel is written only after that kernel."""
        r = RtRequest()
        r.scene = scene
        r.kind = KINDS[kind]
        r.n = n
        r.memory = 1 if device else 0
        r.rays = ptr(rays)
        r.out = ptr(out)
        r.deadline_ms = deadline_ms
        if device and self.ctx is not None:
            from .rt import _stream_handle
            r.stream = _stream_handle(stream)
        t = ctypes.c_uint64()
        check(load().rt_server_submit(self.handle, ctypes.byref(r), ctypes.byref(t)), self._ctxh)
        return t.value

    def cancel(self, ticket):
        check(load().rt_server_cancel(self.handle, ticket), self._ctxh)

    def poll(self, ticket):
        st, status, lat = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(load().rt_server_poll(self.handle, ticket, ctypes.byref(st), ctypes.byref(status), ctypes.byref(lat)),
              self._ctxh)
        return st.value, status.value, lat.value

    def wait(self, ticket, timeout_ms=0.0):
        st, status, lat = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(load().rt_server_wait(self.handle, ticket, timeout_ms, ctypes.byref(st), ctypes.byref(status),
                                    ctypes.byref(lat)), self._ctxh)
        return st.value, status.value, lat.value

    def forget(self, ticket):
        check(load().rt_server_forget(self.handle, ticket), self._ctxh)

    def stats(self):
        s = RtServerStats()
        check(load().rt_server_stats_get(self.handle, ctypes.byref(s)), self._ctxh)
        return s.as_dict()

    def close(self):
        if self.handle:
            load().rt_server_destroy(self.handle)
            self.handle = None

"""Host mirror of the ray-query C ABI (include/rt/rt_c.h).

``Context`` owns a device context, ``Scene`` a device-resident BVH and its
triangle records; queries take device buffers (torch tensors on the GPU) and
a HIP stream.  PyTorch only provides device memory and streams here.

Query semantics are the reference's: traceGeometry / traceOcclusion
(res/shaders/trace.glsl:9-98) over rayIntersectTri
(res/shaders/primitive.glsl:239-284); see include/rt/rt_c.h.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (HIT_DTYPE, RAY_DTYPE, RtBuildOpts, RtCamera, RtPtOpts, RtQuery, RtSceneInfo, RtTraceOpts, check,
                   load, primary_camera, ptr)  # noqa: F401

BUILDERS = {"sah": 0, "lbvh": 1}


def build_opts(width=4, builder="sah", max_leaf=4, sah_bins=32, pad_rel=None, keep_source=True, presplit=1):
    o = RtBuildOpts()
    load().rt_build_opts_default(ctypes.byref(o))
    o.builder = BUILDERS[builder]
    o.node_width = width
    o.max_leaf = max_leaf
    o.sah_bins = sah_bins
    if pad_rel is not None:
        o.pad_rel = pad_rel
    o.keep_source = 1 if keep_source else 0
    o.presplit = presplit
    return o


def trace_opts(lds_stack=None, block_size=256, waves_per_cu=0, schedule="auto", refill=None, chunk=None,
               pool_shift=None):
    o = RtTraceOpts()
    load().rt_trace_opts_default(ctypes.byref(o))
    if lds_stack is not None:
        o.lds_stack = lds_stack
    o.block_size = block_size
    o.waves_per_cu = waves_per_cu
    o.schedule = {"dynamic": 0, "static": 1, "auto": 2}[schedule]
    if refill is not None:
        o.refill = refill
    if chunk is not None:
        o.chunk = chunk
    if pool_shift is not None:
        o.pool_shift = pool_shift
    return o


class HostBvh:
    """A BVH built on the host (no device needed): node bytes and triangle records."""

    def __init__(self, tris, **kw):
        lib = load()
        tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        self.tris_in = tris
        h = ctypes.c_void_p()
        o = build_opts(**kw)
        check(lib.rt_bvh_build(ptr(tris), tris.shape[0], ctypes.byref(o), ctypes.byref(h)))
        self.handle = h
        info = RtSceneInfo()
        check(lib.rt_bvh_info(h, ctypes.byref(info)))
        self.info = info.as_dict()
        pn, pt = ctypes.c_void_p(), ctypes.c_void_p()
        check(lib.rt_bvh_arrays(h, ctypes.byref(pn), ctypes.byref(pt)))
        nb = self.info["nodes_total_bytes"]
        tb = self.info["tris_total_bytes"]
        self.nodes = np.ctypeslib.as_array((ctypes.c_uint8 * nb).from_address(pn.value)).copy() if nb else \
            np.zeros(0, np.uint8)
        self.tris = np.ctypeslib.as_array((ctypes.c_uint8 * tb).from_address(pt.value)).copy().view(
            np.float32).reshape(-1, 12) if tb else np.zeros((0, 12), np.float32)

    def close(self):
        if self.handle:
            load().rt_bvh_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_count():
    n = ctypes.c_int(0)
    rc = load().rt_device_count(ctypes.byref(n))
    return n.value if rc == 0 else 0


class Context:
    def __init__(self, device=0):
        lib = load()
        h = ctypes.c_void_p()
        check(lib.rt_ctx_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            load().rt_ctx_destroy(self.handle)
            self.handle = None

    def check_device_error(self, stream=None):
        code = ctypes.c_uint32(0)
        check(load().rt_check_device_error(self.handle, stream, ctypes.byref(code)), self.handle)
        return code.value

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _stream_handle(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def rays_to_device(rays, device="cuda"):
    """Structured RAY_DTYPE array -> (n, 8) float32 device tensor (bit copy)."""
    import torch
    a = np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)
    return torch.from_numpy(a.copy()).to(device)


def hits_to_numpy(hits):
    """(n, 4) int32 device tensor -> HIT_DTYPE array."""
    return hits.cpu().numpy().view(np.uint32).reshape(-1, 4).copy().view(HIT_DTYPE).reshape(-1)


def make_camera(eye, p0, p1, p2, jitter=(0.5, 0.5)):
    c = RtCamera()
    for k in range(3):
        c.eye[k], c.p0[k], c.p1[k], c.p2[k] = eye[k], p0[k], p1[k], p2[k]
    c.jitter[0], c.jitter[1] = jitter
    return c


def screen_camera(eye, target, vfov_deg, aspect=1.0):
    """Near-plane corners (p0 bottom-left, p1 bottom-right, p2 top-left at
    distance 1) so that rows run top to bottom under the reference's
    cy = 1 - y/height mapping."""
    e = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - e
    f /= np.linalg.norm(f)
    r = np.cross([0.0, 1.0, 0.0], f)
    r /= np.linalg.norm(r)
    u = np.cross(f, r)
    th = np.tan(np.radians(vfov_deg) / 2)
    c = e + f
    p0 = c - aspect * th * r - th * u
    p1 = c + aspect * th * r - th * u
    p2 = c - aspect * th * r + th * u
    return tuple(np.asarray(v, np.float32).tolist() for v in (e, p0, p1, p2))


def lbvh_binary(ctx, device_tris, stream=None):
    """The binary tree every GPU LBVH build starts from, as device tensors
    (morton, order, left, right, boxes); see rt_lbvh_binary."""
    import torch
    n = device_tris.numel() // 9
    dev = device_tris.device
    morton = torch.empty(n, dtype=torch.int32, device=dev)
    order = torch.empty(n, dtype=torch.int32, device=dev)
    left = torch.empty(max(n - 1, 1), dtype=torch.int32, device=dev)
    right = torch.empty(max(n - 1, 1), dtype=torch.int32, device=dev)
    boxes = torch.empty((max(2 * n - 1, 1), 6), dtype=torch.float32, device=dev)
    check(load().rt_lbvh_binary(ctx.handle, ptr(device_tris), n, ptr(morton), ptr(order), ptr(left), ptr(right),
                                ptr(boxes), _stream_handle(stream)), ctx.handle)
    return morton, order, left[:max(n - 1, 0)], right[:max(n - 1, 0)], boxes[:max(2 * n - 1, 0)]


class Scene:
    """A device-resident scene: BVH nodes plus leaf-ordered triangle records."""

    def __init__(self, ctx, tris=None, width=4, builder="sah", max_leaf=4, sah_bins=32, pad_rel=None,
                 keep_source=True, host_bvh=None, device_tris=None, stream=None, packed=None, presplit=1):
        """tris: (n, 9) float32 positions; or packed: the reference's 48-byte
        Triangle records (TRIANGLE_DTYPE), whose normals the scene keeps."""
        lib = load()
        self.ctx = ctx
        h = ctypes.c_void_p()
        if packed is not None:
            packed = np.ascontiguousarray(packed)
            assert packed.dtype.itemsize == 48, "packed records are 48-byte Triangles"
            o = build_opts(width, builder, max_leaf, sah_bins, pad_rel, keep_source, presplit)
            check(lib.rt_scene_create_packed(ctx.handle, ptr(packed) if packed.size else None, packed.shape[0],
                                             ctypes.byref(o), ctypes.byref(h)), ctx.handle)
        elif host_bvh is not None:
            src = host_bvh.tris_in if keep_source else None
            check(lib.rt_scene_from_bvh(ctx.handle, host_bvh.handle, ptr(src), ctypes.byref(h)), ctx.handle)
        elif device_tris is not None:
            o = build_opts(width, "lbvh", max_leaf, sah_bins, pad_rel, keep_source)
            n = device_tris.numel() // 9
            check(lib.rt_scene_create_device(ctx.handle, ptr(device_tris), n, ctypes.byref(o),
                                             _stream_handle(stream), ctypes.byref(h)), ctx.handle)
        else:
            tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
            o = build_opts(width, builder, max_leaf, sah_bins, pad_rel, keep_source, presplit)
            check(lib.rt_scene_create(ctx.handle, ptr(tris), tris.shape[0], ctypes.byref(o), ctypes.byref(h)),
                  ctx.handle)
        self.handle = h
        info = RtSceneInfo()
        check(lib.rt_scene_info_get(h, ctypes.byref(info)))
        self.info = info.as_dict()

    def set_analytic(self, spheres=None, cubes=None, planes=None):
        """Analytic primitives after the triangles in the object-id space
        (trace.glsl:22-46): spheres (n, 4) centre + radius, cubes (n, 6)
        start + end, planes (n, 4) normal + offset.  Replaces the previous set."""
        arrs = []
        for a, w in ((spheres, 4), (cubes, 6), (planes, 4)):
            a = np.zeros((0, w), np.float32) if a is None else np.ascontiguousarray(a, np.float32).reshape(-1, w)
            arrs.append(a)
        sp, cu, pl = arrs
        check(load().rt_scene_set_analytic(self.ctx.handle, self.handle, ptr(sp) if sp.size else None, sp.shape[0],
                                           ptr(cu) if cu.size else None, cu.shape[0],
                                           ptr(pl) if pl.size else None, pl.shape[0]), self.ctx.handle)
        info = RtSceneInfo()
        check(load().rt_scene_info_get(self.handle, ctypes.byref(info)))
        self.info = info.as_dict()

    def download(self):
        """(node bytes, triangle records as float32 (n, 12))."""
        nb, tb = self.info["nodes_total_bytes"], self.info["tris_total_bytes"]
        nodes = np.zeros(nb, np.uint8)
        tris = np.zeros(tb // 4, np.float32)
        check(load().rt_scene_download(self.handle, ptr(nodes) if nb else None, ptr(tris) if tb else None),
              self.ctx.handle)
        return nodes, tris.reshape(-1, 12)

    def close(self):
        if self.handle:
            load().rt_scene_destroy(self.handle)
            self.handle = None

    # ---- queries on device buffers --------------------------------------
    def trace_closest(self, rays, hits=None, stream=None, opts=None, counts=None):
        import torch
        n = rays.shape[0]
        if hits is None:
            hits = torch.empty((n, 4), dtype=torch.int32, device=rays.device)
        lib = load()
        s = _stream_handle(stream)
        if opts is None and counts is None:
            rc = lib.rt_trace_closest(self.ctx.handle, self.handle, ptr(rays), n, ptr(hits), s)
        else:
            rc = lib.rt_trace_closest_ex(self.ctx.handle, self.handle, ptr(rays), n, ptr(hits),
                                         ctypes.byref(opts) if opts is not None else None, ptr(counts), s)
        check(rc, self.ctx.handle)
        return hits

    def trace_occluded(self, rays, flags=None, masks=None, stream=None, opts=None, counts=None, want_flags=True):
        import torch
        n = rays.shape[0]
        if flags is None and want_flags:
            flags = torch.empty(n, dtype=torch.uint8, device=rays.device)
        lib = load()
        s = _stream_handle(stream)
        if opts is None and counts is None and masks is None:
            rc = lib.rt_trace_occluded(self.ctx.handle, self.handle, ptr(rays), n, ptr(flags), s)
        else:
            rc = lib.rt_trace_occluded_ex(self.ctx.handle, self.handle, ptr(rays), n, ptr(flags), ptr(masks),
                                          ctypes.byref(opts) if opts is not None else None, ptr(counts), s)
        check(rc, self.ctx.handle)
        return flags

    def trace_occluded_mask(self, rays, masks=None, stream=None):
        import torch
        n = rays.shape[0]
        if masks is None:
            masks = torch.empty((n + 63) // 64, dtype=torch.int64, device=rays.device)
        check(load().rt_trace_occluded_mask(self.ctx.handle, self.handle, ptr(rays), n, ptr(masks),
                                            _stream_handle(stream)), self.ctx.handle)
        return masks

    def shadow_pass(self, dir_object, eye, light, seed, width, height, samples, tiles=None, stream=None):
        """nv_all.shadow.comp after raygen: tile ballots (int64 words on the
        device) from the dirObject image, the eye, a light dict (make_light)
        and the frame seed (randomX, randomY)."""
        import torch
        from ._lib import make_light
        words = ((width + 15) // 16) * ((height + 3) // 4) * samples
        if tiles is None:
            tiles = torch.empty(words, dtype=torch.int64, device="cuda")
        e = (ctypes.c_float * 3)(*[float(v) for v in eye])
        sd = (ctypes.c_float * 2)(*[float(v) for v in seed])
        check(load().rt_shadow_pass(self.ctx.handle, self.handle, ptr(dir_object), e, ctypes.byref(make_light(light)),
                                    sd, width, height, samples, ptr(tiles), _stream_handle(stream)), self.ctx.handle)
        return tiles

    def raygen_primary(self, cam, width, height, normals=None, dir_object=None, uv_normal=None, stream=None):
        """raygen.comp with calculatePrimary in full; cam is an
        RtPrimaryCamera (primary_camera())."""
        import torch
        if dir_object is None:
            dir_object = torch.empty((height, width, 4), dtype=torch.float32, device="cuda")
        if uv_normal is None:
            uv_normal = torch.empty((height, width, 4), dtype=torch.float32, device="cuda")
        check(load().rt_raygen_primary(self.ctx.handle, self.handle, ctypes.byref(cam), width, height,
                                       ptr(normals), ptr(dir_object), ptr(uv_normal), _stream_handle(stream)),
              self.ctx.handle)
        return dir_object, uv_normal

    def raygen_closest(self, cam, width, height, normals=None, dir_object=None, uv_normal=None, stream=None,
                       jitter=(0.5, 0.5)):
        """raygen.comp: (dir_object, uv_normal) float4 images [h, w, 4] on the device."""
        import torch
        if dir_object is None:
            dir_object = torch.empty((height, width, 4), dtype=torch.float32, device="cuda")
        if uv_normal is None:
            uv_normal = torch.empty((height, width, 4), dtype=torch.float32, device="cuda")
        check(load().rt_raygen_closest(self.ctx.handle, self.handle, ctypes.byref(make_camera(*cam, jitter=jitter)),
                                       width, height,
                                       ptr(normals), ptr(dir_object), ptr(uv_normal), _stream_handle(stream)),
              self.ctx.handle)
        return dir_object, uv_normal

    def shadow_tiles(self, rays, width, height, samples, tiles=None, stream=None):
        """Occlusion ballots in 16x4 pixel tiles; rays ordered [sample][y][x]."""
        import torch
        nwords = ((width + 15) // 16) * ((height + 3) // 4) * samples
        if tiles is None:
            tiles = torch.empty(nwords, dtype=torch.int64, device="cuda")
        check(load().rt_shadow_tiles(self.ctx.handle, self.handle, ptr(rays), width, height, samples, ptr(tiles),
                                     _stream_handle(stream)), self.ctx.handle)
        return tiles

    def trace_closest_brute(self, rays, hits=None, stream=None):
        import torch
        n = rays.shape[0]
        if hits is None:
            hits = torch.empty((n, 4), dtype=torch.int32, device=rays.device)
        check(load().rt_trace_closest_brute(self.ctx.handle, self.handle, ptr(rays), n, ptr(hits),
                                            _stream_handle(stream)), self.ctx.handle)
        return hits

    def trace_occluded_brute(self, rays, flags=None, stream=None):
        import torch
        n = rays.shape[0]
        if flags is None:
            flags = torch.empty(n, dtype=torch.uint8, device=rays.device)
        check(load().rt_trace_occluded_brute(self.ctx.handle, self.handle, ptr(rays), n, ptr(flags),
                                             _stream_handle(stream)), self.ctx.handle)
        return flags


def opts_dict(o):
    return {"block_size": o.block_size, "lds_stack": o.lds_stack, "waves_per_cu": o.waves_per_cu,
            "schedule": o.schedule, "refill": o.refill, "chunk": o.chunk, "pool_shift": o.pool_shift}


def autotune(scene, rays, kind="closest", iters=5, stream=None):
    """Fastest launch shape for this scene and ray sample: (RtTraceOpts, ms)."""
    import torch
    n = rays.shape[0]
    if kind == "closest":
        out = torch.empty((n, 4), dtype=torch.int32, device=rays.device)
    else:
        out = torch.empty(n, dtype=torch.uint8, device=rays.device)
    best = RtTraceOpts()
    ms = ctypes.c_float(0.0)
    check(load().rt_autotune(scene.ctx.handle, scene.handle, 0 if kind == "closest" else 1, ptr(rays), n, ptr(out),
                             iters, ctypes.byref(best), ctypes.byref(ms), _stream_handle(stream)), scene.ctx.handle)
    return best, ms.value


class Batch:
    """Scenes of one node width traced together: every query of a call runs
    in one persistent launch (rt_batch_trace)."""

    def __init__(self, ctx, scenes):
        self.ctx = ctx
        self.scenes = list(scenes)
        arr = (ctypes.c_void_p * len(self.scenes))(*[s.handle.value for s in self.scenes])
        h = ctypes.c_void_p()
        check(load().rt_batch_create(ctx.handle, arr, len(self.scenes), ctypes.byref(h)), ctx.handle)
        self.handle = h

    @staticmethod
    def _table(queries):
        q = (RtQuery * max(len(queries), 1))()
        for i, (si, kind, rays, out) in enumerate(queries):
            q[i].scene = si
            q[i].kind = 0 if kind == "closest" else 1
            q[i].n = rays.shape[0]
            q[i].rays = ptr(rays)
            q[i].out = ptr(out)
        return q

    def trace(self, queries, opts=None, stream=None, table=None):
        """queries: list of (scene index, "closest"|"occluded", device rays, device out)."""
        q = table if table is not None else self._table(queries)
        check(load().rt_batch_trace(self.ctx.handle, self.handle, q, len(queries),
                                    ctypes.byref(opts) if opts is not None else None, _stream_handle(stream)),
              self.ctx.handle)

    def autotune(self, queries, iters=3, stream=None):
        best = RtTraceOpts()
        ms = ctypes.c_float(0.0)
        check(load().rt_batch_autotune(self.ctx.handle, self.handle, self._table(queries), len(queries), iters,
                                       ctypes.byref(best), ctypes.byref(ms), _stream_handle(stream)),
              self.ctx.handle)
        return best, ms.value

    def close(self):
        if self.handle:
            load().rt_batch_destroy(self.handle)
            self.handle = None


class PathTracer:
    """Wavefront diffuse path tracing over a scene (rt_pt_*)."""

    def __init__(self, scene, cam, width, height, spp, max_depth=3, rr_survive=0.8, albedo=0.7, jitter=(0.5, 0.5),
                 chunk=0, refill=0):
        lib = load()
        o = RtPtOpts()
        lib.rt_pt_opts_default(ctypes.byref(o))
        o.width, o.height, o.spp, o.max_depth = width, height, spp, max_depth
        o.rr_survive, o.albedo = rr_survive, albedo
        o.chunk, o.refill = chunk, refill
        self.scene = scene
        self.width, self.height, self.max_depth = width, height, max_depth
        h = ctypes.c_void_p()
        check(lib.rt_pt_create(scene.ctx.handle, scene.handle, ctypes.byref(make_camera(*cam, jitter=jitter)),
                               ctypes.byref(o), ctypes.byref(h)), scene.ctx.handle)
        self.handle = h

    def render(self, seed, image=None, use_graph=True, stream=None):
        import torch
        if image is None:
            image = torch.empty((self.height, self.width), dtype=torch.float32, device="cuda")
        check(load().rt_pt_render(self.handle, seed, ptr(image), 1 if use_graph else 0, _stream_handle(stream)),
              self.scene.ctx.handle)
        return image

    def counts(self):
        import numpy as np
        c = np.zeros(self.max_depth + 1, np.uint32)
        check(load().rt_pt_counts(self.handle, ptr(c), c.size), self.scene.ctx.handle)
        return c

    def queue_rays(self, depth):
        """The rays the last frame traced at `depth` (max_depth or max_depth - 1),
        in queue order, as an (n, 8) float32 device tensor like rays_to_device's."""
        import torch
        n = int(self.counts()[depth])
        out = torch.empty((n, 8), dtype=torch.float32, device="cuda")
        check(load().rt_pt_queue_rays(self.handle, depth, ptr(out) if n else None, n), self.scene.ctx.handle)
        return out

    def close(self):
        if self.handle:
            load().rt_pt_destroy(self.handle)
            self.handle = None

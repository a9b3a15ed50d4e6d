"""ctypes declarations of librt_hip.so (include/rt/rt_c.h, include/rt/rt_workload.h).

The library is built in-tree by ``__graft_entry__.build()`` (make -C
igx_raytracing_amd/csrc).  Loading fails loudly when it is missing: there is
no fallback path.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librt_hip.so")

RT_OK = 0
RT_MISS = 0xFFFFFFFF
RT_NO_HIT_T = np.float32(3.4028235e38)

RAY_DTYPE = np.dtype([("org", "<f4", (3,)), ("tmax", "<f4"), ("dir", "<f4", (3,)), ("prev_id", "<u4")])
HIT_DTYPE = np.dtype([("id", "<u4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])
# The reference's std430 Triangle (primitive.glsl:28-38): positions with
# half2-packed spheremap vertex normals.
TRIANGLE_DTYPE = np.dtype([("p0", "<f4", (3,)), ("n0", "<u4"), ("p1", "<f4", (3,)), ("n1", "<u4"),
                           ("p2", "<f4", (3,)), ("n2", "<u4")])
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 16 and TRIANGLE_DTYPE.itemsize == 48

c_u32 = ctypes.c_uint32
c_p = ctypes.c_void_p


class RtBuildOpts(ctypes.Structure):
    _fields_ = [
        ("builder", c_u32), ("node_width", c_u32), ("max_leaf", c_u32), ("sah_bins", c_u32),
        ("pad_rel", ctypes.c_float), ("keep_source", c_u32), ("presplit", c_u32), ("reserved", c_u32),
    ]


class RtSceneInfo(ctypes.Structure):
    _fields_ = [
        ("num_tris", c_u32), ("node_width", c_u32), ("num_nodes", c_u32), ("root_ref", c_u32),
        ("max_depth", c_u32), ("max_stack", c_u32), ("num_leaves", c_u32), ("node_bytes", c_u32),
        ("nodes_total_bytes", ctypes.c_uint64), ("tris_total_bytes", ctypes.c_uint64),
        ("bounds_lo", ctypes.c_float * 3), ("bounds_hi", ctypes.c_float * 3), ("build_ms", ctypes.c_double),
        ("num_spheres", c_u32), ("num_cubes", c_u32), ("num_planes", c_u32), ("pad_", c_u32),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["bounds_lo"] = list(self.bounds_lo)
        d["bounds_hi"] = list(self.bounds_hi)
        return d


class RtCamera(ctypes.Structure):
    _fields_ = [("eye", ctypes.c_float * 3), ("p0", ctypes.c_float * 3), ("p1", ctypes.c_float * 3),
                ("p2", ctypes.c_float * 3), ("jitter", ctypes.c_float * 2)]


class RtPrimaryCamera(ctypes.Structure):
    """rt_primary_camera: calculatePrimary's camera (camera.glsl:14-44)."""
    _fields_ = [("eye", ctypes.c_float * 3), ("p0", ctypes.c_float * 3), ("p1", ctypes.c_float * 3),
                ("p2", ctypes.c_float * 3), ("p3", ctypes.c_float * 3), ("p4", ctypes.c_float * 3),
                ("p5", ctypes.c_float * 3), ("ipd", ctypes.c_float), ("projection", c_u32),
                ("jitter_mode", c_u32), ("seed", ctypes.c_float * 2), ("jitter", ctypes.c_float * 2)]


PROJECTIONS = {"screen": 0, "omni": 1, "omni_stereo_tb": 2, "stereo_tb": 3, "omni_stereo_lr": 4, "stereo_lr": 5}


def primary_camera(eye, p0, p1, p2, p3=(0, 0, 0), p4=(0, 0, 0), p5=(0, 0, 0), ipd=0.0, projection="screen",
                   seed=(0.0, 0.0), jitter=None):
    """An RtPrimaryCamera; jitter None takes the reference's per-pixel
    rand(loc + seed) offset, a pair a fixed offset in pixels."""
    c = RtPrimaryCamera()
    for name, v in (("eye", eye), ("p0", p0), ("p1", p1), ("p2", p2), ("p3", p3), ("p4", p4), ("p5", p5)):
        arr = getattr(c, name)
        for k in range(3):
            arr[k] = v[k]
    c.ipd = ipd
    c.projection = PROJECTIONS[projection] if isinstance(projection, str) else projection
    c.jitter_mode = 0 if jitter is None else 1
    c.seed[0], c.seed[1] = seed
    if jitter is not None:
        c.jitter[0], c.jitter[1] = jitter
    return c


class RtLight(ctypes.Structure):
    """rt_light: the reference's Light (primitive.glsl:46-54)."""
    _fields_ = [("pos", ctypes.c_float * 3), ("rad_origin", c_u32), ("dir", c_u32 * 2), ("color_type", c_u32 * 2)]


def make_light(light):
    """An RtLight from a dict: pos, rad_origin (half2 bits), dir (2 words), color_type (2 words)."""
    L = RtLight()
    for k in range(3):
        L.pos[k] = float(light["pos"][k])
    L.rad_origin = int(light["rad_origin"])
    L.dir[0], L.dir[1] = (int(x) for x in light["dir"])
    L.color_type[0], L.color_type[1] = (int(x) for x in light["color_type"])
    return L


class RtPtOpts(ctypes.Structure):
    _fields_ = [("width", c_u32), ("height", c_u32), ("spp", c_u32), ("max_depth", c_u32),
                ("rr_survive", ctypes.c_float), ("albedo", ctypes.c_float), ("chunk", c_u32), ("refill", c_u32)]


class RtQuery(ctypes.Structure):
    _fields_ = [("scene", c_u32), ("kind", c_u32), ("n", c_u32), ("reserved", c_u32), ("rays", c_p), ("out", c_p)]


class RtTraceOpts(ctypes.Structure):
    _fields_ = [("block_size", c_u32), ("lds_stack", c_u32), ("waves_per_cu", c_u32), ("schedule", c_u32),
                ("refill", c_u32), ("chunk", c_u32), ("pool_shift", c_u32), ("reserved", c_u32)]


class RtServerOpts(ctypes.Structure):
    _fields_ = [("mem_budget", ctypes.c_uint64), ("streams", c_u32), ("max_tick_rays", c_u32), ("plain", c_u32),
                ("node_width", c_u32), ("max_scenes", c_u32), ("simulate", c_u32), ("launch", RtTraceOpts)]


class RtRequest(ctypes.Structure):
    _fields_ = [("scene", c_u32), ("kind", c_u32), ("n", c_u32), ("memory", c_u32), ("rays", c_p), ("out", c_p),
                ("deadline_ms", ctypes.c_double), ("stream", c_p)]


class RtServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64),
                ("ticks", ctypes.c_uint64), ("rays", ctypes.c_uint64), ("mem_budget", ctypes.c_uint64),
                ("mem_peak", ctypes.c_uint64), ("busy_ms", ctypes.c_double),
                ("mean_requests_per_tick", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# name -> (restype, argtypes); every symbol include/rt/*.h declares.
SIGNATURES = {
    "rt_build_opts_default": (None, [ctypes.POINTER(RtBuildOpts)]),
    "rt_trace_opts_default": (None, [ctypes.POINTER(RtTraceOpts)]),
    "rt_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(c_p)]),
    "rt_ctx_destroy": (ctypes.c_int, [c_p]),
    "rt_last_error": (ctypes.c_char_p, [c_p]),
    "rt_last_error_global": (ctypes.c_char_p, []),
    "rt_sync": (ctypes.c_int, [c_p, c_p]),
    "rt_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "rt_bvh_build": (ctypes.c_int, [c_p, c_u32, ctypes.POINTER(RtBuildOpts), ctypes.POINTER(c_p)]),
    "rt_bvh_info": (ctypes.c_int, [c_p, ctypes.POINTER(RtSceneInfo)]),
    "rt_bvh_arrays": (ctypes.c_int, [c_p, ctypes.POINTER(c_p), ctypes.POINTER(c_p)]),
    "rt_bvh_free": (ctypes.c_int, [c_p]),
    "rt_scene_create": (ctypes.c_int, [c_p, c_p, c_u32, ctypes.POINTER(RtBuildOpts), ctypes.POINTER(c_p)]),
    "rt_scene_create_packed": (ctypes.c_int, [c_p, c_p, c_u32, ctypes.POINTER(RtBuildOpts), ctypes.POINTER(c_p)]),
    "rt_scene_create_device": (ctypes.c_int,
                               [c_p, c_p, c_u32, ctypes.POINTER(RtBuildOpts), c_p, ctypes.POINTER(c_p)]),
    "rt_scene_from_bvh": (ctypes.c_int, [c_p, c_p, c_p, ctypes.POINTER(c_p)]),
    "rt_scene_info_get": (ctypes.c_int, [c_p, ctypes.POINTER(RtSceneInfo)]),
    "rt_scene_set_analytic": (ctypes.c_int, [c_p, c_p, c_p, c_u32, c_p, c_u32, c_p, c_u32]),
    "rt_lbvh_binary": (ctypes.c_int, [c_p, c_p, c_u32, c_p, c_p, c_p, c_p, c_p, c_p]),
    "rt_scene_download": (ctypes.c_int, [c_p, c_p, c_p]),
    "rt_scene_destroy": (ctypes.c_int, [c_p]),
    "rt_trace_closest": (ctypes.c_int, [c_p, c_p, c_p, c_u32, c_p, c_p]),
    "rt_trace_closest_noalloc": (ctypes.c_int, [c_p, c_p, c_p, c_u32, c_p, c_p]),
    "rt_trace_occluded": (ctypes.c_int, [c_p, c_p, c_p, c_u32, c_p, c_p]),
    "rt_trace_occluded_mask": (ctypes.c_int, [c_p, c_p, c_p, c_u32, c_p, c_p]),
    "rt_trace_closest_ex": (ctypes.c_int, [c_p, c_p, c_p, c_u32, c_p, ctypes.POINTER(RtTraceOpts), c_p, c_p]),
    "rt_trace_occluded_ex": (ctypes.c_int,
                             [c_p, c_p, c_p, c_u32, c_p, c_p, ctypes.POINTER(RtTraceOpts), c_p, c_p]),
    "rt_trace_closest_brute": (ctypes.c_int, [c_p, c_p, c_p, c_u32, c_p, c_p]),
    "rt_trace_occluded_brute": (ctypes.c_int, [c_p, c_p, c_p, c_u32, c_p, c_p]),
    "rt_check_device_error": (ctypes.c_int, [c_p, c_p, ctypes.POINTER(c_u32)]),
    "rt_pt_opts_default": (None, [ctypes.POINTER(RtPtOpts)]),
    "rt_pt_create": (ctypes.c_int, [c_p, c_p, ctypes.POINTER(RtCamera), ctypes.POINTER(RtPtOpts),
                                    ctypes.POINTER(c_p)]),
    "rt_pt_render": (ctypes.c_int, [c_p, ctypes.c_uint64, c_p, ctypes.c_int, c_p]),
    "rt_pt_counts": (ctypes.c_int, [c_p, c_p, c_u32]),
    "rt_pt_queue_rays": (ctypes.c_int, [c_p, c_u32, c_p, c_u32]),
    "rt_pt_destroy": (ctypes.c_int, [c_p]),
    "rt_raygen_closest": (ctypes.c_int, [c_p, c_p, ctypes.POINTER(RtCamera), c_u32, c_u32, c_p, c_p, c_p, c_p]),
    "rt_shadow_pass": (ctypes.c_int, [c_p, c_p, c_p, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(RtLight),
                                      ctypes.POINTER(ctypes.c_float), c_u32, c_u32, c_u32, c_p, c_p]),
    "rt_raygen_primary": (ctypes.c_int, [c_p, c_p, ctypes.POINTER(RtPrimaryCamera), c_u32, c_u32, c_p, c_p, c_p,
                                         c_p]),
    "rt_shadow_tiles": (ctypes.c_int, [c_p, c_p, c_p, c_u32, c_u32, c_u32, c_p, c_p]),
    "rt_autotune": (ctypes.c_int, [c_p, c_p, c_u32, c_p, c_u32, c_p, c_u32, ctypes.POINTER(RtTraceOpts),
                                   ctypes.POINTER(ctypes.c_float), c_p]),
    "rt_batch_create": (ctypes.c_int, [c_p, c_p, c_u32, ctypes.POINTER(c_p)]),
    "rt_batch_trace": (ctypes.c_int, [c_p, c_p, c_p, c_u32, ctypes.POINTER(RtTraceOpts), c_p]),
    "rt_batch_destroy": (ctypes.c_int, [c_p]),
    "rt_batch_autotune": (ctypes.c_int, [c_p, c_p, c_p, c_u32, c_u32, ctypes.POINTER(RtTraceOpts),
                                         ctypes.POINTER(ctypes.c_float), c_p]),
    "rt_server_opts_default": (None, [ctypes.POINTER(RtServerOpts)]),
    "rt_server_create": (ctypes.c_int, [c_p, ctypes.POINTER(RtServerOpts), ctypes.POINTER(c_p)]),
    "rt_server_add_scene": (ctypes.c_int, [c_p, c_p, ctypes.POINTER(c_u32)]),
    "rt_server_upload_scene": (ctypes.c_int, [c_p, c_p, c_u32, ctypes.POINTER(c_u32)]),
    "rt_server_submit": (ctypes.c_int, [c_p, ctypes.POINTER(RtRequest), ctypes.POINTER(ctypes.c_uint64)]),
    "rt_server_cancel": (ctypes.c_int, [c_p, ctypes.c_uint64]),
    "rt_server_poll": (ctypes.c_int, [c_p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]),
    "rt_server_wait": (ctypes.c_int, [c_p, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(ctypes.c_int),
                                      ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]),
    "rt_server_forget": (ctypes.c_int, [c_p, ctypes.c_uint64]),
    "rt_server_stats_get": (ctypes.c_int, [c_p, ctypes.POINTER(RtServerStats)]),
    "rt_server_destroy": (ctypes.c_int, [c_p]),
    # rt_workload.h
    "rtw_pcg32_nth": (c_u32, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]),
    "rtw_scene_clusters": (c_u32, [ctypes.c_uint64, c_u32, c_u32, ctypes.c_float, ctypes.c_float,
                                   ctypes.c_float, c_p]),
    "rtw_rays_uniform": (None, [ctypes.c_uint64, c_p, c_p, ctypes.c_uint64, c_u32, ctypes.c_float,
                                ctypes.c_float, c_p]),
    "rtw_scene_mixed4k": (c_u32, [ctypes.c_uint64, c_p]),
    "rtw_rays_pinhole": (None, [c_u32, c_u32, c_p, c_p, ctypes.c_float, c_p]),
    "rtw_bounds": (None, [c_p, c_u32, c_p, c_p]),
    "rtw_scene_icosphere": (c_u32, [ctypes.c_uint64, c_u32, ctypes.c_float, ctypes.c_float, c_p]),
}

_LIB = None


def load():
    """Loads librt_hip.so once; raises if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    # One HIP runtime per process: PyTorch ships its own libamdhip64 with the
    # same soname (libamdhip64.so.7).  Loading torch first makes this library
    # bind to that runtime; loading this library first would make torch load
    # a second runtime, which then finds no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


class RtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rt error {code}: {msg}")
        self.code = code


def check(code, ctx=None):
    if code != RT_OK:
        lib = load()
        msg = lib.rt_last_error(ctx) if ctx else lib.rt_last_error_global()
        raise RtError(code, (msg or b"").decode(errors="replace"))
    return code


def ptr(a):
    """Address of a numpy array or torch tensor (host or device)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()

"""CPU restatement of the reference's ray-query algorithm (ctypes over
liboracle.so, built from oracle.c by oracle/Makefile).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, and only as the checker or the timed CPU
baseline.  The product (igx_raytracing_amd/) never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_LIB = None

c_p = ctypes.c_void_p
c_u32 = ctypes.c_uint32
_F3 = ctypes.POINTER(ctypes.c_float)
_D3 = ctypes.POINTER(ctypes.c_double)


def load():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C oracle)")
        lib = ctypes.CDLL(LIB_PATH)
        lib.orc_mt_f32.restype = ctypes.c_int
        lib.orc_mt_f32.argtypes = [_F3] * 5 + [_F3] * 3
        lib.orc_mt_f64.restype = ctypes.c_int
        lib.orc_mt_f64.argtypes = [_D3] * 5 + [_D3] * 3
        lib.orc_closest_brute.restype = None
        lib.orc_closest_brute.argtypes = [c_p, c_u32, c_p, c_u32, c_p, c_p, ctypes.c_int]
        lib.orc_occluded_brute.restype = None
        lib.orc_occluded_brute.argtypes = [c_p, c_u32, c_p, c_u32, c_p, ctypes.c_int]
        for name in ("orc_closest_brute_guard", "orc_occluded_brute_guard"):
            fn = getattr(lib, name)
            fn.restype = None
            fn.argtypes = [c_p, c_u32, c_p, c_u32, c_p, ctypes.c_double, ctypes.c_int]
        lib.orc_walk.restype = None
        lib.orc_walk.argtypes = [ctypes.c_int, c_p, c_p, c_u32, c_p, c_u32, ctypes.c_int, c_p, c_p, c_p,
                                 ctypes.c_int]
        lib.orc_lbvh_binary.restype = ctypes.c_int
        lib.orc_lbvh_binary.argtypes = [c_p, c_u32, c_p, c_p, c_p, c_p, c_p]
        lib.orc_raygen.restype = None
        lib.orc_raygen.argtypes = [c_p, c_u32, c_p, c_p, c_u32, c_u32, c_p, c_p, ctypes.c_int]
        lib.orc_pathtrace.restype = None
        lib.orc_pathtrace.argtypes = [c_p, c_u32, c_p, c_u32, c_u32, c_u32, c_u32, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_uint64, c_p, c_p, ctypes.c_int]
        lib.orc_pathtrace_pixels.restype = None
        lib.orc_pathtrace_pixels.argtypes = [c_p, c_u32, c_p, c_u32, c_u32, c_u32, c_u32, ctypes.c_float,
                                             ctypes.c_float, ctypes.c_uint64, c_p, c_u32, c_p, ctypes.c_int]
        lib.orc_closest_brute_ana.restype = None
        lib.orc_closest_brute_ana.argtypes = [c_p, c_u32, c_p, c_p, c_u32, c_p, c_p, ctypes.c_int]
        lib.orc_occluded_brute_ana.restype = None
        lib.orc_occluded_brute_ana.argtypes = [c_p, c_u32, c_p, c_p, c_u32, c_p, ctypes.c_int]
        lib.orc_primary_rays.restype = None
        lib.orc_primary_rays.argtypes = [c_p, c_u32, c_u32, c_p]
        lib.orc_raygen_primary.restype = None
        lib.orc_raygen_primary.argtypes = [c_p, c_u32, c_p, c_p, c_p, c_u32, c_u32, c_p, c_p, ctypes.c_int]
        lib.orc_shadow_rays.restype = None
        lib.orc_shadow_rays.argtypes = [c_p, c_p, c_p, c_p, c_u32, c_u32, c_u32, c_p, c_p]
        lib.orc_shadow_pass.restype = None
        lib.orc_shadow_pass.argtypes = [c_p, c_u32, c_p, c_p, c_p, c_p, c_p, c_u32, c_u32, c_u32, c_p, ctypes.c_int]
        lib.orc_raygen_ana.restype = None
        lib.orc_raygen_ana.argtypes = [c_p, c_u32, c_p, c_p, c_p, c_u32, c_u32, c_p, c_p, ctypes.c_int]
        lib.orc_asinf.restype = ctypes.c_float
        lib.orc_asinf.argtypes = [ctypes.c_float]
        lib.orc_atanf.restype = ctypes.c_float
        lib.orc_atanf.argtypes = [ctypes.c_float]
        for name in ("orc_sphere", "orc_cube", "orc_plane"):
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [c_p, c_p, c_p, c_u32, c_u32, c_p, c_p, c_p]
        lib.orc_now.restype = ctypes.c_double
        lib.orc_now.argtypes = []
        _LIB = lib
    return _LIB


def _f3(x):
    a = (ctypes.c_float * 3)(*[float(v) for v in x])
    return a


def mt_f32(o, d, p0, p1, p2):
    """(hit, t, u, v) of one pair in FP32; hit False when u/v reject."""
    t, u, v = (ctypes.c_float(), ctypes.c_float(), ctypes.c_float())
    args = [(ctypes.c_float * 3)(*np.asarray(x, np.float32).tolist()) for x in (o, d, p0, p1, p2)]
    ok = load().orc_mt_f32(*args, ctypes.byref(t), ctypes.byref(u), ctypes.byref(v))
    return bool(ok), np.float32(t.value), np.float32(u.value), np.float32(v.value)


def mt_f64(o, d, p0, p1, p2):
    t, u, v = (ctypes.c_double(), ctypes.c_double(), ctypes.c_double())
    args = [(ctypes.c_double * 3)(*np.asarray(x, np.float64).tolist()) for x in (o, d, p0, p1, p2)]
    ok = load().orc_mt_f64(*args, ctypes.byref(t), ctypes.byref(u), ctypes.byref(v))
    return bool(ok), t.value, u.value, v.value


def _rays_f32(rays):
    return np.ascontiguousarray(rays).view(np.float32).reshape(-1, 8)


class _Analytic(ctypes.Structure):
    _fields_ = [("spheres", c_p), ("num_spheres", c_u32), ("cubes", c_p), ("num_cubes", c_u32),
                ("planes", c_p), ("num_planes", c_u32)]


def _analytic(analytic):
    """(struct, keep-alive arrays) for a dict with optional 'spheres' (n,4),
    'cubes' (n,6) and 'planes' (n,4) float32 arrays; (None, ()) for None."""
    if analytic is None:
        return None, ()
    arrs = []
    st = _Analytic()
    for key, width in (("spheres", 4), ("cubes", 6), ("planes", 4)):
        a = analytic.get(key)
        a = np.zeros((0, width), np.float32) if a is None else np.ascontiguousarray(a, np.float32).reshape(-1, width)
        arrs.append(a)
        setattr(st, key, a.ctypes.data if a.size else None)
        setattr(st, "num_" + key, a.shape[0])
    return st, arrs


def closest_brute(tris, rays, nthreads=8, with_nan=False, analytic=None):
    """trace.glsl:9-66 over input-order triangles, then the analytic
    primitives (spheres, cubes, planes) if given; returns (n, 4) uint32 hits."""
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    r = _rays_f32(rays)
    hits = np.empty((r.shape[0], 4), np.uint32)
    nanf = np.zeros(r.shape[0], np.uint8)
    st, keep = _analytic(analytic)
    load().orc_closest_brute_ana(tris.ctypes.data, tris.shape[0], ctypes.byref(st) if st is not None else None,
                                 r.ctypes.data, r.shape[0], hits.ctypes.data, nanf.ctypes.data, nthreads)
    del keep
    return (hits, nanf) if with_nan else hits


def closest_brute_guard(tris, rays, guard, nthreads=8):
    """closest_brute without FP32 artifacts: pairs whose accepted t is NaN or
    puts the hit point more than `guard` outside the triangle's box."""
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    r = _rays_f32(rays)
    hits = np.empty((r.shape[0], 4), np.uint32)
    load().orc_closest_brute_guard(tris.ctypes.data, tris.shape[0], r.ctypes.data, r.shape[0], hits.ctypes.data,
                                   ctypes.c_double(guard), nthreads)
    return hits


def occluded_brute_guard(tris, rays, guard, nthreads=8):
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    r = _rays_f32(rays)
    flags = np.empty(r.shape[0], np.uint8)
    load().orc_occluded_brute_guard(tris.ctypes.data, tris.shape[0], r.ctypes.data, r.shape[0], flags.ctypes.data,
                                    ctypes.c_double(guard), nthreads)
    return flags


def occluded_brute(tris, rays, nthreads=8, analytic=None):
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    r = _rays_f32(rays)
    flags = np.empty(r.shape[0], np.uint8)
    st, keep = _analytic(analytic)
    load().orc_occluded_brute_ana(tris.ctypes.data, tris.shape[0], ctypes.byref(st) if st is not None else None,
                                  r.ctypes.data, r.shape[0], flags.ctypes.data, nthreads)
    del keep
    return flags


def asinf(x):
    return np.float32(load().orc_asinf(float(x)))


def atanf(x):
    return np.float32(load().orc_atanf(float(x)))


def analytic_test(kind, o, d, prim, obj=0, prev=0xFFFFFFFF, best_t=np.float32(3.4028235e38)):
    """One primitive test ('sphere', 'cube', 'plane'): (accepted, t, u, v, normal)."""
    f3 = lambda x: (ctypes.c_float * len(x))(*np.asarray(x, np.float32).tolist())
    bt = ctypes.c_float(float(best_t))
    uv = (ctypes.c_float * 2)()
    n = (ctypes.c_float * 3)()
    ok = getattr(load(), "orc_" + kind)(f3(o), f3(d), f3(prim), obj, prev, ctypes.byref(bt), uv, n)
    return (bool(ok), np.float32(bt.value), np.float32(uv[0]), np.float32(uv[1]),
            np.array(n[:], np.float32))


def walk(width, nodes, tri_records, root, rays, occlusion=False, nthreads=8, counts=False):
    """Traversal restatement over built node bytes; returns hits (n,4) or flags, and counts (n,2)."""
    r = _rays_f32(rays)
    n = r.shape[0]
    nodes = np.ascontiguousarray(nodes)
    recs = np.ascontiguousarray(tri_records, np.float32)
    hits = None if occlusion else np.empty((n, 4), np.uint32)
    flags = np.empty(n, np.uint8) if occlusion else None
    cnt = np.empty((n, 2), np.uint32) if counts else None
    load().orc_walk(width, nodes.ctypes.data if nodes.size else None, recs.ctypes.data if recs.size else None,
                    root, r.ctypes.data, n, 1 if occlusion else 0,
                    hits.ctypes.data if hits is not None else None,
                    flags.ctypes.data if flags is not None else None,
                    cnt.ctypes.data if cnt is not None else None, nthreads)
    out = flags if occlusion else hits
    return (out, cnt) if counts else out


def lbvh_binary(tris):
    """(morton, order, left, right, boxes) of the binary LBVH, as rt_lbvh_binary returns them."""
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    n = tris.shape[0]
    morton = np.empty(n, np.uint32)
    order = np.empty(n, np.uint32)
    left = np.empty(max(n - 1, 1), np.uint32)
    right = np.empty(max(n - 1, 1), np.uint32)
    boxes = np.empty((max(2 * n - 1, 1), 6), np.float32)
    rc = load().orc_lbvh_binary(tris.ctypes.data, n, morton.ctypes.data, order.ctypes.data, left.ctypes.data,
                                right.ctypes.data, boxes.ctypes.data)
    assert rc == 0
    return morton, order, left[:max(n - 1, 0)], right[:max(n - 1, 0)], boxes[:max(2 * n - 1, 0)]


def raygen(tris, cam, width, height, normals=None, nthreads=8, fill=np.float32(-7.0), analytic=None):
    """(dir_object, uv_normal) float32 [h, w, 4]; uv_normal keeps `fill` where rays miss."""
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    c = np.ascontiguousarray(cam, np.float32).reshape(14)
    dob = np.empty((height, width, 4), np.float32)
    uvn = np.full((height, width, 4), fill, np.float32)
    nrm = None if normals is None else np.ascontiguousarray(normals, np.uint32)
    st, keep = _analytic(analytic)
    load().orc_raygen_ana(tris.ctypes.data if tris.size else None, tris.shape[0],
                          nrm.ctypes.data if nrm is not None else None,
                          ctypes.byref(st) if st is not None else None, c.ctypes.data,
                          width, height, dob.ctypes.data, uvn.ctypes.data, nthreads)
    del keep
    return dob, uvn


class _Camera(ctypes.Structure):
    _fields_ = [("eye", ctypes.c_float * 3), ("p0", ctypes.c_float * 3), ("p1", ctypes.c_float * 3),
                ("p2", ctypes.c_float * 3), ("p3", ctypes.c_float * 3), ("p4", ctypes.c_float * 3),
                ("p5", ctypes.c_float * 3), ("ipd", ctypes.c_float), ("projection", c_u32),
                ("rand_jitter", c_u32), ("seed", ctypes.c_float * 2), ("jitter", ctypes.c_float * 2)]


def _camera(cam):
    """An orc_camera from a dict: eye, p0..p5, ipd, projection (0..5),
    seed (randLoc), and jitter (None: rand(loc + seed) per pixel)."""
    c = _Camera()
    for name in ("eye", "p0", "p1", "p2", "p3", "p4", "p5"):
        v = cam.get(name, (0.0, 0.0, 0.0))
        arr = getattr(c, name)
        for k in range(3):
            arr[k] = float(v[k])
    c.ipd = float(cam.get("ipd", 0.0))
    c.projection = int(cam.get("projection", 0))
    jit = cam.get("jitter")
    c.rand_jitter = 1 if jit is None else 0
    c.seed[0], c.seed[1] = cam.get("seed", (0.0, 0.0))
    if jit is not None:
        c.jitter[0], c.jitter[1] = jit
    return c


def primary_rays(cam, width, height):
    """calculatePrimary's rays [h * w] (RAY_DTYPE-compatible (n, 8) float32)."""
    out = np.empty((width * height, 8), np.float32)
    c = _camera(cam)
    load().orc_primary_rays(ctypes.byref(c), width, height, out.ctypes.data)
    return out


def raygen_primary(tris, cam, width, height, normals=None, nthreads=8, fill=np.float32(-7.0), analytic=None):
    """raygen with calculatePrimary in full: (dir_object, uv_normal) [h, w, 4]."""
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    c = _camera(cam)
    dob = np.empty((height, width, 4), np.float32)
    uvn = np.full((height, width, 4), fill, np.float32)
    nrm = None if normals is None else np.ascontiguousarray(normals, np.uint32)
    st, keep = _analytic(analytic)
    load().orc_raygen_primary(tris.ctypes.data if tris.size else None, tris.shape[0],
                              nrm.ctypes.data if nrm is not None else None,
                              ctypes.byref(st) if st is not None else None, ctypes.byref(c),
                              width, height, dob.ctypes.data, uvn.ctypes.data, nthreads)
    del keep
    return dob, uvn


class _Light(ctypes.Structure):
    _fields_ = [("pos", ctypes.c_float * 3), ("rad_origin", c_u32), ("dir", c_u32 * 2), ("color_type", c_u32 * 2)]


def _light(light):
    """An orc_light from a dict: pos, rad_origin (u32 half2 bits), dir (2 u32),
    color_type (2 u32)."""
    L = _Light()
    for k in range(3):
        L.pos[k] = float(light["pos"][k])
    L.rad_origin = int(light["rad_origin"])
    L.dir[0], L.dir[1] = (int(x) for x in light["dir"])
    L.color_type[0], L.color_type[1] = (int(x) for x in light["color_type"])
    return L


def shadow_rays(dir_object, eye, light, seed, width, height, samples):
    """nv_all.shadow.comp's rays [S, H, W] (n, 8) float32 and trace flags."""
    dob = np.ascontiguousarray(dir_object, np.float32)
    n = width * height * samples
    rays = np.zeros((n, 8), np.float32)
    trace = np.zeros(n, np.uint8)
    L = _light(light)
    load().orc_shadow_rays(dob.ctypes.data, _f3(eye), ctypes.byref(L), (ctypes.c_float * 2)(*seed), width, height,
                           samples, rays.ctypes.data, trace.ctypes.data)
    return rays, trace


def shadow_pass(tris, dir_object, eye, light, seed, width, height, samples, analytic=None, nthreads=8):
    """nv_all.shadow.comp's tile ballots (uint64 words)."""
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    dob = np.ascontiguousarray(dir_object, np.float32)
    words = np.zeros(((width + 15) // 16) * ((height + 3) // 4) * samples, np.uint64)
    L = _light(light)
    st, keep = _analytic(analytic)
    load().orc_shadow_pass(tris.ctypes.data if tris.size else None, tris.shape[0],
                           ctypes.byref(st) if st is not None else None, dob.ctypes.data, _f3(eye), ctypes.byref(L),
                           (ctypes.c_float * 2)(*seed), width, height, samples, words.ctypes.data, nthreads)
    del keep
    return words


def pathtrace(tris, cam, width, height, spp, max_depth=3, rr=0.8, albedo=0.7, seed=0, nthreads=8):
    """(image float32 [h, w], rays traced per depth)."""
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    c = np.ascontiguousarray(cam, np.float32).reshape(14)
    img = np.empty((height, width), np.float32)
    counts = np.zeros(max_depth + 1, np.uint32)
    load().orc_pathtrace(tris.ctypes.data, tris.shape[0], c.ctypes.data, width, height, spp, max_depth, rr, albedo,
                         seed, img.ctypes.data, counts.ctypes.data, nthreads)
    return img, counts


def pathtrace_pixels(tris, cam, width, height, spp, pixels, max_depth=3, rr=0.8, albedo=0.7, seed=0, nthreads=8):
    """Image values of the listed pixel indices (y * width + x) only."""
    tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    c = np.ascontiguousarray(cam, np.float32).reshape(14)
    px = np.ascontiguousarray(pixels, np.uint32)
    vals = np.empty(px.shape[0], np.float32)
    load().orc_pathtrace_pixels(tris.ctypes.data, tris.shape[0], c.ctypes.data, width, height, spp, max_depth, rr,
                                albedo, seed, px.ctypes.data, px.shape[0], vals.ctypes.data, nthreads)
    return vals


def now():
    return load().orc_now()

"""Seeded synthetic workloads (numpy views over librt_hip.so's generators).

The reference ships no scenes or ray sets for its query path; these are the
benchmark inputs.  Triangles are float32 arrays of shape (n, 9) = p0 p1 p2 in
input order; rays are structured arrays of RAY_DTYPE.
"""
import numpy as np

from ._lib import RAY_DTYPE, load, ptr

SEED_BASE = 0x5EED0000


def seed(k):
    """Seed of workload k."""
    return SEED_BASE + k


def clusters(seed_, nclusters=256, tris_per_cluster=4096, range_=50.0, sigma=1.0, edge_sigma=0.05):
    """Gaussian clusters: centres U[-range, range]^3, anchors N(0, sigma), vertices N(0, edge_sigma)."""
    out = np.empty((nclusters * tris_per_cluster, 9), np.float32)
    load().rtw_scene_clusters(seed_, nclusters, tris_per_cluster, range_, sigma, edge_sigma, ptr(out))
    return out


def bounds(tris):
    lo = np.empty(3, np.float32)
    hi = np.empty(3, np.float32)
    tris = np.ascontiguousarray(tris, np.float32)
    load().rtw_bounds(ptr(tris), tris.shape[0], ptr(lo), ptr(hi))
    return lo, hi


def rays_uniform(seed_, lo, hi, n, first=0, tmax=None):
    """Origins uniform in [lo, hi], directions uniform on the sphere.

    tmax None gives closest-hit rays (tmax = FLT_MAX); a pair (a, b) draws
    tmax ~ U[a, b] (occlusion rays)."""
    out = np.empty(n, RAY_DTYPE)
    lo = np.ascontiguousarray(lo, np.float32)
    hi = np.ascontiguousarray(hi, np.float32)
    a, b = (0.0, 0.0) if tmax is None else tmax
    load().rtw_rays_uniform(seed_, ptr(lo), ptr(hi), first, n, a, b, ptr(out))
    return out


def mixed4k(seed_):
    """4,096-triangle scene: 64 cubes, a 32x32 ground grid, 1,280 loose triangles."""
    out = np.empty((4096, 9), np.float32)
    n = load().rtw_scene_mixed4k(seed_, ptr(out))
    assert n == 4096
    return out


def pinhole(width, height, eye=(0.0, 5.0, -25.0), target=(0.0, 0.0, 0.0), vfov_deg=60.0):
    out = np.empty(width * height, RAY_DTYPE)
    e = np.asarray(eye, np.float32)
    t = np.asarray(target, np.float32)
    load().rtw_rays_pinhole(width, height, ptr(e), ptr(t), vfov_deg, ptr(out))
    return out


def icosphere(seed_, subdiv, radius=10.0, disp=0.5):
    n = 20 * 4 ** subdiv
    out = np.empty((n, 9), np.float32)
    m = load().rtw_scene_icosphere(seed_, subdiv, radius, disp, ptr(out))
    assert m == n
    return out


def pcg32_nth(seed_, stream, skip):
    return load().rtw_pcg32_nth(seed_, stream, skip)


# Workload definitions used by the tests and the benchmark.
def c1_scene():
    return mixed4k(seed(0))


def c1_rays(width=512, height=512):
    return pinhole(width, height)


def c2_scene():
    return clusters(seed(1), 256, 4096)


def c2_rays(tris, n=1 << 20, first=0, occlusion=False):
    lo, hi = bounds(tris)
    return rays_uniform(seed(1), lo, hi, n, first=first, tmax=(0.0, 20.0) if occlusion else None)


def c3_scene(subdiv=8, cluster_tris=786_432):
    """Displaced icosphere (radius 10, displacement 0.5 N(0,1) per vertex) plus
    clustered triangles: 20 * 4^8 + 786,432 = 2,097,152 triangles by default
    (no icosphere has exactly 2^21 triangles)."""
    sphere = icosphere(seed(2), subdiv, 10.0, 0.5)
    extra = clusters(seed(3), max(1, cluster_tris // 4096), min(4096, cluster_tris)) if cluster_tris else \
        np.zeros((0, 9), np.float32)
    return np.concatenate([sphere, extra])


def niels_scene():
    """The reference's demo scene (test/scene/niels_scene.cpp): 3 triangles
    in Triangle{p0, p1, p2} argument order (:26-42), and its analytic
    objects (:21-24,44-47,54-56, with the positions NielsScene::update sets
    at :61-71 at time 0, the two-argument Vec3f32 taken as z = 0): spheres
    centre + radius (the shader squares w, primitive.glsl:180), cubes
    start.xyz end.xyz, planes normal.xyz offset.  The igx SceneGraph packing
    order is not in the reference (its submodule is empty): static spheres
    first, then the dynamic ones.  Returns (tris (3, 9), analytic dict)."""
    tris = np.array([
        [1, 1, 0, -1, 1, 0, 1, 0, 1],
        [-1, 4, 0, 1, 4, 0, 1, 3, 1],
        [-1, 7, 0, 1, 7, 0, 1, 5, 1],
    ], np.float32)
    analytic = {
        "spheres": np.array([
            [0, 1, 5, 1], [0, 1, -5, 1], [3, 1, 0, 1], [0, 6, 0, 1],
            [7, 2, 0, 1], [-5, 3, 0, 1], [0, 4, 0, 1],
        ], np.float32),
        "cubes": np.array([[0, 0, 0, 1, 1, 1], [-2, 0, -2, -1, 1, -1]], np.float32),
        "planes": np.array([[0, 1, 0, 0]], np.float32),
    }
    return tris, analytic


def analytic_mix(seed_, nspheres=64, ncubes=32):
    """Random analytic objects over [-14, 14]^3: spheres (radius U[0.2, 2]),
    axis-aligned cubes (extent U[0.3, 3]) and two planes."""
    rng = np.random.default_rng(seed_)
    lo = rng.uniform(-14, 12, (ncubes, 3))
    return {
        "spheres": np.concatenate([rng.uniform(-14, 14, (nspheres, 3)), rng.uniform(0.2, 2.0, (nspheres, 1))],
                                  1).astype(np.float32),
        "cubes": np.concatenate([lo, lo + rng.uniform(0.3, 3.0, (ncubes, 3))], 1).astype(np.float32),
        "planes": np.array([[0, 1, 0, -1.5], [0.3, 1, -0.2, -30]], np.float32),
    }


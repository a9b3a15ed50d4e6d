"""Rays aimed at the weak spots of conservative box culling.

A BVH may only skip a box when no triangle inside can be accepted by the
Moller-Trumbore test (res/shaders/primitive.glsl:239-284) in FP32.  MT can
accept a pair whose exact hit point lies a little outside the triangle's
tight box, and the slab test itself rounds, so these rays probe where the
margins are thinnest:

* grazing: rays at 1e-7 .. 1e-2 rad to a triangle's plane, and exactly in
  it, aimed at a point inside the triangle;
* axis-aligned: directions with exactly zero components (the slab test's
  1/0 case), from origins exactly on triangle vertices, edge midpoints and
  face centres: the cube faces and the ground grid of the 4,096-triangle
  scene;
* far: origins ~8x the scene's scale away (slab distances carry the
  rounding of o/d), aimed at points inside triangles.  The box padding
  (2^-16 of the scene scale) covers the slab test's rounding for origins up
  to about 2^8 scene scales away.

Each set is a RAY_DTYPE array with tmax = FLT_MAX (closest hit); occlusion
variants draw tmax from the distance to the aimed point.
"""
import numpy as np

from igx_raytracing_amd import RAY_DTYPE, RT_MISS, RT_NO_HIT_T


def _rays(org, d, tmax=None):
    out = np.empty(org.shape[0], RAY_DTYPE)
    out["org"] = org.astype(np.float32)
    out["dir"] = d.astype(np.float32)
    out["tmax"] = RT_NO_HIT_T if tmax is None else tmax.astype(np.float32)
    out["prev_id"] = RT_MISS
    return out


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def grazing(tris, n, seed):
    rng = np.random.default_rng(seed)
    t = tris[rng.integers(0, tris.shape[0], n)].astype(np.float64).reshape(n, 3, 3)
    e1, e2 = t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]
    nrm = np.cross(e1, e2)
    ok = np.linalg.norm(nrm, axis=1) > 0
    t, e1, e2, nrm = t[ok], e1[ok], e2[ok], _unit(nrm[ok])
    m = t.shape[0]
    a, b = rng.uniform(0, 1, m), rng.uniform(0, 1, m)
    flip = a + b > 1
    a[flip], b[flip] = 1 - a[flip], 1 - b[flip]
    p = t[:, 0] + a[:, None] * e1 + b[:, None] * e2
    tang = rng.normal(size=(m, 3))
    tang = _unit(tang - (tang * nrm).sum(1, keepdims=True) * nrm)
    ang = 10.0 ** rng.uniform(-7, -2, m)
    ang[rng.uniform(size=m) < 0.15] = 0.0  # exactly in the plane
    sgn = np.where(rng.uniform(size=m) < 0.5, -1.0, 1.0)
    d = _unit(np.cos(ang)[:, None] * tang + (sgn * np.sin(ang))[:, None] * nrm)
    s = rng.uniform(0.05, 5.0, m)
    org = p - s[:, None] * d
    return _rays(org, d), s


AXIS_DIRS = np.array(
    [[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
     [1, 1, 0], [1, -1, 0], [-1, 0, 1], [0, 1, -1], [0, -1, -1], [-1, 0, -1]], np.float64)


def axis_aligned(tris, n, seed):
    rng = np.random.default_rng(seed)
    v = tris.reshape(-1, 3, 3).astype(np.float32)
    pts = [v.reshape(-1, 3),                                  # vertices
           ((v[:, 0] + v[:, 1]) * np.float32(0.5)),           # edge midpoints
           ((v[:, 1] + v[:, 2]) * np.float32(0.5)),           # the hypotenuse of a cube face's half:
           ]                                                  # its midpoint is the face centre
    pts = np.concatenate(pts).astype(np.float32)
    org = pts[rng.integers(0, pts.shape[0], n)]
    d = _unit(AXIS_DIRS[rng.integers(0, AXIS_DIRS.shape[0], n)])
    d32 = d.astype(np.float32)
    d32[np.abs(d) < 1e-30] = 0.0  # exact zeros
    return _rays(org, d32)


def far(tris, n, seed, radius=None):
    """Origins at ~`radius` (default 8x the scene's largest |coordinate| or
    extent) from the scene's centre."""
    rng = np.random.default_rng(seed)
    v = tris.reshape(-1, 3).astype(np.float64)
    lo, hi = v.min(0), v.max(0)
    centre = (lo + hi) / 2
    if radius is None:
        radius = 8.0 * max(np.abs(lo).max(), np.abs(hi).max(), (hi - lo).max())
    t = tris[rng.integers(0, tris.shape[0], n)].astype(np.float64).reshape(n, 3, 3)
    a, b = rng.uniform(0, 1, n), rng.uniform(0, 1, n)
    flip = a + b > 1
    a[flip], b[flip] = 1 - a[flip], 1 - b[flip]
    p = t[:, 0] + a[:, None] * (t[:, 1] - t[:, 0]) + b[:, None] * (t[:, 2] - t[:, 0])
    org = centre + _unit(rng.normal(size=(n, 3))) * radius * rng.uniform(0.8, 1.2, (n, 1))
    d = _unit(p - org)
    # every third ray along a coordinate axis, with exact zeros
    ax = np.arange(n) % 3 == 0
    k = rng.integers(0, 3, n)
    d2 = np.zeros((n, 3))
    d2[np.arange(n), k] = np.sign(p[np.arange(n), k] - org[np.arange(n), k])
    org2 = p.copy()
    org2[np.arange(n), k] = p[np.arange(n), k] - d2[np.arange(n), k] * radius
    org = np.where(ax[:, None], org2, org)
    d = np.where(ax[:, None], d2, d)
    return _rays(org, d), np.linalg.norm(p - org, axis=1)


def all_sets(tris, n_each, seed):
    """Closest-hit and occlusion ray sets: (name, rays) pairs."""
    g, gs = grazing(tris, n_each, seed)
    a = axis_aligned(tris, n_each, seed + 1)
    f, fs = far(tris, n_each, seed + 2)
    rng = np.random.default_rng(seed + 3)
    go = g.copy()
    go["tmax"] = (gs[: go.shape[0]] * rng.uniform(0.5, 1.5, go.shape[0])).astype(np.float32)
    ao = a.copy()
    ao["tmax"] = rng.uniform(0.0, 20.0, ao.shape[0]).astype(np.float32)
    fo = f.copy()
    fo["tmax"] = (fs * rng.uniform(0.99, 1.01, fo.shape[0])).astype(np.float32)
    return {"grazing": (g, go), "axis": (a, ao), "far": (f, fo)}

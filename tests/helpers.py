"""Shared helpers of the tests: comparisons and derived ray sets."""
import numpy as np

from igx_raytracing_amd import RAY_DTYPE, RT_MISS
from igx_raytracing_amd.workload import niels_scene

# The reference's demo scene (NielsScene): 3 triangles and 10 analytic objects.
NIELS_TRIS, NIELS_ANALYTIC = niels_scene()


def hits_u32(h):
    """Any hit representation -> (n, 4) uint32 (id, t bits, u bits, v bits)."""
    if hasattr(h, "cpu"):
        h = h.cpu().numpy()
    a = np.ascontiguousarray(h)
    if a.dtype.names:
        return a.view(np.uint32).reshape(-1, 4)
    return a.view(np.uint32).reshape(-1, 4)


def assert_hits_equal(got, want, what=""):
    g, w = hits_u32(got), hits_u32(want)
    assert g.shape == w.shape, (g.shape, w.shape)
    bad = np.nonzero((g != w).any(axis=1))[0]
    if bad.size:
        i = bad[0]
        raise AssertionError(f"{what}: {bad.size} of {g.shape[0]} rays differ; first ray {i}: got "
                             f"id={g[i, 0]:#x} t={g[i, 1:].view(np.float32)} want id={w[i, 0]:#x} "
                             f"t={w[i, 1:].view(np.float32)}")


def shadow_rays(rays, hits_u, light):
    """Shadow rays from each closest hit toward a point light, as the reference's
    shadow pass builds them: origin = hit position, tmax = distance to the light,
    prev_id = the hit triangle (nv_all.shadow.comp:84,104,119)."""
    h = hits_u32(hits_u)
    hit = h[:, 0] != RT_MISS
    t = h[:, 1].view(np.float32)
    org = rays["org"][hit]
    d = rays["dir"][hit]
    p = (org + t[hit, None] * d).astype(np.float32)
    lvec = (np.asarray(light, np.float32)[None, :] - p).astype(np.float32)
    dist = np.sqrt((lvec.astype(np.float64) ** 2).sum(1)).astype(np.float32)
    out = np.empty(int(hit.sum()), RAY_DTYPE)
    out["org"] = p
    out["dir"] = (lvec / dist[:, None]).astype(np.float32)
    out["tmax"] = dist
    out["prev_id"] = h[hit, 0]
    return out


def to_dev(rays):
    from igx_raytracing_amd import rays_to_device
    return rays_to_device(rays)


def scene_pad(tris, pad_rel=2.0 ** -16):
    """The builders' box padding: pad_rel times the largest |coordinate| or
    extent of the scene's box (bvh_build.cpp pad_and_refit)."""
    v = np.asarray(tris, np.float32).reshape(-1, 3).astype(np.float64)
    lo, hi = v.min(0), v.max(0)
    return float(max(np.abs(lo).max(), np.abs(hi).max(), (hi - lo).max())) * pad_rel


def _key(h):
    """(t, id) sort keys of (n, 4) uint32 hits; a miss sorts last."""
    return h[:, 1].view(np.float32).astype(np.float64), h[:, 0].astype(np.int64)


def assert_hits_adjudicated(got, tris, rays, full, guarded, lower, what=""):
    """BVH hits against the reference's scan, with FP32 artifacts adjudicated.

    `full` is the scan (oracle.closest_brute); `guarded` the scan without
    FP32 artifacts (oracle.closest_brute_guard with half the builders' box
    padding); `lower` the scan without NaN accepts only (guard = inf).  Where
    full and guarded agree, no artifact decides the ray and the BVH must
    equal them bit for bit.  Where they differ, the scan's winner is an
    artifact a BVH cannot be obliged to reach (or a NaN the BVH rejects); the
    BVH's answer must then lie between lower and guarded in (t, id) order
    and be a genuine FP32 accept of its triangle (or a miss).  Returns the
    number of adjudicated rays."""
    import oracle
    g, f, s, lw = hits_u32(got), hits_u32(full), hits_u32(guarded), hits_u32(lower)
    # A NaN accepted anywhere in the scan poisons its answer (every later
    # pair is accepted): those rays are adjudicated too.
    clean = (f == s).all(axis=1) & (f == lw).all(axis=1)
    bad = np.nonzero(clean & (g != f).any(axis=1))[0]
    if bad.size:
        i = bad[0]
        raise AssertionError(f"{what}: {bad.size} rays differ from the scan; first ray {i}: got "
                             f"id={g[i, 0]:#x} t={g[i, 1:].view(np.float32)} want id={f[i, 0]:#x} "
                             f"t={f[i, 1:].view(np.float32)}")
    idx = np.nonzero(~clean)[0]
    (gt, gi), (ft, fi), (st, si) = _key(g[idx]), _key(lw[idx]), _key(s[idx])
    lo_ok = (gt > ft) | ((gt == ft) & (gi >= fi))
    hi_ok = (gt < st) | ((gt == st) & (gi <= si))
    assert (lo_ok & hi_ok).all(), f"{what}: adjudicated rays outside [scan without NaN, scan without artifacts]"
    for j in idx:
        if g[j, 0] == RT_MISS:
            continue
        tri = np.asarray(tris, np.float32)[g[j, 0]]
        ok, t, u, v = oracle.mt_f32(rays["org"][j], rays["dir"][j], tri[:3], tri[3:6], tri[6:])
        assert ok and np.float32([t, u, v]).view(np.uint32).tolist() == g[j, 1:].tolist(), \
            f"{what}: ray {j} reports a pair FP32 MT does not produce"
    return int(idx.size)


def assert_flags_adjudicated(got, full, guarded, what=""):
    """Occlusion flags: equal to the scan's wherever the scan without FP32
    artifacts agrees with it.  Returns the number of adjudicated rays."""
    got, full, guarded = (np.asarray(x).astype(np.uint8) for x in (got, full, guarded))
    clean = full == guarded
    bad = np.nonzero(clean & (got != full))[0]
    assert bad.size == 0, f"{what}: {bad.size} flags differ from the scan; first ray {bad[:1]}"
    return int((~clean).sum())

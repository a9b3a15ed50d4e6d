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

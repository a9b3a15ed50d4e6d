"""CPU checks of the drop-in boundary: the C ABI library loads and exports
every symbol include/rt/*.h declares, host-side error paths return codes
instead of crashing, built BVHs satisfy their structural invariants, and the
traversal restatement over those exact node bytes equals the reference's
linear scan."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
import igx_raytracing_amd as rt
from igx_raytracing_amd import RAY_DTYPE, RT_MISS, RtError
from igx_raytracing_amd import workload as wl
from igx_raytracing_amd._lib import SIGNATURES

from helpers import assert_hits_equal

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("rt_c.h", "rt_workload.h"):
        text = open(os.path.join(ROOT, "include", "rt", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(rtw?_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(rt.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python mirror declares a signature for each of them
    assert names <= set(SIGNATURES), sorted(names - set(SIGNATURES))


def test_host_error_paths_return_codes():
    lib = rt.load()
    tris = wl.c1_scene()
    with pytest.raises(RtError, match="node_width"):
        rt.HostBvh(tris, width=3)
    with pytest.raises(RtError, match="max_leaf"):
        rt.HostBvh(tris, max_leaf=9)
    h = ctypes.c_void_p()
    assert lib.rt_bvh_build(None, 5, None, ctypes.byref(h)) == -1
    assert b"null" in lib.rt_last_error_global()
    assert lib.rt_ctx_create(0, None) == -1
    assert lib.rt_trace_closest(None, None, None, 0, None, None) == -1


def test_context_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    rc = rt.load().rt_ctx_create(0, ctypes.byref(h))
    assert rc in (-3, -1) and not h.value


# ---- structural invariants --------------------------------------------------

def decode_nodes(width, nodes):
    """(child refs [N, W], child boxes lo/hi [N, W, 3] in float64)."""
    if width == 2:
        a = nodes.view(np.float32).reshape(-1, 16)
        lo = np.stack([a[:, 0:3], a[:, 6:9]], 1).astype(np.float64)
        hi = np.stack([a[:, 3:6], a[:, 9:12]], 1).astype(np.float64)
        refs = a[:, 12:14].view(np.uint32)
        return refs, lo, hi
    words = 16 if width == 4 else 32
    a = nodes.view(np.uint32).reshape(-1, words)
    origin = a[:, 0:3].view(np.float32).astype(np.float64)
    ex = np.stack([(a[:, 3] >> (8 * k)) & 0xFF for k in range(3)], 1).astype(np.int64)
    scale = np.ldexp(1.0, ex - 127)
    b = nodes.reshape(-1, words * 4)
    qlo = b[:, 16:16 + 3 * width].reshape(-1, 3, width).transpose(0, 2, 1).astype(np.float64)
    qhi = b[:, 16 + 3 * width:16 + 6 * width].reshape(-1, 3, width).transpose(0, 2, 1).astype(np.float64)
    lo = origin[:, None, :] + qlo * scale[:, None, :]
    hi = origin[:, None, :] + qhi * scale[:, None, :]
    refs = a[:, 4 + 6 * width // 4:4 + 6 * width // 4 + width]
    return refs, lo, hi


def check_bvh(bvh, tris, max_leaf):
    info = bvh.info
    n = tris.shape[0]
    W = info["node_width"]
    recs = bvh.tris
    ids = recs[:, 3].view(np.uint32)
    assert sorted(ids.tolist()) == list(range(n)), "every triangle exactly once"
    # stored edges are p1 - p0 and p2 - p0 rounded in FP32 (bit-exact)
    src = tris[ids]
    np.testing.assert_array_equal(recs[:, 0:3], src[:, 0:3])
    np.testing.assert_array_equal(recs[:, 4:7], (src[:, 3:6] - src[:, 0:3]).astype(np.float32))
    np.testing.assert_array_equal(recs[:, 8:11], (src[:, 6:9] - src[:, 0:3]).astype(np.float32))
    if info["num_nodes"] == 0:
        return
    refs, lo, hi = decode_nodes(W, bvh.nodes)
    seen_nodes = np.zeros(info["num_nodes"], np.int32)
    covered = np.zeros(n, np.int32)
    verts = src.reshape(-1, 3, 3).astype(np.float64)
    for i in range(info["num_nodes"]):
        for c in range(W):
            r = int(refs[i, c])
            if r == 0xFFFFFFFF:
                assert lo[i, c, 0] > hi[i, c, 0] or W == 2, "empty slot must have an empty box"
                continue
            if r & 0x80000000:
                first, cnt = r & 0x0FFFFFFF, ((r >> 28) & 7) + 1
                assert cnt <= max_leaf
                covered[first:first + cnt] += 1
                v = verts[first:first + cnt].reshape(-1, 3)
                assert (v >= lo[i, c]).all() and (v <= hi[i, c]).all(), "leaf box must contain its triangles"
            else:
                assert r < info["num_nodes"]
                seen_nodes[r] += 1
                if W == 2:  # float boxes: child's children inside the child's box
                    assert (lo[r] >= lo[i, c]).all() and (hi[r] <= hi[i, c]).all()
    assert (covered == 1).all(), "every record in exactly one leaf"
    assert seen_nodes[0] == 0 and (seen_nodes[1:] == 1).all(), "nodes form a tree"


@pytest.mark.parametrize("width", (2, 4, 8))
@pytest.mark.parametrize("max_leaf", (1, 4, 8))
def test_bvh_invariants_mixed_scene(width, max_leaf):
    tris = wl.c1_scene()
    b = rt.HostBvh(tris, width=width, max_leaf=max_leaf)
    check_bvh(b, tris, max_leaf)
    assert b.info["max_stack"] >= b.info["max_depth"] - 1


@pytest.mark.parametrize("width", (2, 4, 8))
def test_bvh_invariants_clusters_and_degenerate(width):
    tris = wl.clusters(wl.seed(9), 16, 2048)
    check_bvh(rt.HostBvh(tris, width=width), tris, 4)
    # coincident triangles (all centroids equal) still split into <= 4 per leaf
    same = np.tile(np.float32([0, 0, 0, 1, 0, 0, 0, 1, 0]), (37, 1))
    check_bvh(rt.HostBvh(same, width=width), same, 4)
    # one triangle: the root is a leaf
    b = rt.HostBvh(tris[:1], width=width)
    assert b.info["num_nodes"] == 0 and b.info["root_ref"] & 0x80000000
    e = rt.HostBvh(np.zeros((0, 9), np.float32), width=width)
    assert e.info["root_ref"] == 0xFFFFFFFF


# ---- traversal restatement == linear scan, on the CPU -------------------

@pytest.mark.parametrize("width", (2, 4, 8))
def test_walk_equals_linear_scan_mixed(width):
    tris = wl.c1_scene()
    rays = wl.c1_rays(128, 128)
    b = rt.HostBvh(tris, width=width)
    want = oracle.closest_brute(tris, rays)
    assert_hits_equal(oracle.walk(width, b.nodes, b.tris, b.info["root_ref"], rays), want, f"walk BVH{width}")
    occ = rays.copy()
    occ["tmax"] = np.random.default_rng(0).uniform(0, 40, occ.size).astype(np.float32)
    np.testing.assert_array_equal(oracle.walk(width, b.nodes, b.tris, b.info["root_ref"], occ, occlusion=True),
                                  oracle.occluded_brute(tris, occ))


@pytest.mark.parametrize("width", (2, 4, 8))
def test_walk_equals_linear_scan_clusters(width):
    tris = wl.clusters(wl.seed(1), 64, 1024)
    rays = wl.c2_rays(tris, 4096)
    occ = wl.c2_rays(tris, 4096, occlusion=True)
    b = rt.HostBvh(tris, width=width)
    assert_hits_equal(oracle.walk(width, b.nodes, b.tris, b.info["root_ref"], rays),
                      oracle.closest_brute(tris, rays), f"walk BVH{width}")
    np.testing.assert_array_equal(oracle.walk(width, b.nodes, b.tris, b.info["root_ref"], occ, occlusion=True),
                                  oracle.occluded_brute(tris, occ))


def test_walk_tie_rule_on_duplicates():
    """Stacked duplicates (same t on many ids): the lowest id wins in any node
    width, as in the scan; with prev_id the next lowest wins."""
    rng = np.random.default_rng(4)
    base = rng.normal(0, 1, (200, 9)).astype(np.float32)
    tris = np.concatenate([base, base[::-1], base])  # each triangle three times
    rays = np.zeros(3000, RAY_DTYPE)
    rays["org"] = rng.normal(0, 3, (3000, 3)).astype(np.float32)
    tgt = rng.normal(0, 0.5, (3000, 3))
    d = tgt - rays["org"]
    rays["dir"] = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    rays["tmax"] = np.float32(3.4028235e38)
    rays["prev_id"] = np.where(rng.random(3000) < 0.3, rng.integers(0, 600, 3000), RT_MISS).astype(np.uint32)
    want = oracle.closest_brute(tris, rays)
    assert (want[:, 0] != RT_MISS).mean() > 0.3
    for width in (2, 4, 8):
        b = rt.HostBvh(tris, width=width)
        assert_hits_equal(oracle.walk(width, b.nodes, b.tris, b.info["root_ref"], rays), want, f"dups BVH{width}")


def test_triangle_record_layout_is_std430():
    """TRIANGLE_DTYPE follows std430 for the reference's struct Triangle
    { vec3 p0; uint n0; vec3 p1; uint n1; vec3 p2; uint n2; }
    (primitive.glsl:28-38): vec3 aligns to 16 bytes and occupies 12, uint
    aligns to 4, and the struct's size rounds up to its largest alignment."""
    from igx_raytracing_amd import TRIANGLE_DTYPE
    members = [("p0", 16, 12), ("n0", 4, 4), ("p1", 16, 12), ("n1", 4, 4), ("p2", 16, 12), ("n2", 4, 4)]
    off, want = 0, {}
    for name, align, size in members:
        off = (off + align - 1) // align * align
        want[name] = off
        off += size
    size = (off + 15) // 16 * 16
    assert {k: TRIANGLE_DTYPE.fields[k][1] for k in want} == want
    assert TRIANGLE_DTYPE.itemsize == size == 48

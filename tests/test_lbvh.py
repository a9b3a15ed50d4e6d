"""GPU LBVH builder: the binary tree (Morton codes, sort order, Karras
children, boxes) is bit-exact against the oracle's CPU restatement, and
queries over LBVH scenes equal the reference's linear scan."""
import numpy as np
import pytest

import oracle
from igx_raytracing_amd import RT_MISS, Scene, lbvh_binary, trace_opts
from igx_raytracing_amd import workload as wl

from helpers import assert_hits_equal, hits_u32, to_dev


def check_binary_structure(morton, order, left, right, boxes, tris):
    n = tris.shape[0]
    assert sorted(order.tolist()) == list(range(n))
    m = morton.astype(np.int64)
    assert (np.diff(m) >= 0).all()
    same = np.diff(m) == 0
    assert (np.diff(order.astype(np.int64))[same] > 0).all(), "equal codes keep index order"
    if n == 1:
        return
    refs = np.concatenate([left, right])
    leaf = refs >= 0x80000000
    assert sorted((refs[leaf] & 0x7FFFFFFF).tolist()) == list(range(n))
    assert sorted(refs[~leaf].tolist()) == list(range(1, n - 1))
    # every internal box is the exact union of its children's boxes
    for i in range(n - 1):
        kids = [(n - 1 + (c & 0x7FFFFFFF)) if c >= 0x80000000 else c for c in (left[i], right[i])]
        np.testing.assert_array_equal(boxes[i, :3], np.minimum(boxes[kids[0], :3], boxes[kids[1], :3]))
        np.testing.assert_array_equal(boxes[i, 3:], np.maximum(boxes[kids[0], 3:], boxes[kids[1], 3:]))
    v = tris[order].reshape(n, 3, 3)
    np.testing.assert_array_equal(boxes[n - 1:, :3], v.min(1))
    np.testing.assert_array_equal(boxes[n - 1:, 3:], v.max(1))


@pytest.mark.parametrize("scene", ["mixed", "clusters", "coincident", "tiny"])
def test_oracle_lbvh_structure(scene):
    if scene == "mixed":
        tris = wl.c1_scene()
    elif scene == "clusters":
        tris = wl.clusters(wl.seed(5), 32, 512)
    elif scene == "coincident":
        tris = np.tile(np.float32([0, 0, 0, 1, 0, 0, 0, 1, 0]), (100, 1))
    else:
        tris = wl.c1_scene()[:3]
    check_binary_structure(*oracle.lbvh_binary(tris), tris)


def _scenes():
    dup = np.concatenate([wl.c1_scene()] * 3)  # equal Morton codes by construction
    return {"mixed4k": wl.c1_scene(), "dup12k": dup, "one": wl.c1_scene()[:1], "two": wl.c1_scene()[:2],
            "five": wl.c1_scene()[:5], "clusters1M": wl.c2_scene()}


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["one", "two", "five", "mixed4k", "dup12k", "clusters1M"])
def test_gpu_lbvh_binary_bitexact(gpu_ctx, name):
    import torch
    tris = _scenes()[name]
    print(f"lbvh binary {name} n={tris.shape[0]}", flush=True)
    d = torch.from_numpy(tris.copy()).cuda()
    got = [t.cpu().numpy() for t in lbvh_binary(gpu_ctx, d)]
    want = oracle.lbvh_binary(tris)
    for g, w, what in zip(got, want, ("morton", "order", "left", "right", "boxes")):
        np.testing.assert_array_equal(g.view(w.dtype).reshape(w.shape), w, err_msg=what)


@pytest.mark.gpu
@pytest.mark.parametrize("width", (2, 4, 8))
def test_lbvh_scene_queries_match_linear_scan(gpu_ctx, c2_data, width):
    import torch
    tris, rays, occ = c2_data
    n = 1 << 16
    d = torch.from_numpy(tris.copy()).cuda()
    sc = Scene(gpu_ctx, device_tris=d, width=width)
    info = sc.info
    assert info["num_nodes"] > 0 and info["max_stack"] > 0
    dr, do = to_dev(rays[:n]), to_dev(occ[:n])
    ref_h = sc.trace_closest_brute(dr)
    ref_o = sc.trace_occluded_brute(do)
    assert_hits_equal(sc.trace_closest(dr), ref_h, f"LBVH{width} closest")
    np.testing.assert_array_equal(sc.trace_occluded(do).cpu().numpy(), ref_o.cpu().numpy())
    # the oracle's traversal over the same node bytes: same results and counts
    nodes, recs = sc.download()
    counts = torch.empty((n, 2), dtype=torch.int32, device="cuda")
    got = sc.trace_closest(dr, opts=trace_opts(), counts=counts)
    want, wcnt = oracle.walk(width, nodes, recs, info["root_ref"], rays[:n], counts=True, nthreads=16)
    assert_hits_equal(got, want, "LBVH walk")
    np.testing.assert_array_equal(counts.cpu().numpy().view(np.uint32), wcnt)
    assert gpu_ctx.check_device_error() == 0
    sc.close()


@pytest.mark.gpu
def test_lbvh_small_and_mixed_scenes(gpu_ctx):
    import torch
    for name in ("one", "two", "five", "mixed4k", "dup12k"):
        tris = _scenes()[name]
        rays = wl.c1_rays(96, 96)
        want = oracle.closest_brute(tris, rays)
        for width in (2, 4, 8):
            sc = Scene(gpu_ctx, device_tris=torch.from_numpy(tris.copy()).cuda(), width=width, max_leaf=4)
            assert_hits_equal(sc.trace_closest(to_dev(rays)), want, f"LBVH{width} {name}")
            sc.close()
    # the host-pointer entry point with the GPU builder
    tris = wl.c1_scene()
    sc = Scene(gpu_ctx, tris, width=4, builder="lbvh")
    rays = wl.c1_rays(64, 64)
    assert_hits_equal(sc.trace_closest(to_dev(rays)), oracle.closest_brute(tris, rays), "LBVH host entry")
    sc.close()


@pytest.mark.gpu
def test_lbvh_refuses_infinite_coordinates(gpu_ctx):
    """The GPU builder reads the scene bounds back before it quantises any
    box, and refuses a scene with an infinite coordinate (a box with an
    infinite side has no quantised form); the context stays usable."""
    import torch
    from igx_raytracing_amd import RtError
    tris = wl.c1_scene().copy()
    tris[7, 2] = np.inf
    for width in (2, 4, 8):
        with pytest.raises(RtError, match="infinite coordinate"):
            Scene(gpu_ctx, device_tris=torch.from_numpy(tris.copy()).cuda(), width=width)
    ok = wl.c1_scene()
    sc = Scene(gpu_ctx, device_tris=torch.from_numpy(ok.copy()).cuda(), width=4)
    rays = wl.c1_rays(32, 32)
    assert_hits_equal(sc.trace_closest(to_dev(rays)), oracle.closest_brute(ok, rays), "after a refused build")
    sc.close()
    assert gpu_ctx.check_device_error() == 0


def _half_area(box):
    """Half surface area in float64 from a float32 box (lo xyz, hi xyz), in
    the builders' operation order: dx*dy + dy*dz + dz*dx."""
    d = box[3:].astype(np.float64) - box[:3].astype(np.float64)
    return d[0] * d[1] + d[1] * d[2] + d[2] * d[0]


def expected_leaves(left, right, boxes, n, max_leaf):
    """The (first, count) leaves of the final tree, from the binary tree: a
    subtree of at most max_leaf triangles is a leaf when the host SAH
    builder's leaf test accepts it (bvh_build.cpp:182-183: count <=
    1 + (A_l n_l + A_r n_r) / A), else its children are examined; binary
    leaves are one-triangle leaves.  A scene of at most max_leaf triangles
    is one leaf."""
    if n <= max_leaf:
        return {(0, n)}
    leaf = lambda c: c >= 0x80000000
    # sorted-index range of every internal node, children before parents
    first = np.zeros(n - 1, np.int64)
    last = np.zeros(n - 1, np.int64)
    order, stack = [], [0]
    while stack:
        i = stack.pop()
        order.append(i)
        stack.extend(int(c) for c in (left[i], right[i]) if not leaf(c))
    for i in reversed(order):
        lo_hi = []
        for c in (int(left[i]), int(right[i])):
            lo_hi.append((c & 0x7FFFFFFF, c & 0x7FFFFFFF) if leaf(c) else (first[c], last[c]))
        first[i], last[i] = min(lo_hi[0][0], lo_hi[1][0]), max(lo_hi[0][1], lo_hi[1][1])
    count = lambda c: 1 if leaf(c) else int(last[c] - first[c] + 1)
    box = lambda c: boxes[n - 1 + (c & 0x7FFFFFFF)] if leaf(c) else boxes[c]

    def is_leaf(c):
        if leaf(c):
            return True
        k = count(c)
        if k > max_leaf:
            return False
        a = _half_area(box(c))
        if not a > 0.0:
            return True
        l, r = int(left[c]), int(right[c])
        split = 1.0 + (_half_area(box(l)) * float(count(l)) + _half_area(box(r)) * float(count(r))) / a
        return float(k) <= split

    out, stack = set(), [0]
    while stack:
        i = stack.pop()
        for c in (int(left[i]), int(right[i])):
            if is_leaf(c):
                out.add((c & 0x7FFFFFFF, 1) if leaf(c) else (int(first[c]), count(c)))
            else:
                stack.append(c)
    return out


def emitted_leaves(nodes, width, root_ref):
    """(first, count) of every leaf reference in a scene's node bytes."""
    if root_ref >= 0x80000000:
        refs = np.array([root_ref], np.uint32)
    else:
        stride, off = (128, 64) if width == 8 else (64, 48 if width == 2 else 40)
        refs = nodes.reshape(-1, stride)[:, off:off + 4 * width].copy().view(np.uint32).reshape(-1)
    refs = refs[(refs & 0x80000000 != 0) & (refs != 0xFFFFFFFF)].astype(np.int64)
    got = list(zip((refs & 0x0FFFFFFF).tolist(), (((refs >> 28) & 7) + 1).tolist()))
    assert len(got) == len(set(got)), "a leaf is referenced twice"
    return set(got)


def test_expected_leaves_rule():
    """The leaf rule on hand-made binary trees: far-apart triangles split,
    a tight pair stays one leaf, degenerate boxes make leaves, and
    max_leaf bounds every leaf."""
    L = 0x80000000
    # 3 triangles: root 0 = (leaf 0, node 1); node 1 = (leaf 1, leaf 2)
    left, right = np.array([L | 0, L | 1], np.uint32), np.array([1, L | 2], np.uint32)

    def boxes_for(leaf_boxes):
        lb = np.float32(leaf_boxes)
        n1 = np.concatenate([np.minimum(lb[1, :3], lb[2, :3]), np.maximum(lb[1, 3:], lb[2, 3:])])
        root = np.concatenate([np.minimum(lb[0, :3], n1[:3]), np.maximum(lb[0, 3:], n1[3:])])
        return np.vstack([root, n1, lb]).astype(np.float32)

    far = boxes_for([[0, 0, 0, 1, 1, 1], [10, 0, 0, 11, 1, 1], [20, 0, 0, 21, 1, 1]])
    # node 1 spans [10,21]x[0,1]x[0,1]: A = 11 + 1 + 11 = 23; its unit-cube
    # children have A = 3 each: split = 1 + (3 + 3)/23 < 2 triangles -> split
    assert expected_leaves(left, right, far, 3, 2) == {(0, 1), (1, 1), (2, 1)}
    tight = boxes_for([[0, 0, 0, 1, 1, 1], [10, 0, 0, 11, 1, 1], [10, 0, 0, 11, 1, 1]])
    # both children fill node 1's box: split = 1 + (3 + 3)/3 = 3 >= 2 -> leaf
    assert expected_leaves(left, right, tight, 3, 2) == {(0, 1), (1, 2)}
    assert expected_leaves(left, right, tight, 3, 1) == {(0, 1), (1, 1), (2, 1)}
    flat = boxes_for([[0, 0, 0, 1, 1, 1], [5, 5, 5, 5, 5, 5], [5, 5, 5, 5, 5, 5]])
    assert expected_leaves(left, right, flat, 3, 2) == {(0, 1), (1, 2)}  # zero area: one leaf
    assert expected_leaves(left, right, far, 3, 3) == {(0, 3)}  # the whole scene fits one leaf


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["five", "mixed4k", "dup12k", "clusters16k"])
def test_lbvh_leaves_follow_the_sah_leaf_test(gpu_ctx, name):
    """The emitted tree's leaves are exactly the subtrees the leaf test
    selects on the oracle's binary tree, for every node width."""
    import torch
    tris = wl.clusters(wl.seed(5), 32, 512) if name == "clusters16k" else _scenes()[name]
    n = tris.shape[0]
    _, _, left, right, boxes = oracle.lbvh_binary(tris)
    for max_leaf in (1, 4, 8):
        want = expected_leaves(left, right, boxes.reshape(-1, 6), n, max_leaf)
        assert sum(c for _, c in want) == n
        for width in (2, 4, 8):
            print(f"lbvh leaves {name} max_leaf={max_leaf} width={width}", flush=True)
            sc = Scene(gpu_ctx, device_tris=torch.from_numpy(tris.copy()).cuda(), width=width, max_leaf=max_leaf)
            nodes, _ = sc.download()
            assert emitted_leaves(nodes, width, sc.info["root_ref"]) == want
            sc.close()

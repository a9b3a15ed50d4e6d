"""Multi-scene batches and the autotuner: one persistent launch over many
scenes and uneven ray batches gives the same bits as one plain launch per
query; tuned launch shapes give the same bits as the defaults."""
import numpy as np
import pytest

from igx_raytracing_amd import Batch, Scene, autotune, trace_opts
from igx_raytracing_amd import workload as wl

from helpers import assert_hits_equal, hits_u32, to_dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("width", (2, 4, 8))
def test_batch_equals_plain_launches(gpu_ctx, width):
    import torch
    rng = np.random.default_rng(7)
    sizes = [1, 5, 3000, 20000, 150000]
    scenes, tris_list = [], []
    for k, nt in enumerate(sizes):
        tris = wl.clusters(wl.seed(40 + k), max(1, nt // 1000), min(nt, 1000)) if nt >= 1000 else wl.c1_scene()[:nt]
        tris_list.append(tris)
        scenes.append(Scene(gpu_ctx, device_tris=torch.from_numpy(tris.copy()).cuda(), width=width,
                            keep_source=False))
    batch = Batch(gpu_ctx, scenes)
    queries, plain = [], []
    for qi in range(24):
        si = int(rng.integers(0, len(scenes)))
        n = int(rng.choice([0, 1, 63, 64, 65, 777, 5000, 40000]))
        kind = "closest" if qi % 2 == 0 else "occluded"
        lo, hi = wl.bounds(tris_list[si])
        rays = wl.rays_uniform(wl.seed(60 + qi), lo - 1, hi + 1, max(n, 1), tmax=None if kind == "closest" else (0, 5))[:n]
        d = to_dev(rays) if n else torch.empty((0, 8), dtype=torch.float32, device="cuda")
        if kind == "closest":
            out = torch.empty((n, 4), dtype=torch.int32, device="cuda")
            ref = scenes[si].trace_closest(d) if n else out
        else:
            out = torch.empty(n, dtype=torch.uint8, device="cuda")
            ref = scenes[si].trace_occluded(d) if n else out
        queries.append((si, kind, d, out))
        plain.append(ref)
    for lds in (8, 16, 32):
        for block in (128, 256):
            for q in queries:
                q[3].fill_(0x55)
            batch.trace(queries, opts=trace_opts(lds_stack=lds, block_size=block))
            torch.cuda.synchronize()
            for (si, kind, d, out), ref in zip(queries, plain):
                if kind == "closest":
                    assert_hits_equal(out, ref, f"batch BVH{width} scene {si} n={d.shape[0]}")
                else:
                    np.testing.assert_array_equal(out.cpu().numpy(), ref.cpu().numpy())
    assert gpu_ctx.check_device_error() == 0
    batch.close()
    for s in scenes:
        s.close()


def test_autotune_never_changes_results(gpu_ctx, c2_data):
    tris, rays, occ = c2_data
    n = 1 << 17
    sc = Scene(gpu_ctx, tris, width=4, keep_source=False)
    d, do = to_dev(rays[:n]), to_dev(occ[:n])
    best, ms = autotune(sc, d, "closest", iters=2)
    assert ms > 0 and best.block_size in (128, 256) and best.lds_stack in (8, 16, 32)
    assert_hits_equal(sc.trace_closest(d, opts=best), sc.trace_closest(d, opts=trace_opts(schedule="static")),
                      "tuned closest")
    best_o, _ = autotune(sc, do, "occluded", iters=2)
    np.testing.assert_array_equal(sc.trace_occluded(do, opts=best_o).cpu().numpy(),
                                  sc.trace_occluded(do, opts=trace_opts(schedule="static")).cpu().numpy())
    # every 128-thread shape matches too
    ref = hits_u32(sc.trace_closest(d))
    for lds in (8, 16, 32):
        for sched in ("static", "dynamic"):
            assert_hits_equal(sc.trace_closest(d, opts=trace_opts(lds_stack=lds, block_size=128, schedule=sched)),
                              ref, f"block 128 lds {lds} {sched}")
    sc.close()


def test_back_to_back_batches_with_different_tables(gpu_ctx):
    """Six calls with six different query tables and no synchronisation in
    between: each call's table is staged through one of two pinned slots
    while earlier calls may still be queued, and two calls outgrow the
    table.  Every call's outputs equal plain launches of its queries."""
    import torch
    tris_list = [wl.clusters(wl.seed(90 + k), 32 * (k + 1), 1000) for k in range(3)]
    scenes = [Scene(gpu_ctx, device_tris=torch.from_numpy(t.copy()).cuda(), width=4, keep_source=False)
              for t in tris_list]
    batch = Batch(gpu_ctx, scenes)
    rng = np.random.default_rng(21)
    calls = []
    for c, nq in enumerate((4, 8, 70, 8, 4, 90)):
        queries, plain = [], []
        for qi in range(nq):
            si = int(rng.integers(0, len(scenes)))
            n = int(rng.integers(1, 30000))
            kind = "closest" if rng.integers(0, 2) == 0 else "occluded"
            lo, hi = wl.bounds(tris_list[si])
            rays = wl.rays_uniform(wl.seed(1000 + 100 * c + qi), lo, hi, n,
                                   tmax=None if kind == "closest" else (0, 5))
            d = to_dev(rays)
            if kind == "closest":
                out = torch.full((n, 4), 0x55, dtype=torch.int32, device="cuda")
                ref = scenes[si].trace_closest(d)
            else:
                out = torch.full((n,), 0x55, dtype=torch.uint8, device="cuda")
                ref = scenes[si].trace_occluded(d)
            queries.append((si, kind, d, out))
            plain.append(ref)
        calls.append((queries, plain))
    torch.cuda.synchronize()
    for queries, _ in calls:
        batch.trace(queries)
    torch.cuda.synchronize()
    for c, (queries, plain) in enumerate(calls):
        for (si, kind, d, out), ref in zip(queries, plain):
            if kind == "closest":
                assert_hits_equal(out, ref, f"call {c} scene {si} n={d.shape[0]}")
            else:
                np.testing.assert_array_equal(out.cpu().numpy(), ref.cpu().numpy(), err_msg=f"call {c}")
    assert gpu_ctx.check_device_error() == 0
    batch.close()
    for s in scenes:
        s.close()

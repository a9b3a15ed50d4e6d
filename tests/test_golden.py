"""Golden vectors: the PCG32 generator against its published reference
output, the workload generators and the oracle against the committed
fixtures (tests/golden/make_fixtures.py), and the GPU kernels against the
same fixtures."""
import hashlib
import os

import numpy as np
import pytest

import oracle
from igx_raytracing_amd import Scene
from igx_raytracing_amd import workload as wl

from helpers import assert_hits_equal, to_dev

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture(name):
    return np.load(os.path.join(HERE, name), allow_pickle=False)


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_pcg32_published_vector():
    # pcg32-demo (O'Neill): pcg32_srandom_r(42, 54) then pcg32_random_r.
    want = [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]
    assert [wl.pcg32_nth(42, 54, k) for k in range(6)] == want


def test_generators_reproduce_fixture_inputs():
    f = fixture("mixed4k.npz")
    tris, rays = wl.c1_scene(), wl.c1_rays()
    assert sha(tris) == str(f["scene_sha"]) and sha(rays) == str(f["rays_sha"])
    np.testing.assert_array_equal(rays[f["ray_index"]].view(np.uint32), f["rays"].view(np.uint32))
    g = fixture("clusters1m.npz")
    t2 = wl.c2_scene()
    assert sha(t2) == str(g["scene_sha"])
    np.testing.assert_array_equal(t2[:64], g["first_tris"])
    np.testing.assert_array_equal(wl.c2_rays(t2, 512).view(np.uint32), g["rays"].view(np.uint32))


def test_oracle_reproduces_fixture_outputs():
    f = fixture("mixed4k.npz")
    tris = wl.c1_scene()
    np.testing.assert_array_equal(oracle.closest_brute(tris, f["rays"]), f["hits"])
    np.testing.assert_array_equal(oracle.occluded_brute(tris, f["shadow_rays"]), f["shadow_flags"])
    n = fixture("niels3.npz")
    np.testing.assert_array_equal(oracle.closest_brute(n["tris"], n["rays"]), n["hits"])


@pytest.mark.gpu
@pytest.mark.parametrize("width", (2, 4, 8))
@pytest.mark.parametrize("builder", ("sah", "lbvh"))
def test_gpu_matches_fixtures(gpu_ctx, width, builder):
    f = fixture("mixed4k.npz")
    tris = wl.c1_scene()
    sc = Scene(gpu_ctx, tris, width=width, builder=builder)
    assert_hits_equal(sc.trace_closest(to_dev(f["rays"])), f["hits"], f"mixed4k {builder} BVH{width}")
    np.testing.assert_array_equal(sc.trace_occluded(to_dev(f["shadow_rays"])).cpu().numpy(), f["shadow_flags"])
    sc.close()
    g = fixture("clusters1m.npz")
    sc = Scene(gpu_ctx, wl.c2_scene(), width=width, builder=builder, keep_source=False)
    assert_hits_equal(sc.trace_closest(to_dev(g["rays"])), g["hits"], f"clusters1m {builder} BVH{width}")
    np.testing.assert_array_equal(sc.trace_occluded(to_dev(g["occ_rays"])).cpu().numpy(), g["occ_flags"])
    sc.close()
    n = fixture("niels3.npz")
    sc = Scene(gpu_ctx, n["tris"], width=width, builder=builder)
    assert_hits_equal(sc.trace_closest(to_dev(n["rays"])), n["hits"], "niels3")
    sc.close()

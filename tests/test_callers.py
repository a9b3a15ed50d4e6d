"""The callers of the query path: raygen.comp's packing of closest hits into
dirObject / uvNormal (with shading normals) and nv_all.shadow.comp's wave64
ballots in 16x4 pixel tiles, against the oracle's restatement; the tile
words are read back with the reference's own decoder rule (didHit,
light_rt.glsl:30-39, indexToLight light.glsl:221-229)."""
import numpy as np
import pytest

import oracle
from igx_raytracing_amd import RAY_DTYPE, RT_MISS, Scene, screen_camera
from igx_raytracing_amd import workload as wl

from helpers import to_dev

pytestmark = pytest.mark.gpu
FILL = np.float32(-7.0)


def random_spheremap_normals(n, seed):
    """3 packed half2 spheremap normals per triangle, |nn| < 0.7."""
    rng = np.random.default_rng(seed)
    h = rng.uniform(-0.7, 0.7, (n, 3, 2)).astype(np.float16).view(np.uint16).astype(np.uint32)
    return (h[..., 0] | (h[..., 1] << 16)).reshape(-1)


def cam_array(cam, jitter):
    return np.concatenate([np.float32(x) for x in cam] + [np.float32(jitter)])


@pytest.mark.parametrize("width", (2, 4, 8))
@pytest.mark.parametrize("with_normals", (False, True))
def test_raygen_packing_matches_oracle(gpu_ctx, width, with_normals):
    import torch
    tris = wl.c1_scene()
    W, H = 100, 70  # edge tiles are partial
    cam = screen_camera((0.0, 5.0, -25.0), (0.0, 0.0, 0.0), 60.0, aspect=W / H)
    normals = random_spheremap_normals(tris.shape[0], 3) if with_normals else None
    sc = Scene(gpu_ctx, tris, width=width)
    for jitter in ((0.5, 0.5), (0.1, 0.9)):
        d_norm = torch.from_numpy(normals.view(np.int32)).cuda() if with_normals else None
        uvn = torch.full((H, W, 4), float(FILL), dtype=torch.float32, device="cuda")
        dob, uvn = sc.raygen_closest(cam, W, H, normals=d_norm, uv_normal=uvn, jitter=jitter)
        want_d, want_u = oracle.raygen(tris, cam_array(cam, jitter), W, H, normals=normals, fill=FILL)
        np.testing.assert_array_equal(dob.cpu().numpy().view(np.uint32), want_d.view(np.uint32))
        np.testing.assert_array_equal(uvn.cpu().numpy().view(np.uint32), want_u.view(np.uint32))
    hit = want_d[..., 3].view(np.uint32) != RT_MISS
    assert 0.2 < hit.mean() < 0.9
    sc.close()


@pytest.mark.parametrize("projection", ["screen", "omni", "omni_stereo_tb", "stereo_tb", "omni_stereo_lr",
                                        "stereo_lr"])
def test_raygen_primary_projections_match_oracle(gpu_ctx, projection):
    """calculatePrimary in full (camera.glsl:49-138): every projection, with
    the reference's per-pixel rand(loc + seed) offset and with a fixed one.
    The offset and the omni directions go through sin/cos, evaluated on both
    sides as the FP32 rounding of FP64 sin/cos; a double-rounding tie could
    make one pixel differ, so at most 2 pixels per image may (none expected)."""
    import torch
    from igx_raytracing_amd import primary_camera
    from igx_raytracing_amd._lib import PROJECTIONS
    tris = wl.c1_scene()
    W, H = 96, 64
    eye, p0, p1, p2 = screen_camera((0.0, 5.0, -25.0), (0.0, 0.0, 0.0), 60.0, aspect=W / H)
    _, p3, p4, p5 = screen_camera((1.5, 5.0, -25.0), (1.5, 0.0, 0.0), 60.0, aspect=W / H)
    normals = random_spheremap_normals(tris.shape[0], 4)
    d_norm = torch.from_numpy(normals.view(np.int32)).cuda()
    sc = Scene(gpu_ctx, tris, width=4)
    for jitter, seed in ((None, (0.3125, 0.71875)), (None, (17.25, 3.5)), ((0.5, 0.5), (0.0, 0.0))):
        c = primary_camera(eye, p0, p1, p2, p3, p4, p5, ipd=64.0, projection=projection, seed=seed,
                           jitter=jitter)
        uvn = torch.full((H, W, 4), float(FILL), dtype=torch.float32, device="cuda")
        dob, uvn = sc.raygen_primary(c, W, H, normals=d_norm, uv_normal=uvn)
        cam = {"eye": eye, "p0": p0, "p1": p1, "p2": p2, "p3": p3, "p4": p4, "p5": p5, "ipd": 64.0,
               "projection": PROJECTIONS[projection], "seed": seed, "jitter": jitter}
        want_d, want_u = oracle.raygen_primary(tris, cam, W, H, normals=normals, fill=FILL, nthreads=16)
        got_d, got_u = dob.cpu().numpy(), uvn.cpu().numpy()
        bad = ((got_d.view(np.uint32) != want_d.view(np.uint32)).any(-1) |
               (got_u.view(np.uint32) != want_u.view(np.uint32)).any(-1))
        print(f"{projection} jitter={jitter} seed={seed}: {int(bad.sum())} pixels differ, "
              f"hits {np.mean(want_d[..., 3].view(np.uint32) != RT_MISS):.3f}", flush=True)
        assert bad.sum() <= 2
    assert gpu_ctx.check_device_error() == 0
    sc.close()


def did_hit(words, x, y, sample, W, H):
    """light_rt.glsl:30-39 (64-bit tiles; the 64-bit shift the writer uses)."""
    tiles_x = (W >> 4) + (1 if W & 15 else 0)
    tiles_y = (H >> 2) + (1 if H & 3 else 0)
    idx = (x >> 4) + (y >> 2) * tiles_x + sample * tiles_y * tiles_x
    bit = (x & 15) | ((y & 3) << 4)
    return (int(words[idx]) >> bit) & 1


@pytest.mark.parametrize("width", (2, 4, 8))
def test_shadow_tiles_layout(gpu_ctx, width):
    tris = wl.c1_scene()
    W, H, S = 50, 30, 3
    rays = wl.c1_rays(W, H)
    prim = oracle.closest_brute(tris, rays)
    rng = np.random.default_rng(11)
    all_rays = np.zeros((S, H * W), RAY_DTYPE)
    for s in range(S):
        light = np.float32([rng.uniform(-10, 10), 15.0, rng.uniform(-15, 0)])
        hit = prim[:, 0] != RT_MISS
        t = prim[:, 1].view(np.float32)
        p = (rays["org"] + t[:, None] * rays["dir"]).astype(np.float32)
        lv = (light[None, :] - p).astype(np.float32)
        dist = np.sqrt((lv.astype(np.float64) ** 2).sum(1)).astype(np.float32)
        r = all_rays[s]
        r["org"] = np.where(hit[:, None], p, rays["org"])
        r["dir"] = np.where(hit[:, None], lv / dist[:, None], rays["dir"]).astype(np.float32)
        r["tmax"] = np.where(hit, dist, np.float32(0.0))  # no primary hit: nothing to shade (nv_all.shadow.comp:108)
        r["prev_id"] = np.where(hit, prim[:, 0], RT_MISS)
    flat = all_rays.reshape(-1)
    want = oracle.occluded_brute(tris, flat).reshape(S, H, W)
    assert 0.0 < want.mean() < 1.0
    sc = Scene(gpu_ctx, tris, width=width)
    words = sc.shadow_tiles(to_dev(flat), W, H, S).cpu().numpy().view(np.uint64)
    assert words.size == ((W + 15) // 16) * ((H + 3) // 4) * S
    for s in range(S):
        for y in range(H):
            for x in range(W):
                assert did_hit(words, x, y, s, W, H) == want[s, y, x], (s, x, y)
    # bits of pixels outside the image are zero
    tiles_x = (W + 15) // 16
    for ty in range((H + 3) // 4):
        for tx in range(tiles_x):
            for bit in range(64):
                x, y = tx * 16 + (bit & 15), ty * 4 + (bit >> 4)
                if x >= W or y >= H:
                    for s in range(S):
                        assert (int(words[tx + ty * tiles_x + s * tiles_x * ((H + 3) // 4)]) >> bit) & 1 == 0
    sc.close()


@pytest.mark.parametrize("width", (2, 4))
def test_packed_triangle_scene(gpu_ctx, width):
    """A scene made from the reference's 48-byte Triangle records keeps their
    normals: raygen without a normals argument equals raygen on the positions
    with the same normals passed explicitly, and the oracle."""
    import torch
    from igx_raytracing_amd import TRIANGLE_DTYPE
    tris = wl.c1_scene()
    normals = random_spheremap_normals(tris.shape[0], 5).reshape(-1, 3)
    packed = np.zeros(tris.shape[0], TRIANGLE_DTYPE)
    for v in range(3):
        packed[f"p{v}"] = tris[:, 3 * v:3 * v + 3]
        packed[f"n{v}"] = normals[:, v]
    W, H = 64, 40
    cam = screen_camera((0.0, 5.0, -25.0), (0.0, 0.0, 0.0), 60.0, aspect=W / H)
    sp = Scene(gpu_ctx, packed=packed, width=width)
    sx = Scene(gpu_ctx, tris, width=width)
    assert sp.info["num_tris"] == tris.shape[0]
    d_norm = torch.from_numpy(normals.reshape(-1).view(np.int32)).cuda()
    outs = []
    for sc, nrm in ((sp, None), (sx, d_norm)):
        uvn = torch.full((H, W, 4), float(FILL), dtype=torch.float32, device="cuda")
        dob, uvn = sc.raygen_closest(cam, W, H, normals=nrm, uv_normal=uvn)
        outs.append((dob.cpu().numpy().view(np.uint32), uvn.cpu().numpy().view(np.uint32)))
    want_d, want_u = oracle.raygen(tris, cam_array(cam, (0.5, 0.5)), W, H, normals=normals.reshape(-1), fill=FILL)
    for dob, uvn in outs:
        np.testing.assert_array_equal(dob, want_d.view(np.uint32))
        np.testing.assert_array_equal(uvn, want_u.view(np.uint32))
    rays = to_dev(wl.c1_rays(64, 64))
    np.testing.assert_array_equal(sp.trace_closest(rays).cpu().numpy(), sx.trace_closest(rays).cpu().numpy())
    sp.close()
    sx.close()


@pytest.mark.parametrize("light_kind", ["point", "directional"])
def test_shadow_pass_matches_oracle(gpu_ctx, light_kind):
    """nv_all.shadow.comp in full after raygen: the GPU's own dirObject
    image, a point light (sphere of radius 0.75 inside a [0.75, 60) window)
    or a directional light with an angular radius, 3 samples per pixel,
    analytic objects included.  Every tile word equals the oracle's built
    from the same dirObject; light samples go through sin/cos/acos, so at
    most 2 pixel bits per image may differ on a double-rounding tie (none
    expected)."""
    from igx_raytracing_amd import primary_camera
    tris, ana = wl.c1_scene(), wl.analytic_mix(8, nspheres=8, ncubes=4)
    W, H, S = 90, 62, 3
    eye, p0, p1, p2 = screen_camera((0.0, 5.0, -25.0), (0.0, 0.0, 0.0), 60.0, aspect=W / H)
    sc = Scene(gpu_ctx, tris, width=4)
    sc.set_analytic(**ana)
    dob, _ = sc.raygen_primary(primary_camera(eye, p0, p1, p2, seed=(0.375, 0.125)), W, H)

    def h16(x):
        return int(np.float16(x).view(np.uint16))
    if light_kind == "point":
        light = {"pos": (6.0, 15.0, -12.0), "rad_origin": h16(60.0) | (h16(0.75) << 16), "dir": (0, 0),
                 "color_type": (0, 2 << 16)}
    else:
        n = np.array([0.3, 1.0, -0.4]) / np.linalg.norm([0.3, 1.0, -0.4])
        hh = [int((c * 0.5 + 0.5) * 65535) for c in n]
        light = {"pos": (0.0, 0.0, 0.0), "rad_origin": h16(0.05), "dir": ((hh[0] << 16) | hh[1], hh[2]),
                 "color_type": (0, 0)}
    seed = (0.6875, 0.1875)
    got = sc.shadow_pass(dob, eye, light, seed, W, H, S).cpu().numpy().view(np.uint64)
    want = oracle.shadow_pass(tris, dob.cpu().numpy(), eye, light, seed, W, H, S, analytic=ana, nthreads=16)
    diff = int(sum(bin(int(x)).count("1") for x in (got ^ want)))
    bits = int(sum(bin(int(x)).count("1") for x in want))
    print(f"shadow pass {light_kind}: {bits} occluded pixel-samples, {diff} bits differ", flush=True)
    assert diff <= 2 and bits > 100
    assert gpu_ctx.check_device_error() == 0
    sc.close()

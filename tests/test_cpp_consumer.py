"""The boundary from C++: include/rt/ray_query.hpp and rt_c.h compile in a
plain g++ translation unit, link against librt_hip.so, and the host builder
runs without a device."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_cpp_consumer_builds_and_runs(tmp_path):
    lib_dir = os.path.join(ROOT, "igx_raytracing_amd")
    exe = tmp_path / "consumer"
    cmd = ["g++", "-std=c++17", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__", f"-I{ROOT}/include",
           "-I/opt/rocm/include", os.path.join(ROOT, "tests", "cpp", "host_consumer.cpp"), f"-L{lib_dir}",
           "-lrt_hip", f"-Wl,-rpath,{lib_dir}", "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    assert out.startswith("nodes ") and "depth" in out


@pytest.mark.gpu
@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_cpp_mirror_drives_raygen_on_the_gpu(tmp_path):
    """tests/cpp/gpu_consumer.cpp, built with g++ against the C++ mirror,
    makes a scene from the reference's 48-byte Triangle records with their
    spheremap normals, adds the demo scene's analytic objects, and runs
    RaygenTask; both images equal the oracle's raygen bit for bit."""
    import numpy as np

    import oracle
    from igx_raytracing_amd import TRIANGLE_DTYPE, screen_camera
    from igx_raytracing_amd import workload as wl

    lib_dir = os.path.join(ROOT, "igx_raytracing_amd")
    exe = tmp_path / "gpu_consumer"
    cmd = ["g++", "-std=c++17", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__", f"-I{ROOT}/include",
           "-I/opt/rocm/include", os.path.join(ROOT, "tests", "cpp", "gpu_consumer.cpp"), f"-L{lib_dir}",
           "-lrt_hip", "-L/opt/rocm/lib", "-lamdhip64", f"-Wl,-rpath,{lib_dir}", "-Wl,-rpath,/opt/rocm/lib",
           "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True)
    tris, ana = wl.niels_scene()
    extra = wl.mixed4k(wl.seed(0))[:500]  # more triangles behind the demo objects
    tris = np.concatenate([tris, extra])
    rng = np.random.default_rng(5)
    h = rng.uniform(-0.7, 0.7, (tris.shape[0], 3, 2)).astype(np.float16).view(np.uint16).astype(np.uint32)
    normals = (h[..., 0] | (h[..., 1] << 16))
    packed = np.zeros(tris.shape[0], TRIANGLE_DTYPE)
    for v in range(3):
        packed[f"p{v}"] = tris[:, 3 * v:3 * v + 3]
        packed[f"n{v}"] = normals[:, v]
    W, H = 96, 64
    cam = screen_camera((0.0, 5.0, -9.0), (0.0, 4.0, 0.0), 60.0, aspect=W / H)
    cam14 = np.concatenate([np.float32(x) for x in cam] + [np.float32((0.5, 0.5))]).astype(np.float32)
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(np.array([tris.shape[0], len(ana["spheres"]), len(ana["cubes"]), len(ana["planes"]), W, H],
                         np.uint32).tobytes())
        f.write(packed.tobytes())
        for key in ("spheres", "cubes", "planes"):
            f.write(np.ascontiguousarray(ana[key], np.float32).tobytes())
        f.write(cam14.tobytes())
    outp = tmp_path / "out.bin"
    res = subprocess.run([str(exe), str(inp), str(outp)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    got = np.fromfile(outp, np.float32).reshape(2, H, W, 4)
    want_d, want_u = oracle.raygen(tris, cam14, W, H, normals=normals.reshape(-1), fill=np.float32(0.0),
                                   analytic=ana)
    np.testing.assert_array_equal(got[0].view(np.uint32), want_d.view(np.uint32))
    np.testing.assert_array_equal(got[1].view(np.uint32), want_u.view(np.uint32))
    ids = want_d[..., 3].view(np.uint32)
    assert (ids < tris.shape[0]).sum() > 50 and ((ids >= tris.shape[0]) & (ids != 0xFFFFFFFF)).sum() > 500


def _shadow_rays(tris, W, H, S, seed):
    """Shadow rays [S][H][W] from the C1 camera's primary hits toward S
    point lights, as the reference's shadow pass aims them (origin = hit,
    tmax = distance, prev = the hit triangle; no hit: tmax 0)."""
    import numpy as np

    import oracle
    from igx_raytracing_amd import RAY_DTYPE, RT_MISS
    from igx_raytracing_amd import workload as wl
    rays = wl.c1_rays(W, H)
    prim = oracle.closest_brute(tris, rays)
    hit = prim[:, 0] != RT_MISS
    t = prim[:, 1].view(np.float32)
    p = (rays["org"] + t[:, None] * rays["dir"]).astype(np.float32)
    rng = np.random.default_rng(seed)
    out = np.zeros((S, H * W), RAY_DTYPE)
    for s in range(S):
        light = np.float32([rng.uniform(-10, 10), 15.0, rng.uniform(-15, 0)])
        lv = (light[None, :] - p).astype(np.float32)
        dist = np.sqrt((lv.astype(np.float64) ** 2).sum(1)).astype(np.float32)
        r = out[s]
        r["org"] = np.where(hit[:, None], p, rays["org"])
        r["dir"] = np.where(hit[:, None], lv / dist[:, None], rays["dir"]).astype(np.float32)
        r["tmax"] = np.where(hit, dist, np.float32(0.0))
        r["prev_id"] = np.where(hit, prim[:, 0], RT_MISS)
    return out.reshape(-1)


@pytest.mark.gpu
@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_cpp_shadow_task_on_the_gpu(tmp_path):
    """tests/cpp/shadow_consumer.cpp drives the C++ ShadowTask from g++:
    switchToScene, resize at two sizes, one pass each.  Its ballot buffer
    holds ceil(W/16) * ceil(H/4) * samples * 8 bytes (shadow_task.cpp:
    146-165), and every pixel's bit, read with the reference's didHit rule
    (light_rt.glsl:30-39, indexToLight light.glsl:221-229), equals the
    oracle's occlusion of that pixel's ray; bits outside the image are 0."""
    import numpy as np

    import oracle
    from igx_raytracing_amd import workload as wl

    lib_dir = os.path.join(ROOT, "igx_raytracing_amd")
    exe = tmp_path / "shadow_consumer"
    cmd = ["g++", "-std=c++17", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__", f"-I{ROOT}/include",
           "-I/opt/rocm/include", os.path.join(ROOT, "tests", "cpp", "shadow_consumer.cpp"), f"-L{lib_dir}",
           "-lrt_hip", "-L/opt/rocm/lib", "-lamdhip64", f"-Wl,-rpath,{lib_dir}", "-Wl,-rpath,/opt/rocm/lib",
           "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True)
    tris = wl.c1_scene()
    S = 2
    sizes = [(50, 30), (33, 9)]
    ray_sets = [_shadow_rays(tris, W, H, S, 11 + k) for k, (W, H) in enumerate(sizes)]
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(np.array([tris.shape[0], S, len(sizes)], np.uint32).tobytes())
        f.write(np.array([v for wh in sizes for v in wh], np.uint32).tobytes())
        f.write(np.ascontiguousarray(tris, np.float32).tobytes())
        for r in ray_sets:
            f.write(np.ascontiguousarray(r).tobytes())
    outp = tmp_path / "out.bin"
    res = subprocess.run([str(exe), str(inp), str(outp)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    blob = outp.read_bytes()
    pos = 0
    for (W, H), rays in zip(sizes, ray_sets):
        nbytes = int(np.frombuffer(blob, np.uint64, 1, pos)[0])
        pos += 8
        assert nbytes == -(-W // 16) * -(-H // 4) * S * 8
        words = np.frombuffer(blob, np.uint64, nbytes // 8, pos)
        pos += nbytes
        want = oracle.occluded_brute(tris, rays).reshape(S, H, W)
        assert 0 < want.mean() < 1
        tiles_x, tiles_y = (W >> 4) + (1 if W & 15 else 0), (H >> 2) + (1 if H & 3 else 0)
        for s in range(S):
            for ty in range(tiles_y):
                for tx in range(tiles_x):
                    word = int(words[tx + ty * tiles_x + s * tiles_y * tiles_x])
                    for bit in range(64):
                        x, y = tx * 16 + (bit & 15), ty * 4 + (bit >> 4)
                        expect = int(want[s, y, x]) if (x < W and y < H) else 0
                        assert (word >> bit) & 1 == expect, (W, H, s, x, y)
    assert pos == len(blob)

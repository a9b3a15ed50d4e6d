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

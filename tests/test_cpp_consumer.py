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

"""The host BVH builders (bvh_build.cpp, no device code) under
AddressSanitizer and UndefinedBehaviorSanitizer, built with g++."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_host_builders_under_sanitizers(tmp_path):
    """tests/cpp/builder_sanitize.cpp links the host builders' source
    (bvh_build.cpp, no device code) with AddressSanitizer and
    UndefinedBehaviorSanitizer, float-to-int overflow included, and builds
    small, long, NaN-bearing and degenerate soups at every width and
    splitting budget; any report aborts it."""
    exe = tmp_path / "builder_sanitize"
    src = os.path.join(ROOT, "igx_raytracing_amd", "csrc")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined,float-cast-overflow",
           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", f"-I{ROOT}/include", f"-I{src}",
           os.path.join(ROOT, "tests", "cpp", "builder_sanitize.cpp"), os.path.join(src, "bvh_build.cpp"),
           "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([str(exe)], capture_output=True, text=True, env=env)
    assert run.returncode == 0, run.stdout[-2000:] + run.stderr[-4000:]
    assert run.stdout.strip().endswith("ok")
    assert run.stdout.count("records") == 48
    assert "runtime error" not in run.stderr

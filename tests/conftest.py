import os
import sys

import pytest

try:  # import torch once at collection: the first import on a fresh box is slow
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on cuda:0)")


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch
    import igx_raytracing_amd as rt
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    ctx = rt.Context(0)
    yield ctx
    torch.cuda.synchronize()
    ctx.close()


@pytest.fixture(scope="session")
def c2_data():
    """1,048,576-triangle clustered scene and its incoherent rays (closest and occlusion)."""
    from igx_raytracing_amd import workload as wl
    tris = wl.c2_scene()
    rays = wl.c2_rays(tris, 1 << 20)
    occ = wl.c2_rays(tris, 1 << 20, occlusion=True)
    return tris, rays, occ

"""MI355X ray-query engine: closest-hit and occlusion queries over triangle
scenes, traced by hand-written HIP kernels for gfx950 behind a C ABI
(include/rt/rt_c.h).  This package is the Python host mirror used by the
tests and the benchmark; the native library is librt_hip.so, built in-tree.
"""
from ._lib import (HIT_DTYPE, RAY_DTYPE, RT_MISS, TRIANGLE_DTYPE, RT_NO_HIT_T, LIB_PATH, RtError, load,  # noqa: F401
                   primary_camera)
from .rt import (Batch, Context, PathTracer, HostBvh, Scene, autotune, build_opts, device_count, hits_to_numpy,  # noqa: F401
                 lbvh_binary, make_camera, opts_dict, rays_to_device, screen_camera, trace_opts)
from . import workload  # noqa: F401
from .serve import Server  # noqa: F401

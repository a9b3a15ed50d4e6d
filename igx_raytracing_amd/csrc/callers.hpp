// Launch interface of the caller kernels (callers.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace rtx {

// The screen camera of camera.glsl:72-87 (eye and three near-plane corners)
// plus the per-frame pixel offset of camera.glsl:115 and 1/resolution.
struct CameraDev {
    float eye[3], p0[3], p1[3], p2[3];
    float jitter[2];
    float inv_res[2];
};

struct CallerLaunch {
    CameraDev cam;
    uint32_t width = 0, height = 0, samples = 1;
    const uint32_t *normals = nullptr;  // raygen: 3 packed vertex normals per input triangle
    float4 *dir_object = nullptr;       // raygen outputs, [height][width]
    float4 *uv_normal = nullptr;
    const rt_ray *rays = nullptr;       // shadow tiles input, [samples][height][width]
    unsigned long long *tiles = nullptr;
    uint2 *gstack = nullptr;
    uint32_t gstack_depth = 0;
    uint32_t *err = nullptr;
};

enum class CallerKind { Raygen, ShadowTiles };
using CallerFn = void (*)(const DevScene, const CallerLaunch);

CallerFn caller_kernel(uint32_t width, CallerKind k);
uint32_t caller_lds_stack(CallerKind k);
int launch_caller(uint32_t width, CallerKind k, const DevScene &s, const CallerLaunch &a, uint32_t blocks,
                  hipStream_t stream);

}  // namespace rtx

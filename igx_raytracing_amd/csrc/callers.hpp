// Launch interface of the caller kernels (callers.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace rtx {

// The fields of the reference's Camera (camera.glsl:14-44) that
// calculatePrimary reads (camera.glsl:113-138): eye, the left screen
// p0..p2, the right screen p3..p5 (stereo), ipd (omnidirectional), the
// projection type, and 1/resolution; plus the pixel offset: rand(loc +
// seed) per pixel (camera.glsl:115, Seed.randomX/Y of rand_util.glsl:5-10)
// or a fixed one.
struct CameraDev {
    float eye[3], p0[3], p1[3], p2[3];
    float p3[3], p4[3], p5[3];
    float jitter[2];    // fixed pixel offset (rand_jitter == 0)
    float inv_res[2];
    float seed[2];      // randLoc of calculatePrimary (rand_jitter == 1)
    float ipd = 0.f;
    uint32_t projection = 0;   // RT_PROJ_*
    uint32_t rand_jitter = 0;
};

// The reference's Light (primitive.glsl:46-54): position, radius/origin as
// half2, an encodeNormal-packed direction, and colour + type (type =
// color_type[1] >> 16; 0 directional, 2 point).
struct LightDev {
    float pos[3];
    uint32_t rad_origin;
    uint32_t dir[2];
    uint32_t color_type[2];
};

struct CallerLaunch {
    CameraDev cam;
    LightDev light;                     // shadow pass
    const float4 *dir_in = nullptr;     // shadow pass input: raygen's dirObject [height][width]
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

enum class CallerKind { Raygen, ShadowTiles, ShadowPass };
using CallerFn = void (*)(const DevScene, const CallerLaunch);

CallerFn caller_kernel(uint32_t width, CallerKind k);
uint32_t caller_lds_stack(CallerKind k);
int launch_caller(uint32_t width, CallerKind k, const DevScene &s, const CallerLaunch &a, uint32_t blocks,
                  hipStream_t stream);

}  // namespace rtx

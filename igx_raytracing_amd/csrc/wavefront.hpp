// Launch interface of the wavefront path tracer (wavefront.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "callers.hpp"
#include "kernels.hpp"

namespace rtx {

// State of one path, 32 bytes.
struct PathState {
    float throughput;
    float radiance;
    uint32_t pixel;
    uint32_t pad;
    uint64_t rng_state;
    uint64_t rng_inc;
};
static_assert(sizeof(PathState) == 32, "PathState must be 32 bytes");

struct PtLaunch {
    CameraDev cam;
    uint32_t width = 0, height = 0, spp = 1, max_depth = 3;
    float rr_survive = 0.8f, albedo = 0.7f;
    const uint64_t *frame_seed = nullptr;  // device word, set before each frame
    PathState *paths = nullptr;
    rt_ray *rays[2] = {nullptr, nullptr};    // queue-ordered rays, by depth parity
    uint32_t *queue[2] = {nullptr, nullptr}; // path index of each queued ray
    uint32_t *counts = nullptr;              // rays queued at depth d, max_depth + 2 words
    float *image = nullptr;
    uint2 *gstack = nullptr;
    uint32_t gstack_depth = 0;
    int32_t dual_max = -1;  // see TraceLaunch::dual_max
    uint32_t *err = nullptr;
    uint32_t *work = nullptr;
    uint32_t chunk = 64, refill = 8;  // bounce feed (WaveFeed)
};

using PtBounceFn = void (*)(const DevScene, const PtLaunch, uint32_t);
PtBounceFn pt_bounce_kernel(uint32_t width);
// Enqueues one frame: camera, max_depth + 1 bounces, resolve.
int pt_enqueue_frame(const DevScene &s, const PtLaunch &a, uint32_t bounce_blocks, hipStream_t stream);

}  // namespace rtx

// Launch interface of the query kernels (implemented in trace.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "rt/rt_c.h"
#include "rt_format.h"

namespace rtx {

// Analytic primitives of a scene, after the triangles in the object-index
// space (trace.glsl:22-46): ids num_tris.. are spheres, then cubes, then
// planes.
struct AnaDev {
    const float4 *spheres = nullptr;  // centre.xyz, radius
    const float *cubes = nullptr;     // start.xyz, end.xyz (6 floats each)
    const float4 *planes = nullptr;   // normal.xyz, offset
    uint32_t num_spheres = 0, num_cubes = 0, num_planes = 0;
    uint32_t count = 0;               // sum of the three
    uint64_t gen = 0;                 // host-side identity of the set (rt_scene_set_analytic call)
};

// Device-resident scene: what one query launch reads.
struct DevScene {
    const void *nodes = nullptr;    // NodeB2 / NodeB4 / NodeB8 array
    const float4 *tris = nullptr;   // TriRec array (3 float4 per triangle)
    const float *src = nullptr;     // input-order p0 e1 e2 (9 floats), optional
    uint32_t num_tris = 0;
    uint32_t width = 4;
    uint32_t root = kRefNone;
    uint32_t max_stack = 0;
    AnaDev ana;
    float lo[3] = {0.f, 0.f, 0.f};  // the triangles' bounding box (ray-order heuristics only)
    float hi[3] = {0.f, 0.f, 0.f};
};

struct TraceLaunch {
    const rt_ray *rays = nullptr;
    uint32_t n = 0;
    rt_hit *hits = nullptr;          // closest-hit output
    uint8_t *flags = nullptr;        // occlusion output (bytes)
    unsigned long long *masks = nullptr;  // occlusion output (wave64 ballots)
    uint32_t *counts = nullptr;      // optional: 2 per ray
    uint2 *gstack = nullptr;         // spill stack, gstack_depth * stride entries
    uint32_t gstack_depth = 0;
    int32_t dual_max = -1;           // closest hit: stack size up to which a step may test two nodes
    uint32_t *err = nullptr;         // device error word
    uint32_t *work = nullptr;        // dynamic schedule: [next ray, finished waves], zero between launches
    uint32_t refill = 16;            // dynamic schedule: idle lanes before a wave hands out rays
    uint32_t chunk = 64;             // dynamic schedule: rays a wave reserves at a time
    uint32_t pool_shift = 3;         // dynamic schedule: n >> pool_shift rays form the shared pool
};

enum class Query { Closest, Occluded };

// Grid and LDS configuration a launch will use.
enum class Schedule : uint32_t { Dynamic = 0, Static = 1 };

struct LaunchShape {
    uint32_t block = 256;
    uint32_t lds_stack = 16;
    uint32_t blocks = 0;  // persistent grid size
    Schedule schedule = Schedule::Dynamic;
};

// Maximum resident blocks per CU for the kernel variant (0 if not built).
int trace_occupancy(uint32_t width, Query q, bool count, uint32_t lds_stack, uint32_t block, Schedule sched);

// Launches one query kernel.  Returns a hipError_t value.
int launch_trace(const DevScene &s, Query q, const TraceLaunch &a, const LaunchShape &shape,
                 hipStream_t stream);

// ---- multi-scene batches -----------------------------------------------

// One scene of a batch (the per-scene BVH offset table).
struct BatchScene {
    const void *nodes;
    const float4 *tris;
    uint32_t root;
    uint32_t pad;
};

// A batch scene's analytic primitives (beside the BVH offset table; read
// only by the analytic post-pass).  Their object ids start at num_tris.
struct BatchAna {
    AnaDev ana;
    uint32_t num_tris = 0;
    uint32_t pad = 0;
};

// One query of a batch: rays [start, start + n) of the launch's index space.
struct BatchQuery {
    uint32_t scene;
    uint32_t kind;   // 0 closest (out = rt_hit*), 1 occlusion (out = uint8_t*)
    uint32_t n;
    uint32_t start;  // exclusive prefix of n over the queries
    const rt_ray *rays;
    void *out;
};

struct BatchLaunch {
    const BatchScene *scenes = nullptr;
    const BatchQuery *queries = nullptr;
    uint32_t nq = 0;
    uint32_t total = 0;          // sum of n
    uint2 *gstack = nullptr;
    uint32_t gstack_depth = 0;
    int32_t dual_max = -1;
    uint32_t *err = nullptr;
    uint32_t *work = nullptr;
    uint32_t refill = 16;
    uint32_t chunk = 64;
    uint32_t pool_shift = 3;
};

// Spill-stack depth per lane for a launch whose tracers may test two nodes
// per step, and the stack size up to which they may (TraceLaunch::dual_max).
// A step that tests two nodes grows the stack by at most 2W - 2 entries; a
// one-node walk from any stack never grows it by more than max_stack
// (the builders' bound over root-to-node paths of sum(children - 1)).  So
// with capacity C = lds_stack + depth, taking two nodes only while the stack
// holds at most C - max_stack - (2W - 2) entries keeps every later size
// within (that + 2W - 2) + max_stack <= C.  The extra max_stack / 2 + 2W
// entries of depth let two-node steps run while the stack is up to about
// half its worst case.
struct DualStack {
    uint32_t depth;
    int32_t dual_max;
};
inline DualStack dual_stack(uint32_t max_stack, uint32_t lds_stack, uint32_t width, bool dual) {
    const uint32_t base = (max_stack > lds_stack ? max_stack - lds_stack : 0u) + 1u;
    if (!dual || width > 4) return {base, -1};
    const uint32_t depth = base + max_stack / 2u + 2u * width;
    const int64_t cap = static_cast<int64_t>(lds_stack) + depth;
    const int64_t lim = cap - max_stack - (2 * static_cast<int64_t>(width) - 2);
    return {depth, static_cast<int32_t>(lim)};
}

int batch_occupancy(uint32_t width, uint32_t lds_stack, uint32_t block);
// One persistent launch over every query of every scene (dynamic schedule).
int launch_batch(uint32_t width, const BatchLaunch &a, const LaunchShape &shape, hipStream_t stream);
// After launch_batch on the same stream: the analytic primitives of the
// queries whose scene has any (ana[scene]), continuing each ray's result.
int launch_batch_analytic(const BatchLaunch &a, const BatchAna *ana, hipStream_t stream);

// The reference's linear scan over input-order triangles, then the
// analytic primitives.
int launch_brute(const DevScene &s, Query q, const rt_ray *rays, uint32_t n, rt_hit *hits,
                 uint8_t *flags, hipStream_t stream);

}  // namespace rtx

// Wavefront path tracing over the query kernels: diffuse bounces with
// per-bounce ray queues.
//
//   k_pt_camera   one path per (pixel, sample): PCG32 seeded per pixel
//                 (stream = sample), jittered screen-camera ray, queue 0
//   k_pt_bounce   one persistent launch per bounce: each lane takes a queued
//                 ray (dynamic schedule), traces the closest hit, shades it
//                 (sky on a miss; Russian roulette, cosine-weighted
//                 reflection off the geometric normal on a hit) and stages
//                 the surviving rays per wave in LDS (ballot + mbcnt
//                 compaction), appending them to the next queue 64 at a
//                 time with one atomic, so the next bounce gets a dense,
//                 uneven-sized queue
//   k_pt_resolve  per pixel, the mean of its samples in sample order
// Queue counts live in device memory, so the whole frame (camera, bounces,
// resolve) is captured once in a hipGraph and replayed; the frame seed is a
// device word set before each replay.  A path's radiance depends only on its
// own rays and random numbers, so queue order (which depends on timing)
// never changes the image.  Sampling uses no transcendental functions (disk
// by rejection, the orthonormal basis of Duff et al. 2017), so the image is
// reproducible bit for bit on any IEEE host.
#include <hip/hip_runtime.h>

#include "traverse.hpp"
#include "wavefront.hpp"

namespace rtx {
namespace {

using namespace dev;

struct Pcg {
    uint64_t state, inc;
    __device__ __forceinline__ uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        const uint32_t xs = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = static_cast<uint32_t>(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    // [0, 1) with 24 bits.
    __device__ __forceinline__ float uniform() { return static_cast<float>(next() >> 8) * (1.0f / 16777216.0f); }
};

__device__ __forceinline__ Pcg pcg_init(uint64_t seed, uint64_t stream) {
    Pcg g;
    g.state = 0;
    g.inc = (stream << 1u) | 1u;
    g.next();
    g.state += seed;
    g.next();
    return g;
}

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return (ax * bx + ay * by) + az * bz;
}

__global__ void __launch_bounds__(256) k_pt_camera(const PtLaunch a) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    const uint32_t npaths = a.width * a.height * a.spp;
    if (p == 0) a.counts[0] = npaths;
    if (p >= npaths) return;
    const uint32_t pixel = p / a.spp, sample = p % a.spp;
    const uint32_t x = pixel % a.width, y = pixel / a.width;
    Pcg g = pcg_init(*a.frame_seed + pixel, sample);
    const CameraDev &c = a.cam;
    const float jx = g.uniform(), jy = g.uniform();
    const float cx = (static_cast<float>(x) + jx) * c.inv_res[0];
    float cy = (static_cast<float>(y) + jy) * c.inv_res[1];
    cy = 1.f - cy;
    float d[3];
    for (int k = 0; k < 3; ++k) {
        const float right = c.p1[k] - c.p0[k], up = c.p2[k] - c.p0[k];
        d[k] = (c.p0[k] + cx * right) + cy * up - c.eye[k];
    }
    const float l = sqrtf(dot3(d[0], d[1], d[2], d[0], d[1], d[2]));
    PathState ps;
    ps.throughput = 1.f;
    ps.radiance = 0.f;
    ps.pixel = pixel;
    ps.pad = 0;
    ps.rng_state = g.state;
    ps.rng_inc = g.inc;
    a.paths[p] = ps;
    float4 *r = reinterpret_cast<float4 *>(a.rays[0] + p);
    r[0] = make_float4(c.eye[0], c.eye[1], c.eye[2], 3.4028235e38f);
    r[1] = make_float4(d[0] / l, d[1] / l, d[2] / l, __uint_as_float(RT_MISS));
    a.queue[0][p] = p;
}

__global__ void k_pt_zero(uint32_t *p) { *p = 0u; }

// Shading of one finished query; returns true and the next ray if the path
// continues.
__device__ __forceinline__ bool shade(const PtLaunch &a, uint32_t depth, PathState &ps, const RayD &r, bool hit,
                                      float t, uint32_t id, const float4 &E1, const float4 &E2, RayD &nr) {
    if (!hit) {
        ps.radiance = ps.radiance + ps.throughput * (0.5f + 0.5f * r.dy);  // sky
        return false;
    }
    if (depth >= a.max_depth) return false;
    Pcg g{ps.rng_state, ps.rng_inc};
    if (depth >= 1) {  // Russian roulette
        if (g.uniform() >= a.rr_survive) {
            ps.rng_state = g.state;
            return false;
        }
        ps.throughput = ps.throughput / a.rr_survive;
    }
    // Geometric normal facing the incoming ray.
    float nx = E1.y * E2.z - E2.y * E1.z, ny = E1.z * E2.x - E2.z * E1.x, nz = E1.x * E2.y - E2.x * E1.y;
    if (dot3(nx, ny, nz, r.dx, r.dy, r.dz) > 0.f) {
        nx = -nx;
        ny = -ny;
        nz = -nz;
    }
    const float nl = sqrtf(dot3(nx, ny, nz, nx, ny, nz));
    nx = nx / nl;
    ny = ny / nl;
    nz = nz / nl;
    // Cosine-weighted direction: a point of the unit disk by rejection,
    // lifted to the hemisphere.
    float u, v, r2;
    do {
        u = 2.0f * g.uniform() - 1.0f;
        v = 2.0f * g.uniform() - 1.0f;
        r2 = u * u + v * v;
    } while (r2 >= 1.0f);
    const float w = sqrtf(1.0f - r2);
    // Orthonormal basis around n (Duff et al. 2017).
    const float sg = copysignf(1.0f, nz);
    const float A = -1.0f / (sg + nz);
    const float B = nx * ny * A;
    const float b1x = 1.0f + sg * nx * nx * A, b1y = sg * B, b1z = -sg * nx;
    const float b2x = B, b2y = sg + ny * ny * A, b2z = -ny;
    nr.dx = (u * b1x + v * b2x) + w * nx;
    nr.dy = (u * b1y + v * b2y) + w * ny;
    nr.dz = (u * b1z + v * b2z) + w * nz;
    nr.ox = r.ox + t * r.dx;
    nr.oy = r.oy + t * r.dy;
    nr.oz = r.oz + t * r.dz;
    ps.throughput = ps.throughput * a.albedo;
    ps.rng_state = g.state;
    (void)id;
    return true;
}

template <int W, int S>
__global__ void __launch_bounds__(256) k_pt_bounce(const DevScene s, const PtLaunch a, uint32_t depth) {
    // Bounce rays mostly hit and end soon after their first leaf: the
    // two-fetch step (see Tracer::advance) measured 5% faster for them.
    using TR = Tracer<W, 0, false, S, 256, false>;
    __shared__ typename TR::Entry lds_stack[S * 256];
    __shared__ MergeSlots<256> slots;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t total_waves = gridDim.x * 4u;
    const uint32_t in = depth & 1u, out = in ^ 1u;
    const uint32_t n = a.counts[depth];
    TR t;
    init_stack(t, lds_stack, a.gstack, a.gstack_depth, &slots, a.dual_max);
    // A bounce ray tests the root (wave registers) and the root's child it
    // visits first (block LDS) when it begins, as the any-hit kernels do:
    // bounce rays are short, so the step saved counts (+1.9% per frame,
    // BVH4, when the queue appends were the bottleneck; within noise now).
    // BVH8 keeps the plain begin: the root copy's registers cost its
    // closest-hit walks more than the step saved.
    constexpr bool kTop = W <= 4;
    RootWords<W> rw;
    if (kTop) rw.load(s.nodes, s.root); else rw.ok = false;
    const TopNodes<W> *top = nullptr;
    if constexpr (kTop) {
        __shared__ TopNodes<W> top_lds;
        top_lds.load(s.nodes, rw);
        top = &top_lds;
    }
    WaveFeed feed;
    feed.init(n, 3u, blockIdx.x * 4u + threadIdx.x / 64u, total_waves);
    // Rays that continue are staged per wave in LDS and appended to the next
    // depth's queue kStage at a time.  Every wave's append is an atomic on
    // the one queue counter, and single-address atomics serialise across the
    // device: with one append per iteration that had a continuing lane, the
    // bounces that spawn rays ran 3-4x slower per ray than the last one,
    // which spawns none (frame 23.7 ms; staged, 9.8 ms, identical images).
    // The queue's order changes, and nothing depends on it.
    constexpr uint32_t kStage = 64;
    // After a flush the area takes every lane of the wave at once.
    static_assert(kStage >= 64, "a wave can stage one ray per lane in one iteration");
    __shared__ float4 stage_ray[4][kStage * 2];
    __shared__ uint32_t stage_path[4][kStage];
    __shared__ uint32_t stage_n[4];  // staged rays per wave
    const uint32_t wib = threadIdx.x / 64u;
    if (lane == 0u) stage_n[wib] = 0u;
    // Appends the wave's cnt staged rays; the active lanes share the copy.
    auto flush = [&](uint32_t cnt) RTX_INLINE {
        __builtin_amdgcn_wave_barrier();
        const unsigned long long act = __ballot(true);
        const uint32_t nact = static_cast<uint32_t>(__popcll(act)), r = lane_rank(act);
        uint32_t base = 0;
        if (r == 0u) base = atomicAdd(&a.counts[depth + 1], cnt);
        base = __builtin_amdgcn_readfirstlane(base);
        for (uint32_t j = r; j < cnt; j += nact) {
            float4 *dst = reinterpret_cast<float4 *>(a.rays[out] + base + j);
            dst[0] = stage_ray[wib][2 * j];
            dst[1] = stage_ray[wib][2 * j + 1];
            a.queue[out][base + j] = stage_path[wib][j];
        }
    };
    uint32_t i = 0;
    RayRec rr;
    dynamic_loop(
        t, feed, a.work, a.chunk, a.refill,
        [&](uint32_t k) RTX_INLINE {
            i = k;
            rr = load_ray(a.rays[in] + i);
        },
        [&]() RTX_INLINE {
            if constexpr (kTop)
                t.begin_at_top(rr, s.root, rw, *top);
            else
                t.begin(rr, s.root);
        },
        [&](typename TR::Fetched &f) RTX_INLINE { t.fetch(s.nodes, s.tris, f); },
        [&](const typename TR::Fetched &f) RTX_INLINE { return t.visit(s.nodes, s.tris, f); },
        [&]() RTX_INLINE { return t.advance(s.nodes, s.tris); },
        [&]() RTX_INLINE {
            // Finished lanes of this iteration shade and compact together.
            if (t.overflow) atomicOr(a.err, 1u);
            const uint32_t p = a.queue[in][i];
            PathState ps = a.paths[p];
            const bool hit = t.best_t < t.tlim;
            float4 E1 = make_float4(0.f, 0.f, 0.f, 0.f), E2 = E1;
            if (hit) {
                const float4 *tp = s.tris + static_cast<size_t>(t.best_rec) * 3;
                E1 = tp[1];
                E2 = tp[2];
            }
            RayD nr;
            const bool keep = shade(a, depth, ps, t.rd, hit, t.best_t, t.best_id, E1, E2, nr);
            a.paths[p] = ps;
            const unsigned long long m = __ballot(keep);
            if (m) {
                const uint32_t k = static_cast<uint32_t>(__popcll(m));
                uint32_t n0 = stage_n[wib];
                if (n0 + k > kStage) {
                    flush(n0);
                    n0 = 0u;
                }
                if (keep) {
                    const uint32_t slot = n0 + lane_rank(m);
                    stage_ray[wib][2 * slot] = make_float4(nr.ox, nr.oy, nr.oz, 3.4028235e38f);
                    stage_ray[wib][2 * slot + 1] = make_float4(nr.dx, nr.dy, nr.dz, __uint_as_float(t.best_id));
                    stage_path[wib][slot] = p;
                }
                __builtin_amdgcn_wave_barrier();
                if (lane_rank(__ballot(true)) == 0u) stage_n[wib] = n0 + k;
            }
        },
        [](uint32_t, bool) RTX_INLINE {},
        [&]() RTX_INLINE {
            keep(rr.a);
            keep(rr.b);
        });
    {
        const uint32_t n0 = stage_n[wib];
        if (n0 != 0u) flush(n0);
    }
    release_work(a.work, total_waves);
}

__global__ void __launch_bounds__(256) k_pt_resolve(const PtLaunch a) {
    const uint32_t pixel = blockIdx.x * 256 + threadIdx.x;
    if (pixel >= a.width * a.height) return;
    float sum = 0.f;
    for (uint32_t s = 0; s < a.spp; ++s) sum = sum + a.paths[static_cast<size_t>(pixel) * a.spp + s].radiance;
    a.image[pixel] = sum / static_cast<float>(a.spp);
}

template <int W>
PtBounceFn pick_w() {
    return &k_pt_bounce<W, 16>;
}

}  // namespace

PtBounceFn pt_bounce_kernel(uint32_t width) {
    switch (width) {
        case 2: return pick_w<2>();
        case 4: return pick_w<4>();
        case 8: return pick_w<8>();
        default: return nullptr;
    }
}

int pt_enqueue_frame(const DevScene &s, const PtLaunch &a, uint32_t bounce_blocks, hipStream_t st) {
    PtBounceFn fn = pt_bounce_kernel(s.width);
    if (!fn) return hipErrorInvalidValue;
    const uint32_t npaths = a.width * a.height * a.spp;
    hipLaunchKernelGGL(k_pt_camera, dim3((npaths + 255) / 256), dim3(256), 0, st, a);
    for (uint32_t d = 0; d <= a.max_depth; ++d) {
        hipLaunchKernelGGL(k_pt_zero, dim3(1), dim3(1), 0, st, a.counts + d + 1);
        hipLaunchKernelGGL(fn, dim3(bounce_blocks), dim3(256), 0, st, s, a, d);
    }
    hipLaunchKernelGGL(k_pt_resolve, dim3((a.width * a.height + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace rtx

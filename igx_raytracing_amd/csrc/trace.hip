// Ray-query kernels for gfx950: stack-based traversal of BVH2 / quantised
// BVH4 / BVH8 nodes with Moller-Trumbore triangle tests, closest-hit and
// any-hit (occlusion), plus the reference's own linear scan.
//
// Build with -ffp-contract=off: the triangle test must round every product
// and sum exactly as the reference states them (primitive.glsl:239-284).  The
// box tests use explicit fmaf, which is conservative by construction (boxes
// are padded outward when built) and identical on any IEEE host.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "kernels.hpp"
#include "traverse.hpp"

namespace rtx {
namespace {

using namespace dev;

// Static schedule.  Persistent grid: wave w handles ray chunks w, w + waves,
// ... of 64 consecutive rays (one ray per lane), so the occlusion ballot of a
// chunk is one 64-bit word (nv_all.shadow.comp:128-140).
template <int W, bool ANY, bool COUNT, int S, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_trace(const DevScene s, const TraceLaunch a) {
    using TR = Tracer<W, ANY ? 1 : 0, COUNT, S, BLOCK>;
    constexpr bool kSteal = !COUNT;  // counts follow the single-lane walk
    __shared__ typename TR::Entry lds_stack[S * BLOCK];
    const uint32_t lane = threadIdx.x & 63u;
    constexpr uint32_t kWavesPerBlock = BLOCK / 64;
    const uint32_t total_waves = gridDim.x * kWavesPerBlock;
    const uint32_t wave = blockIdx.x * kWavesPerBlock + threadIdx.x / 64u;
    const uint32_t nchunks = (a.n + 63u) / 64u;
    __shared__ MergeSlots<BLOCK> slots;
    TR t;
    init_stack(t, lds_stack, a.gstack, a.gstack_depth, kSteal ? &slots : nullptr, a.dual_max);
    // Any-hit walks test the root from the wave's copy of its words (+3-5%
    // occlusion throughput), and BVH2/BVH4 ones the next level from the
    // block's LDS copy too (+2-3.5% more; BVH8 lost 3% to it).  Closest-hit
    // walks gained nothing from the root copy (BVH4) or lost to the
    // registers it holds (BVH8, -6%).
    constexpr bool kTop = ANY && W <= 4;
    RootWords<W> rw;
    if (ANY) rw.load(s.nodes, s.root); else rw.ok = false;
    const TopNodes<W> *top = nullptr;
    if constexpr (kTop) {
        __shared__ TopNodes<W> top_lds;
        top_lds.load(s.nodes, rw);
        top = &top_lds;
    }
    for (uint32_t chunk = wave; chunk < nchunks; chunk += total_waves) {
        const uint32_t i = chunk * 64u + lane;
        bool alive = i < a.n, occluded = false;
        if (alive) {
            if constexpr (kTop)
                t.begin_at_top(load_ray(a.rays + i), s.root, rw, *top);
            else if (ANY)
                t.begin_at_root(load_ray(a.rays + i), s.root, rw);
            else
                t.begin(load_ray(a.rays + i), s.root);
            if (kSteal) t.claim_slot();
        }
        for (;;) {
            if (kSteal)
                steal(t, !alive && !t.helping, (alive || t.helping) && t.cur != kRefNone && t.st.lc >= 1,
                      [](uint32_t, bool) RTX_INLINE {});
            if ((alive || t.helping) && t.cur != kRefNone) t.advance(s.nodes, s.tris);
            if (kSteal) merge_helpers(t);
            if (alive && t.cur == kRefNone && (!kSteal || t.helpers_left() == 0u)) {
                if (kSteal) t.collect();
                t.finish(a, i);
                occluded = t.occluded;
                alive = false;
            }
            if (__ballot(alive || t.helping) == 0ull) break;
        }
        if (ANY && a.masks) {
            const unsigned long long m = __ballot(occluded);
            if (lane == 0) a.masks[chunk] = m;
        }
    }
}

// Dynamic schedule.  Each lane traces one ray at a time.  A wave reserves
// `chunk` consecutive ray indices (from its own share, then from the
// launch's pool) and hands them to its idle lanes by lane rank whenever at
// least `refill` lanes are idle, so lanes do not wait for the slowest ray of
// a group.  A lane handed a ray issues the load of its record after the
// wave's fetch and starts the ray in the next iteration, so the record's
// latency hides behind the node fetches instead of stalling the wave.  The
// last wave to finish zeroes the counter for the next launch on the stream.
template <int W, bool ANY, int S, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_trace_dyn(const DevScene s, const TraceLaunch a) {
    using TR = Tracer<W, ANY ? 1 : 0, false, S, BLOCK>;
    __shared__ typename TR::Entry lds_stack[S * BLOCK];
    __shared__ MergeSlots<BLOCK> slots;
    const uint32_t total_waves = gridDim.x * (BLOCK / 64);
    const uint32_t wave = blockIdx.x * (BLOCK / 64) + threadIdx.x / 64u;
    TR t;
    init_stack(t, lds_stack, a.gstack, a.gstack_depth, &slots, a.dual_max);
    constexpr bool kTop = ANY && W <= 4;  // as in k_trace
    RootWords<W> rw;
    if (ANY) rw.load(s.nodes, s.root); else rw.ok = false;
    const TopNodes<W> *top = nullptr;
    if constexpr (kTop) {
        __shared__ TopNodes<W> top_lds;
        top_lds.load(s.nodes, rw);
        top = &top_lds;
    }
    __shared__ uint32_t feed_lists[BLOCK / 64][RankedFeed::kWin];
    RankedFeed feed;
    feed.init(a.n, a.pool_shift, wave, total_waves, feed_lists[threadIdx.x / 64u], a.rays, s.lo, s.hi);
    uint32_t i = 0;
    RayRec rr;
    dynamic_loop(
        t, feed, a.work, a.chunk, a.refill,
        [&](uint32_t k) RTX_INLINE {
            i = k;
            rr = load_ray(a.rays + i);
        },
        [&]() RTX_INLINE {
            if constexpr (kTop)
                t.begin_at_top(rr, s.root, rw, *top);
            else if (ANY)
                t.begin_at_root(rr, s.root, rw);
            else
                t.begin(rr, s.root);
        },
        [&](typename TR::Fetched &f) RTX_INLINE { t.fetch(s.nodes, s.tris, f); },
        [&](const typename TR::Fetched &f) RTX_INLINE { return t.visit(s.nodes, s.tris, f); },
        [&]() RTX_INLINE { return t.advance(s.nodes, s.tris); },
        [&]() RTX_INLINE { t.finish(a, i); }, [](uint32_t, bool) RTX_INLINE {},
        [&]() RTX_INLINE {
            keep(rr.a);
            keep(rr.b);
        });
    release_work(a.work, total_waves);
}

// Multi-scene batch: the dynamic schedule over the concatenated index space
// of all queries.  A lane maps its index to a query (its current query's
// range, else a binary search on the queries' start offsets), then traces
// against that query's scene (per-lane node/triangle pointers from the scene
// table) with that query's kind, so scenes of uneven size and batches of
// uneven count share one launch.  Analytic primitives follow in a separate
// pass (k_batch_analytic_post).  Its lanes take one fetch per step, any-hit
// lanes included: on the 64-scene batch that measured 0.815 against 0.840
// ms for the two-fetch step (it had lost before steps marked their loads as
// used, see Tracer::settle).
template <int W, int S, int BLOCK>
__global__ void __launch_bounds__(BLOCK) k_trace_batch(const BatchLaunch a) {
    using TR = Tracer<W, 2, false, S, BLOCK, true>;
    __shared__ typename TR::Entry lds_stack[S * BLOCK];
    __shared__ MergeSlots<BLOCK> slots;
    const uint32_t total_waves = gridDim.x * (BLOCK / 64);
    const uint32_t wave = blockIdx.x * (BLOCK / 64) + threadIdx.x / 64u;
    TR t;
    init_stack(t, lds_stack, a.gstack, a.gstack_depth, &slots, a.dual_max);
    WaveFeed feed;
    feed.init(a.total, a.pool_shift, wave, total_waves);
    uint32_t i = 0, root = kRefNone;
    bool any = false;
    RayRec rr;
    const void *nodes = nullptr;
    const float4 *tris = nullptr;
    void *out = nullptr;
    // The lane's current query: rays [q_lo, q_hi) of the index space.  A
    // lane's next ray usually comes from the same dealt chunk, hence the same
    // query, and then needs only its ray record: the search and the query and
    // scene records are a chain of dependent loads the wave would wait on.
    uint32_t q_lo = 0, q_hi = 0;
    const rt_ray *q_rays = nullptr;
    dynamic_loop(
        t, feed, a.work, a.chunk, a.refill,
        [&](uint32_t g) RTX_INLINE {
            if (g < q_lo || g >= q_hi) {
                // Largest q with start <= g (queries[0].start == 0).
                uint32_t lo = 0, hi = a.nq - 1;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if (a.queries[mid].start <= g) lo = mid; else hi = mid - 1;
                }
                const BatchQuery q = a.queries[lo];
                const BatchScene sc = a.scenes[q.scene];
                q_lo = q.start;
                q_hi = q.start + q.n;
                q_rays = q.rays;
                nodes = sc.nodes;
                tris = sc.tris;
                root = sc.root;
                out = q.out;
                any = q.kind != 0;
            }
            i = g - q_lo;
            rr = load_ray(q_rays + i);
        },
        [&]() RTX_INLINE { t.begin(rr, root, any); },
        [&](typename TR::Fetched &f) RTX_INLINE { t.fetch(nodes, tris, f); },
        [&](const typename TR::Fetched &f) RTX_INLINE { return t.visit(nodes, tris, f); },
        [&]() RTX_INLINE { return t.advance(nodes, tris); },
        [&]() RTX_INLINE { t.store(static_cast<rt_hit *>(out), static_cast<uint8_t *>(out), i, a.err); },
        [&](uint32_t d, bool take) RTX_INLINE {
            // A thief traverses the donor's scene.
            const uint64_t n2 = __shfl(static_cast<unsigned long long>(reinterpret_cast<uint64_t>(nodes)),
                                       static_cast<int>(d));
            const uint64_t t2 = __shfl(static_cast<unsigned long long>(reinterpret_cast<uint64_t>(tris)),
                                       static_cast<int>(d));
            if (take) {
                nodes = reinterpret_cast<const void *>(n2);
                tris = reinterpret_cast<const float4 *>(t2);
                q_hi = 0;  // the scene is the donor's now
            }
        },
        [&]() RTX_INLINE {
            keep(rr.a);
            keep(rr.b);
        });
    release_work(a.work, total_waves);
}

// The reference's linear scan (trace.glsl:9-66 and :70-98): every ray tests
// every triangle in input order, rejecting t <= 0, the previous object and
// t >= hit.hitT (primitive.glsl:276).  Triangle loads are wave-uniform.
template <bool ANY>
__global__ void __launch_bounds__(256) k_brute(const float *__restrict__ src, uint32_t ntri, const AnaDev ana,
                                               const rt_ray *__restrict__ rays, uint32_t n,
                                               rt_hit *__restrict__ hits, uint8_t *__restrict__ flags) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const float4 *rp = reinterpret_cast<const float4 *>(rays + i);
    const float4 r0 = rp[0], r1 = rp[1];
    const RayD rd{r0.x, r0.y, r0.z, r1.x, r1.y, r1.z};
    const uint32_t prev = __float_as_uint(r1.w);
    float hit_t = kNoHit, hu = 0.f, hv = 0.f;  // trace.glsl:14-16
    uint32_t obj = 0;
    for (uint32_t j = 0; j < ntri; ++j) {
        const float *p = src + static_cast<size_t>(j) * 9;
        const float4 A = make_float4(p[0], p[1], p[2], 0.f);
        const float4 B = make_float4(p[3], p[4], p[5], 0.f);
        const float4 C = make_float4(p[6], p[7], p[8], 0.f);
        float t, u, v;
        if (!mt_test(rd, A, B, C, t, u, v)) continue;
        if (t <= 0.f || j == prev || t >= hit_t) continue;  // primitive.glsl:276
        hit_t = t;
        hu = u;
        hv = v;
        obj = j;
    }
    // trace.glsl:30-46 (:82-95): spheres, cubes, planes after the triangles.
    analytic_scan(ana, ntri, AnaRay{rd.ox, rd.oy, rd.oz, rd.dx, rd.dy, rd.dz}, prev, hit_t, obj, hu, hv);
    if (ANY) {
        flags[i] = hit_t < r0.w ? 1 : 0;  // trace.glsl:97
    } else {
        // raygen.comp:41 (miss iff hitT == noHit), then the query's tmax.
        bool hit = hit_t != kNoHit;
        if (!(r0.w >= kNoHit) && !(hit_t < r0.w)) hit = false;
        const float4 o = hit ? make_float4(__uint_as_float(obj), hit_t, hu, hv)
                             : make_float4(__uint_as_float(RT_MISS), kNoHit, 0.f, 0.f);
        *reinterpret_cast<float4 *>(hits + i) = o;
    }
}

// The scene's analytic primitives after a triangle query (trace.glsl:30-46):
// one thread per ray continues the reference's scan from the result the
// traversal stored, so the traversal kernels carry none of this code.
// Closest hit resumes from the stored (t, id, u, v), or from t = tlim on a
// miss, exactly the walk's final state.  Occlusion scans rays no triangle
// occluded; ballot output rewrites each 64-ray word from its own lanes.
template <bool ANY>
__global__ void __launch_bounds__(256) k_analytic_post(const AnaDev ana, uint32_t base, const TraceLaunch a) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const bool in = i < a.n;
    bool occluded = false;
    if (in) {
        const RayRec rr = load_ray(a.rays + i);
        const AnaRay r{rr.a.x, rr.a.y, rr.a.z, rr.b.x, rr.b.y, rr.b.z};
        const uint32_t prev = __float_as_uint(rr.b.w);
        const float tmax = rr.a.w;
        if (!ANY) {
            const float tlim = tmax == tmax ? (tmax < kNoHit ? tmax : kNoHit) : tmax;
            const float4 h = *reinterpret_cast<const float4 *>(a.hits + i);
            const bool hit = __float_as_uint(h.x) != RT_MISS;
            float t = hit ? h.y : tlim, u = hit ? h.z : 0.f, v = hit ? h.w : 0.f;
            uint32_t id = hit ? __float_as_uint(h.x) : 0u;
            analytic_scan(ana, base, r, prev, t, id, u, v);
            *reinterpret_cast<float4 *>(a.hits + i) =
                t < tlim ? make_float4(__uint_as_float(id), t, u, v)
                         : make_float4(__uint_as_float(RT_MISS), kNoHit, 0.f, 0.f);
        } else {
            occluded = a.flags ? a.flags[i] != 0 : ((a.masks[i >> 6] >> (i & 63u)) & 1ull) != 0ull;
            if (!occluded && tmax == tmax) {
                float t = kNoHit, u, v;
                uint32_t id = 0;
                analytic_scan(ana, base, r, prev, t, id, u, v);
                occluded = t < tmax;
                if (a.flags && occluded) a.flags[i] = 1;
            }
        }
    }
    if (ANY && a.masks) {
        const unsigned long long m = __ballot(occluded);
        if ((threadIdx.x & 63u) == 0u && (i & ~63u) < a.n) a.masks[i >> 6] = m;
    }
}

// The same pass for a batch: one thread per ray of the launch's index
// space, mapped to its query by binary search; queries whose scene has no
// analytic primitives are left as the traversal wrote them.  Batch
// occlusion output is one byte per ray.
__global__ void __launch_bounds__(256) k_batch_analytic_post(const BatchLaunch a, const BatchAna *__restrict__ ana) {
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    if (g >= a.total) return;
    uint32_t lo = 0, hi = a.nq - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (a.queries[mid].start <= g) lo = mid; else hi = mid - 1;
    }
    const BatchQuery q = a.queries[lo];
    const BatchAna sa = ana[q.scene];
    if (sa.ana.count == 0u) return;
    const uint32_t i = g - q.start;
    const RayRec rr = load_ray(q.rays + i);
    const AnaRay r{rr.a.x, rr.a.y, rr.a.z, rr.b.x, rr.b.y, rr.b.z};
    const uint32_t prev = __float_as_uint(rr.b.w);
    const float tmax = rr.a.w;
    if (q.kind == 0u) {
        rt_hit *hits = static_cast<rt_hit *>(q.out);
        const float tlim = tmax == tmax ? (tmax < kNoHit ? tmax : kNoHit) : tmax;
        const float4 h = *reinterpret_cast<const float4 *>(hits + i);
        const bool hit = __float_as_uint(h.x) != RT_MISS;
        float t = hit ? h.y : tlim, u = hit ? h.z : 0.f, v = hit ? h.w : 0.f;
        uint32_t id = hit ? __float_as_uint(h.x) : 0u;
        analytic_scan(sa.ana, sa.num_tris, r, prev, t, id, u, v);
        *reinterpret_cast<float4 *>(hits + i) = t < tlim ? make_float4(__uint_as_float(id), t, u, v)
                                                         : make_float4(__uint_as_float(RT_MISS), kNoHit, 0.f, 0.f);
    } else {
        uint8_t *flags = static_cast<uint8_t *>(q.out);
        if (flags[i] == 0 && tmax == tmax) {
            float t = kNoHit, u, v;
            uint32_t id = 0;
            analytic_scan(sa.ana, sa.num_tris, r, prev, t, id, u, v);
            if (t < tmax) flags[i] = 1;
        }
    }
}

using TraceFn = void (*)(const DevScene, const TraceLaunch);

template <int W, bool ANY, int S, int B>
TraceFn pick_sched(Schedule sched) {
    if (sched == Schedule::Dynamic) return &k_trace_dyn<W, ANY, S, B>;
    return &k_trace<W, ANY, false, S, B>;
}

template <int W, bool ANY>
TraceFn pick_s(uint32_t S, uint32_t block, bool count, Schedule sched) {
    if (count) return (S == 16 && block == 256) ? &k_trace<W, ANY, true, 16, 256> : nullptr;
    if (block == 256) {
        switch (S) {
            case 1: return pick_sched<W, ANY, 1, 256>(sched);
            case 4: return pick_sched<W, ANY, 4, 256>(sched);
            case 8: return pick_sched<W, ANY, 8, 256>(sched);
            case 16: return pick_sched<W, ANY, 16, 256>(sched);
            case 32: return pick_sched<W, ANY, 32, 256>(sched);
            default: return nullptr;
        }
    }
    if (block == 128) {
        switch (S) {
            case 8: return pick_sched<W, ANY, 8, 128>(sched);
            case 16: return pick_sched<W, ANY, 16, 128>(sched);
            case 32: return pick_sched<W, ANY, 32, 128>(sched);
            default: return nullptr;
        }
    }
    return nullptr;
}

template <int W>
TraceFn pick_w(Query q, bool count, uint32_t S, uint32_t block, Schedule sched) {
    return q == Query::Closest ? pick_s<W, false>(S, block, count, sched) : pick_s<W, true>(S, block, count, sched);
}

TraceFn pick(uint32_t width, Query q, bool count, uint32_t S, uint32_t block, Schedule sched) {
    switch (width) {
        case 2: return pick_w<2>(q, count, S, block, sched);
        case 4: return pick_w<4>(q, count, S, block, sched);
        case 8: return pick_w<8>(q, count, S, block, sched);
        default: return nullptr;
    }
}

using BatchFn = void (*)(const BatchLaunch);

template <int W, int B>
BatchFn pick_batch_s(uint32_t S) {
    switch (S) {
        case 8: return &k_trace_batch<W, 8, B>;
        case 16: return &k_trace_batch<W, 16, B>;
        case 32: return &k_trace_batch<W, 32, B>;
        default: return nullptr;
    }
}

template <int B>
BatchFn pick_batch_w(uint32_t width, uint32_t S) {
    switch (width) {
        case 2: return pick_batch_s<2, B>(S);
        case 4: return pick_batch_s<4, B>(S);
        case 8: return pick_batch_s<8, B>(S);
        default: return nullptr;
    }
}

BatchFn pick_batch(uint32_t width, uint32_t S, uint32_t block) {
    if (block == 256) return pick_batch_w<256>(width, S);
    if (block == 128) return pick_batch_w<128>(width, S);
    return nullptr;
}

}  // namespace

int batch_occupancy(uint32_t width, uint32_t lds_stack, uint32_t block) {
    BatchFn fn = pick_batch(width, lds_stack, block);
    if (!fn) return 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void *>(fn),
                                                     static_cast<int>(block), 0) != hipSuccess)
        return 0;
    return nb;
}

int launch_batch(uint32_t width, const BatchLaunch &a, const LaunchShape &shape, hipStream_t stream) {
    BatchFn fn = pick_batch(width, shape.lds_stack, shape.block);
    if (!fn) return hipErrorInvalidValue;
    if (a.total == 0 || a.nq == 0) return hipSuccess;
    hipLaunchKernelGGL(fn, dim3(shape.blocks), dim3(shape.block), 0, stream, a);
    return hipGetLastError();
}

int launch_batch_analytic(const BatchLaunch &a, const BatchAna *ana, hipStream_t stream) {
    if (a.total == 0 || a.nq == 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_analytic_post, dim3((a.total + 255u) / 256u), dim3(256), 0, stream, a, ana);
    return hipGetLastError();
}

int trace_occupancy(uint32_t width, Query q, bool count, uint32_t lds_stack, uint32_t block, Schedule sched) {
    TraceFn fn = pick(width, q, count, lds_stack, block, sched);
    if (!fn) return 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void *>(fn),
                                                     static_cast<int>(block), 0) != hipSuccess)
        return 0;
    return nb;
}

int launch_trace(const DevScene &s, Query q, const TraceLaunch &a, const LaunchShape &shape,
                 hipStream_t stream) {
    TraceFn fn = pick(s.width, q, a.counts != nullptr, shape.lds_stack, shape.block, shape.schedule);
    if (!fn) return hipErrorInvalidValue;
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(fn, dim3(shape.blocks), dim3(shape.block), 0, stream, s, a);
    if (s.ana.count) {
        const dim3 grid((a.n + 255u) / 256u);
        if (q == Query::Closest)
            hipLaunchKernelGGL(k_analytic_post<false>, grid, dim3(256), 0, stream, s.ana, s.num_tris, a);
        else
            hipLaunchKernelGGL(k_analytic_post<true>, grid, dim3(256), 0, stream, s.ana, s.num_tris, a);
    }
    return hipGetLastError();
}

int launch_brute(const DevScene &s, Query q, const rt_ray *rays, uint32_t n, rt_hit *hits, uint8_t *flags,
                 hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const dim3 grid((n + 255u) / 256u);
    if (q == Query::Closest)
        hipLaunchKernelGGL(k_brute<false>, grid, dim3(256), 0, stream, s.src, s.num_tris, s.ana, rays, n, hits,
                           nullptr);
    else
        hipLaunchKernelGGL(k_brute<true>, grid, dim3(256), 0, stream, s.src, s.num_tris, s.ana, rays, n, nullptr,
                           flags);
    return hipGetLastError();
}

}  // namespace rtx

// C ABI of the ray-query engine: contexts, scenes and query launches.
//
// The reference records compute dispatches into a command list per render
// task (RaygenTask::prepareCommandList, src/rt/task/raygen_task.cpp:88-94;
// ShadowTask::prepareCommandList, src/rt/task/shadow_task.cpp:194-215) and
// binds scene buffers packed by its SceneGraph; here a scene is an opaque
// handle holding device-resident nodes and triangle records, and each query
// is one kernel launch on a caller-given HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "alloc.hpp"
#include "bvh_build.hpp"
#include "kernels.hpp"
#include "lbvh.hpp"
#include "rt/rt_c.h"
#include "internal.hpp"
#include "callers.hpp"
#include "noalloc.hpp"

namespace {
thread_local std::string g_err;
}  // namespace

namespace rtx {

int fail_global(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int fail(rt_ctx ctx, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (ctx) {
        std::lock_guard<std::mutex> g(ctx->mu);
        ctx->err = buf;
    }
    g_err = buf;
    return code;
}

int hip_fail(rt_ctx ctx, hipError_t e, const char *what) {
    return fail(ctx, e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

void set_error(rt_ctx ctx, const std::string &msg) {
    if (ctx) {
        std::lock_guard<std::mutex> g(ctx->mu);
        ctx->err = msg;
    }
    g_err = msg;
}

}  // namespace rtx

using rtx::fail;
using rtx::fail_global;
using rtx::hip_fail;

namespace {

void fill_info(const rtx::HostBvh &h, rt_scene_info &info) {
    std::memset(&info, 0, sizeof info);
    info.num_tris = h.num_tris;
    info.node_width = h.width;
    info.num_nodes = h.num_nodes;
    info.root_ref = h.root_ref;
    info.max_depth = h.max_depth;
    info.max_stack = h.max_stack;
    info.num_leaves = h.num_leaves;
    info.node_bytes = h.node_bytes;
    info.nodes_total_bytes = static_cast<uint64_t>(h.num_nodes) * h.node_bytes;
    info.tris_total_bytes = static_cast<uint64_t>(h.tris.size()) * sizeof(rtx::TriRec);
    for (int k = 0; k < 3; ++k) {
        info.bounds_lo[k] = h.lo[k];
        info.bounds_hi[k] = h.hi[k];
    }
    info.build_ms = h.build_ms;
}

// Uploads a host-built layout (and optionally the input-order source) into a
// new scene.
int upload_scene(rt_ctx ctx, const rtx::HostBvh &h, const float *tri_xyz, bool keep_source, rt_scene *out) {
    rt_scene s = new (std::nothrow) rt_scene_s;
    if (!s) return fail(ctx, RT_ERR_OOM, "host allocation failed");
    s->ctx = ctx;
    fill_info(h, s->info);
    hipError_t e;
    const size_t nb = h.nodes.size(), tb = h.tris.size() * sizeof(rtx::TriRec);
    if (nb) {
        if ((e = rtx::dev_malloc(&s->d_nodes, nb)) != hipSuccess) goto err;
        if ((e = hipMemcpy(s->d_nodes, h.nodes.data(), nb, hipMemcpyHostToDevice)) != hipSuccess) goto err;
    }
    if (tb) {
        if ((e = rtx::dev_malloc(&s->d_tris, tb)) != hipSuccess) goto err;
        if ((e = hipMemcpy(s->d_tris, h.tris.data(), tb, hipMemcpyHostToDevice)) != hipSuccess) goto err;
    }
    if (keep_source && h.num_tris) {
        // Input-order p0, p1-p0, p2-p0 for the linear-scan entry points.
        std::vector<float> src(static_cast<size_t>(h.num_tris) * 9);
        for (size_t i = 0; i < h.num_tris; ++i) {
            const float *t = tri_xyz + i * 9;
            for (int k = 0; k < 3; ++k) {
                src[i * 9 + k] = t[k];
                src[i * 9 + 3 + k] = t[3 + k] - t[k];
                src[i * 9 + 6 + k] = t[6 + k] - t[k];
            }
        }
        if ((e = rtx::dev_malloc(&s->d_src, src.size() * sizeof(float))) != hipSuccess) goto err;
        if ((e = hipMemcpy(s->d_src, src.data(), src.size() * sizeof(float), hipMemcpyHostToDevice)) !=
            hipSuccess)
            goto err;
    }
    s->dev.nodes = s->d_nodes;
    s->dev.tris = static_cast<const float4 *>(s->d_tris);
    s->dev.src = s->d_src;
    s->dev.num_tris = h.num_tris;
    s->dev.width = h.width;
    s->dev.root = h.root_ref;
    s->dev.max_stack = h.max_stack;
    for (int k = 0; k < 3; ++k) {
        s->dev.lo[k] = h.lo[k];
        s->dev.hi[k] = h.hi[k];
    }
    *out = s;
    return RT_OK;
err:
    if (s->d_nodes) (void)hipFree(s->d_nodes);
    if (s->d_tris) (void)hipFree(s->d_tris);
    if (s->d_src) (void)hipFree(s->d_src);
    delete s;
    return hip_fail(ctx, e, "scene upload");
}

bool valid_stack(uint32_t s) { return s == 1 || s == 4 || s == 8 || s == 16 || s == 32; }

}  // namespace

namespace rtx {
// The stream's spill stack (grown to at least `entries`) and work counter.
int stream_resources(rt_ctx ctx, hipStream_t stream, size_t entries, uint2 **gstack, uint32_t **work) {
    std::lock_guard<std::mutex> g(ctx->mu);
    rt_ctx_s::Spill &sp = ctx->spill[stream];
    hipError_t e = hipSuccess;
    if (!sp.work) {
        // Zeroed on the stream that uses it: a plain hipMemset runs on the
        // null stream, which a non-blocking stream does not wait for.
        if ((e = rtx::dev_malloc(&sp.work, 2 * sizeof(uint32_t))) == hipSuccess)
            e = hipMemsetAsync(sp.work, 0, 2 * sizeof(uint32_t), stream);
        if (e != hipSuccess) {
            if (sp.work) (void)hipFree(sp.work);
            sp.work = nullptr;
            ctx->err = std::string("work counter: ") + hipGetErrorString(e);
            g_err = ctx->err;
            return e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP;
        }
    }
    if (sp.entries < entries) {
        if (sp.buf) {
            e = hipStreamSynchronize(stream);
            if (e == hipSuccess) e = hipFree(sp.buf);
            sp.buf = nullptr;
            sp.entries = 0;
            if (e != hipSuccess) {
                ctx->err = std::string("spill buffer release: ") + hipGetErrorString(e);
                g_err = ctx->err;
                return RT_ERR_HIP;
            }
        }
        e = rtx::dev_malloc(&sp.buf, entries * sizeof(uint2));
        if (e != hipSuccess) {
            sp.buf = nullptr;
            ctx->err = std::string("spill buffer: ") + hipGetErrorString(e);
            g_err = ctx->err;
            return e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP;
        }
        sp.entries = entries;
    }
    *gstack = sp.buf;
    *work = sp.work;
    return RT_OK;
}

void ana_hold(rt_ctx ctx, rt_scene scene, hipStream_t stream) {
    if (scene->ana_cur < 0 || !scene->dev.ana.count) return;
    rt_scene_s::AnaBuf &b = scene->ana_bufs[static_cast<size_t>(scene->ana_cur)];
    hipEvent_t ev = nullptr;
    if (!ctx->ev_pool.empty()) {
        ev = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        b.stuck = true;  // the use cannot be tracked: never rewrite this buffer
        return;
    }
    if (hipEventRecord(ev, stream) != hipSuccess) {
        ctx->ev_pool.push_back(ev);
        b.stuck = true;
        return;
    }
    b.uses.push_back(ev);
}

int batch_prepare(rt_ctx ctx, uint32_t width, uint32_t max_stack, uint32_t total, const rt_trace_opts *opts,
                  hipStream_t stream, LaunchShape &shape, BatchLaunch &a) {
    rt_trace_opts o;
    rt_trace_opts_default(&o);
    if (opts) o = *opts;
    if (o.block_size == 0) o.block_size = 256;
    if (o.lds_stack == 0) o.lds_stack = 16;
    if (o.refill == 0) o.refill = 16;
    if (o.chunk == 0) o.chunk = 64;
    if (o.pool_shift == 0) o.pool_shift = 3;
    if ((o.block_size != 256 && o.block_size != 128) || (o.lds_stack != 8 && o.lds_stack != 16 && o.lds_stack != 32) ||
        o.refill > 64 || o.chunk < 16 || o.chunk > 65536 || o.pool_shift > 32)
        return fail(ctx, RT_ERR_UNSUPPORTED, "batch launch shape block=%u lds_stack=%u not built", o.block_size,
                    o.lds_stack);
    if (total > 0x7FFFFFFFu) return fail(ctx, RT_ERR_INVALID, "at most 2^31-1 rays per batch launch");
    shape.block = o.block_size;
    shape.lds_stack = o.lds_stack;
    shape.schedule = Schedule::Dynamic;
    const uint32_t waves_per_block = shape.block / 64;
    // Resident waves per CU: full occupancy unless given.  Unlike a single
    // 1M-ray query, the 64-scene batch measured ~45% slower when capped at
    // the single-query default of 16.
    const int occ = batch_occupancy(width, shape.lds_stack, shape.block);
    uint64_t cap = static_cast<uint64_t>(ctx->num_cu) * std::max(occ, 1);
    if (o.waves_per_cu)
        cap = std::min<uint64_t>(
            cap, std::max<uint64_t>(1, static_cast<uint64_t>(ctx->num_cu) * o.waves_per_cu / waves_per_block));
    const uint64_t want = ((static_cast<uint64_t>(total) + 63) / 64 + waves_per_block - 1) / waves_per_block;
    shape.blocks = static_cast<uint32_t>(std::max<uint64_t>(1, std::min(want, cap)));
    const uint64_t max_blocks = (0xFFFFFFFFull - total) / (static_cast<uint64_t>(waves_per_block) * o.chunk);
    if (shape.blocks > max_blocks) shape.blocks = static_cast<uint32_t>(std::max<uint64_t>(1, max_blocks));
    // Closest-hit lanes may test two nodes per step; max_stack is the largest
    // over the batch's scenes, so the bound holds for each of them.
    const DualStack ds = dual_stack(max_stack, shape.lds_stack, width, true);
    const uint32_t depth = ds.depth;
    a.gstack_depth = depth;
    a.dual_max = ds.dual_max;
    a.total = total;
    a.refill = o.refill;
    a.chunk = o.chunk;
    a.pool_shift = o.pool_shift;
    a.err = ctx->d_err;
    return stream_resources(ctx, stream, static_cast<size_t>(depth) * shape.blocks * shape.block, &a.gstack, &a.work);
}

}  // namespace rtx

using rtx::stream_resources;

namespace {

// Frees a batch query table's device table and pinned slots (not its
// events); the caller has waited for the stream's use of them.
void free_table_buffers(rt_batch_s::Table &t) {
    if (t.d) (void)hipFree(t.d);
    for (rtx::BatchQuery *&h : t.h)
        if (h) (void)hipHostFree(h);
    t.d = nullptr;
    t.h[0] = t.h[1] = nullptr;
    t.used[0] = t.used[1] = false;
    t.cap = 0;
}

// Resolves the persistent grid, the schedule and the per-stream spill buffer
// and work counter for one launch.
// With `fixed` (the allocation-free entry point), the spill stack and work
// counter come from that scratch instead of the stream's buffers, and the
// grid is capped to the blocks whose spill stacks fit it.
struct FixedScratch {
    uint2 *stack = nullptr;
    size_t entries = 0;
    uint32_t *work = nullptr;
};

int prepare_launch(rt_ctx ctx, rt_scene scene, rtx::Query q, bool count, bool masks, uint32_t n,
                   const rt_trace_opts *opts, hipStream_t stream, rtx::LaunchShape &shape,
                   rtx::TraceLaunch &a, const FixedScratch *fixed = nullptr) {
    rt_trace_opts o;
    rt_trace_opts_default(&o);
    if (opts) o = *opts;
    // Per-query defaults measured on MI355X with 1M incoherent rays: closest
    // hit gains from idle lanes taking new rays; occlusion rays end early and
    // gain more from the occupancy of the smaller ring.
    if (o.block_size == 0) o.block_size = 256;
    if (o.schedule == 2) o.schedule = q == rtx::Query::Closest ? 0u : 1u;
    if (o.lds_stack == 0) o.lds_stack = q == rtx::Query::Closest ? 16u : 8u;
    // Closest hit: refill 12 measured 1.1% faster than 16 on 1M incoherent
    // rays (six interleaved runs; 8-14 within noise of each other); with
    // every-step tail stealing, 8 measured 0.3-0.5% faster than 12 (three
    // runs of interleaved launches).
    if (o.refill == 0) o.refill = q == rtx::Query::Closest ? 8u : 16u;
    if (o.chunk == 0) o.chunk = 64;
    // A small pool (1/256 of the rays): each wave hands out its dealt rays
    // longest first, and pool rays, which come last and in no order, put
    // long rays back at the end.  On 1M incoherent rays: 1/256 or no pool
    // 3,594-3,614 Mrays/s, 1/32 3,555, 1/16 3,559, 1/8 3,334.
    if (o.pool_shift == 0) o.pool_shift = 8;
    const bool shape_ok = o.block_size == 256 ? valid_stack(o.lds_stack)
                          : o.block_size == 128 && (o.lds_stack == 8 || o.lds_stack == 16 || o.lds_stack == 32);
    if (!shape_ok || o.schedule > 1 || o.refill > 64 || o.chunk < 16 || o.chunk > 65536 || o.pool_shift > 32)
        return fail(ctx, RT_ERR_UNSUPPORTED,
                    "launch shape block=%u lds_stack=%u schedule=%u refill=%u chunk=%u not built", o.block_size,
                    o.lds_stack, o.schedule, o.refill, o.chunk);
    if (n > 0x7FFFFFFFu) return fail(ctx, RT_ERR_INVALID, "at most 2^31-1 rays per launch");
    shape.block = o.block_size;
    shape.lds_stack = o.lds_stack;
    shape.schedule = o.schedule == 1 ? rtx::Schedule::Static : rtx::Schedule::Dynamic;
    if (count) {  // the counting variant is built for one shape, static schedule
        shape.block = 256;
        shape.lds_stack = 16;
        shape.schedule = rtx::Schedule::Static;
    }
    if (masks) shape.schedule = rtx::Schedule::Static;  // one ballot word per 64-ray chunk
    a.refill = o.refill;
    a.chunk = o.chunk;
    a.pool_shift = o.pool_shift;
    const uint32_t waves_per_block = shape.block / 64;
    const uint64_t nchunks = (static_cast<uint64_t>(n) + 63) / 64;
    uint64_t want = (nchunks + waves_per_block - 1) / waves_per_block;
    int occ = rtx::trace_occupancy(scene->dev.width, q, count, shape.lds_stack, shape.block, shape.schedule);
    if (occ <= 0) occ = 1;
    uint64_t cap = static_cast<uint64_t>(ctx->num_cu) * occ;
    // Resident waves per CU.  Measured on MI355X with 1M incoherent rays on
    // 1M clustered triangles: closest hit (dynamic schedule) is fastest at
    // 12 waves per CU and ~30% slower at its full occupancy of 20; occlusion
    // (static) at 24 of 32.  More waves mean fewer rays per wave, so a larger
    // share of each wave's life is the tail in which a few long rays run
    // with the other lanes idle.
    uint32_t wpc = o.waves_per_cu;
    if (wpc == 0 && !count) wpc = shape.schedule == rtx::Schedule::Dynamic ? 12u : 24u;
    if (wpc)
        cap = std::min<uint64_t>(cap, std::max<uint64_t>(1, static_cast<uint64_t>(ctx->num_cu) * wpc / waves_per_block));
    shape.blocks = static_cast<uint32_t>(std::max<uint64_t>(1, std::min(want, cap)));
    // The dynamic work counter reaches at most n + waves * chunk; keep it
    // below 2^32.
    const uint64_t max_blocks = (0xFFFFFFFFull - n) / (static_cast<uint64_t>(waves_per_block) * o.chunk);
    if (shape.blocks > max_blocks) shape.blocks = static_cast<uint32_t>(std::max<uint64_t>(1, max_blocks));
    // Closest-hit walks (not the counting one) may test two nodes per step.
    const rtx::DualStack ds =
        rtx::dual_stack(scene->dev.max_stack, shape.lds_stack, scene->dev.width, q == rtx::Query::Closest && !count);
    if (fixed) {
        const size_t per_block = static_cast<size_t>(ds.depth) * shape.block;
        const size_t fit = fixed->entries / per_block;
        if (fit == 0)
            return fail(ctx, RT_ERR_INVALID,
                        "scene's traversal stack needs %zu bytes of spill per block; the allocation-free "
                        "scratch holds %zu",
                        per_block * sizeof(uint2), fixed->entries * sizeof(uint2));
        if (shape.blocks > fit) shape.blocks = static_cast<uint32_t>(fit);
        a.gstack = fixed->stack;
        a.work = fixed->work;
    } else {
        const size_t entries = static_cast<size_t>(ds.depth) * shape.blocks * shape.block;
        int rc = stream_resources(ctx, stream, entries, &a.gstack, &a.work);
        if (rc != RT_OK) return rc;
    }
    a.gstack_depth = ds.depth;
    a.dual_max = ds.dual_max;
    a.err = ctx->d_err;
    return RT_OK;
}

int trace_common(rt_ctx ctx, rt_scene scene, rtx::Query q, const rt_ray *rays, uint32_t n, rt_hit *hits,
                 uint8_t *flags, uint64_t *masks, const rt_trace_opts *opts, uint32_t *counts, void *stream) {
    if (!ctx || !scene) return fail_global(RT_ERR_INVALID, "null context or scene");
    if (scene->ctx != ctx) return fail(ctx, RT_ERR_INVALID, "scene belongs to another context");
    if (n == 0) return RT_OK;
    if (!rays) return fail(ctx, RT_ERR_INVALID, "null rays");
    if (q == rtx::Query::Closest && !hits) return fail(ctx, RT_ERR_INVALID, "null hits");
    if (q == rtx::Query::Occluded && !flags && !masks) return fail(ctx, RT_ERR_INVALID, "null flags and masks");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    hipStream_t st = static_cast<hipStream_t>(stream);
    rtx::TraceLaunch a;
    a.rays = rays;
    a.n = n;
    a.hits = hits;
    a.flags = flags;
    a.masks = reinterpret_cast<unsigned long long *>(masks);
    a.counts = counts;
    rtx::LaunchShape shape;
    std::lock_guard<std::mutex> g(ctx->launch_mu);
    int rc = prepare_launch(ctx, scene, q, counts != nullptr, masks != nullptr, n, opts, st, shape, a);
    if (rc != RT_OK) return rc;
    e = static_cast<hipError_t>(rtx::launch_trace(scene->dev, q, a, shape, st));
    if (e != hipSuccess) return hip_fail(ctx, e, "trace launch");
    rtx::ana_hold(ctx, scene, st);
    return RT_OK;
}

}  // namespace

extern "C" {

void rt_trace_opts_default(rt_trace_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->block_size = 256;
    o->lds_stack = 0;
    o->waves_per_cu = 0;
    o->schedule = 2;
    o->refill = 0;  // the query's default, resolved at launch
    o->chunk = 64;
    o->pool_shift = 0;  // the query's default, resolved at launch
}

const char *rt_last_error_global(void) { return g_err.c_str(); }

const char *rt_last_error(rt_ctx ctx) {
    if (!ctx) return g_err.c_str();
    std::lock_guard<std::mutex> g(ctx->mu);
    thread_local std::string copy;
    copy = ctx->err;
    return copy.c_str();
}

int rt_device_count(int *count) {
    if (!count) return fail_global(RT_ERR_INVALID, "null count");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        *count = 0;
        return fail_global(RT_ERR_HIP, "hipGetDeviceCount: %s", hipGetErrorString(e));
    }
    *count = c;
    return RT_OK;
}

int rt_ctx_create(int device, rt_ctx *out) {
    if (!out) return fail_global(RT_ERR_INVALID, "null output");
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0)
        return fail_global(RT_ERR_HIP, "no HIP device: %s", e == hipSuccess ? "count is 0" : hipGetErrorString(e));
    if (device < 0 || device >= ndev) return fail_global(RT_ERR_INVALID, "device %d out of range", device);
    if ((e = hipSetDevice(device)) != hipSuccess)
        return fail_global(RT_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e));
    rt_ctx c = new (std::nothrow) rt_ctx_s;
    if (!c) return fail_global(RT_ERR_OOM, "host allocation failed");
    c->device = device;
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, device)) != hipSuccess) {
        delete c;
        return fail_global(RT_ERR_HIP, "hipGetDeviceProperties: %s", hipGetErrorString(e));
    }
    c->num_cu = prop.multiProcessorCount;
    {  // the allocation-free entry point's scratch on this device
        std::string serr;
        if (rtx::noalloc_scratch_init(device, serr) != 0) {
            delete c;
            return fail_global(RT_ERR_HIP, "%s", serr.c_str());
        }
    }
    if ((e = rtx::dev_malloc(&c->d_err, sizeof(uint32_t))) != hipSuccess ||
        (e = hipMemset(c->d_err, 0, sizeof(uint32_t))) != hipSuccess ||
        (e = hipDeviceSynchronize()) != hipSuccess) {  // visible to every stream before first use
        if (c->d_err) (void)hipFree(c->d_err);
        delete c;
        return fail_global(e == hipErrorOutOfMemory ? RT_ERR_OOM : RT_ERR_HIP, "error word: %s",
                           hipGetErrorString(e));
    }
    *out = c;
    return RT_OK;
}

int rt_ctx_destroy(rt_ctx ctx) {
    if (!ctx) return RT_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    for (auto &kv : ctx->spill) {
        if (kv.second.buf) (void)hipFree(kv.second.buf);
        if (kv.second.work) (void)hipFree(kv.second.work);
    }
    if (ctx->d_err) (void)hipFree(ctx->d_err);
    for (hipEvent_t ev : ctx->ev_pool) (void)hipEventDestroy(ev);
    if (ctx->upload) (void)hipStreamDestroy(ctx->upload);
    rtx::lbvh_workspace_free(ctx->lbvh);
    delete ctx;
    return RT_OK;
}

int rt_sync(rt_ctx ctx, void *stream) {
    if (!ctx) return fail_global(RT_ERR_INVALID, "null context");
    hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    return RT_OK;
}

int rt_check_device_error(rt_ctx ctx, void *stream, uint32_t *code) {
    if (!ctx || !code) return fail_global(RT_ERR_INVALID, "null argument");
    hipStream_t st = static_cast<hipStream_t>(stream);
    uint32_t v = 0, zero = 0;
    hipError_t e = hipMemcpyAsync(&v, ctx->d_err, sizeof v, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(ctx->d_err, &zero, sizeof zero, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(ctx, e, "device error word");
    *code = v;
    return RT_OK;
}

int rt_bvh_build(const float *tri_xyz, uint32_t n, const rt_build_opts *opts, rt_bvh *out) {
    if (!out) return fail_global(RT_ERR_INVALID, "null output");
    *out = nullptr;
    if (n && !tri_xyz) return fail_global(RT_ERR_INVALID, "null triangles");
    rt_build_opts o;
    std::string err;
    if (!rtx::resolve_opts(opts, o, err)) return fail_global(RT_ERR_INVALID, "%s", err.c_str());
    if (o.builder != RT_BUILD_SAH_CPU)
        return fail_global(RT_ERR_UNSUPPORTED, "rt_bvh_build is the host SAH builder");
    rt_bvh b = new (std::nothrow) rt_bvh_s;
    if (!b) return fail_global(RT_ERR_OOM, "host allocation failed");
    try {
        if (!rtx::build_host(tri_xyz, n, o, b->h, err)) {
            delete b;
            return fail_global(RT_ERR_INVALID, "%s", err.c_str());
        }
    } catch (const std::bad_alloc &) {
        delete b;
        return fail_global(RT_ERR_OOM, "host allocation failed during build");
    }
    *out = b;
    return RT_OK;
}

int rt_bvh_info(rt_bvh bvh, rt_scene_info *info) {
    if (!bvh || !info) return fail_global(RT_ERR_INVALID, "null argument");
    fill_info(bvh->h, *info);
    return RT_OK;
}

int rt_bvh_arrays(rt_bvh bvh, const void **nodes, const void **tris) {
    if (!bvh) return fail_global(RT_ERR_INVALID, "null bvh");
    if (nodes) *nodes = bvh->h.nodes.data();
    if (tris) *tris = bvh->h.tris.data();
    return RT_OK;
}

int rt_bvh_free(rt_bvh bvh) {
    delete bvh;
    return RT_OK;
}

int rt_scene_create(rt_ctx ctx, const float *tri_xyz, uint32_t n, const rt_build_opts *opts, rt_scene *out) {
    if (!ctx || !out) return fail_global(RT_ERR_INVALID, "null argument");
    *out = nullptr;
    if (n && !tri_xyz) return fail(ctx, RT_ERR_INVALID, "null triangles");
    rt_build_opts o;
    std::string err;
    if (!rtx::resolve_opts(opts, o, err)) return fail(ctx, RT_ERR_INVALID, "%s", err.c_str());
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    if (o.builder == RT_BUILD_LBVH_GPU) {
        // Stage the host triangles, then build on the device.
        float *d = nullptr;
        size_t bytes = static_cast<size_t>(n) * 9 * sizeof(float);
        if (n) {
            if ((e = rtx::dev_malloc(&d, bytes)) != hipSuccess) return hip_fail(ctx, e, "staging");
            if ((e = hipMemcpy(d, tri_xyz, bytes, hipMemcpyHostToDevice)) != hipSuccess) {
                (void)hipFree(d);
                return hip_fail(ctx, e, "staging copy");
            }
        }
        int rc = rt_scene_create_device(ctx, d, n, &o, nullptr, out);
        if (d) {
            (void)hipDeviceSynchronize();
            (void)hipFree(d);
        }
        return rc;
    }
    rtx::HostBvh h;
    try {
        if (!rtx::build_host(tri_xyz, n, o, h, err)) return fail(ctx, RT_ERR_INVALID, "%s", err.c_str());
    } catch (const std::bad_alloc &) {
        return fail(ctx, RT_ERR_OOM, "host allocation failed during build");
    }
    return upload_scene(ctx, h, tri_xyz, o.keep_source != 0, out);
}

int rt_scene_create_packed(rt_ctx ctx, const void *triangles, uint32_t n, const rt_build_opts *opts,
                           rt_scene *out) {
    if (!ctx || !out) return fail_global(RT_ERR_INVALID, "null argument");
    *out = nullptr;
    if (n && !triangles) return fail(ctx, RT_ERR_INVALID, "null triangles");
    // Triangle, std430 (primitive.glsl:28-38): vec3 p0 @0, uint n0 @12,
    // vec3 p1 @16, uint n1 @28, vec3 p2 @32, uint n2 @44.
    std::vector<float> xyz;
    std::vector<uint32_t> nrm;
    try {
        xyz.resize(static_cast<size_t>(n) * 9);
        nrm.resize(static_cast<size_t>(n) * 3);
    } catch (const std::bad_alloc &) {
        return fail(ctx, RT_ERR_OOM, "host allocation failed");
    }
    const unsigned char *src = static_cast<const unsigned char *>(triangles);
    for (size_t i = 0; i < n; ++i) {
        const unsigned char *r = src + i * 48;
        for (int v = 0; v < 3; ++v) {
            std::memcpy(&xyz[i * 9 + 3 * v], r + 16 * v, 12);
            std::memcpy(&nrm[i * 3 + v], r + 16 * v + 12, 4);
        }
    }
    int rc = rt_scene_create(ctx, xyz.data(), n, opts, out);
    if (rc != RT_OK || n == 0) return rc;
    rt_scene s = *out;
    hipError_t e = rtx::dev_malloc(&s->d_normals, nrm.size() * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemcpy(s->d_normals, nrm.data(), nrm.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        rt_scene_destroy(s);
        *out = nullptr;
        return hip_fail(ctx, e, "normal upload");
    }
    return RT_OK;
}

int rt_scene_from_bvh(rt_ctx ctx, rt_bvh bvh, const float *tri_xyz, rt_scene *out) {
    if (!ctx || !bvh || !out) return fail_global(RT_ERR_INVALID, "null argument");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    return upload_scene(ctx, bvh->h, tri_xyz, tri_xyz != nullptr, out);
}

int rt_scene_create_device(rt_ctx ctx, const float *d_tri_xyz, uint32_t n, const rt_build_opts *opts,
                           void *stream, rt_scene *out) {
    if (!ctx || !out) return fail_global(RT_ERR_INVALID, "null argument");
    *out = nullptr;
    rt_build_opts o;
    std::string err;
    if (!rtx::resolve_opts(opts, o, err)) return fail(ctx, RT_ERR_INVALID, "%s", err.c_str());
    if (o.builder != RT_BUILD_LBVH_GPU)
        return fail(ctx, RT_ERR_INVALID, "rt_scene_create_device needs builder RT_BUILD_LBVH_GPU");
    if (n && !d_tri_xyz) return fail(ctx, RT_ERR_INVALID, "null triangles");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    rt_scene s = new (std::nothrow) rt_scene_s;
    if (!s) return fail(ctx, RT_ERR_OOM, "host allocation failed");
    s->ctx = ctx;
    rtx::LbvhResult res;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        if (!ctx->lbvh) ctx->lbvh = rtx::lbvh_workspace_new();
    }
    std::string why;
    int rc = rtx::lbvh_build_device(ctx->lbvh, d_tri_xyz, n, o, static_cast<hipStream_t>(stream), res, why);
    if (rc != RT_OK) {
        delete s;
        return fail(ctx, rc, "LBVH build: %s", why.c_str());
    }
    s->d_nodes = res.d_nodes;
    s->d_tris = res.d_tris;
    s->d_src = res.d_src;
    s->dev.nodes = res.d_nodes;
    s->dev.tris = static_cast<const float4 *>(res.d_tris);
    s->dev.src = res.d_src;
    s->dev.num_tris = n;
    s->dev.width = o.node_width;
    s->dev.root = res.root_ref;
    s->dev.max_stack = res.max_stack;
    std::memset(&s->info, 0, sizeof s->info);
    s->info.num_tris = n;
    s->info.node_width = o.node_width;
    s->info.num_nodes = res.num_nodes;
    s->info.root_ref = res.root_ref;
    s->info.max_depth = res.max_depth;
    s->info.max_stack = res.max_stack;
    s->info.num_leaves = res.num_leaves;
    s->info.node_bytes = o.node_width == 8 ? 128u : 64u;
    s->info.nodes_total_bytes = static_cast<uint64_t>(res.num_nodes) * s->info.node_bytes;
    s->info.tris_total_bytes = static_cast<uint64_t>(n) * sizeof(rtx::TriRec);
    for (int k = 0; k < 3; ++k) {
        s->info.bounds_lo[k] = res.lo[k];
        s->info.bounds_hi[k] = res.hi[k];
        s->dev.lo[k] = res.lo[k];
        s->dev.hi[k] = res.hi[k];
    }
    s->info.build_ms = res.build_ms;
    *out = s;
    return RT_OK;
}

int rt_lbvh_binary(rt_ctx ctx, const float *d_tri_xyz, uint32_t n, uint32_t *morton, uint32_t *order,
                   uint32_t *left, uint32_t *right, float *boxes, void *stream) {
    if (!ctx) return fail_global(RT_ERR_INVALID, "null context");
    if (n && (!d_tri_xyz || !morton || !order || !boxes || (n > 1 && (!left || !right))))
        return fail(ctx, RT_ERR_INVALID, "null buffer");
    if (n > rtx::kMaxLeafOffset) return fail(ctx, RT_ERR_INVALID, "too many triangles");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        if (!ctx->lbvh) ctx->lbvh = rtx::lbvh_workspace_new();
    }
    std::string why;
    int rc = rtx::lbvh_binary(ctx->lbvh, d_tri_xyz, n, static_cast<hipStream_t>(stream), morton, order, left, right,
                              boxes, why);
    if (rc != RT_OK) return fail(ctx, rc, "LBVH: %s", why.c_str());
    return RT_OK;
}

int rt_scene_info_get(rt_scene scene, rt_scene_info *info) {
    if (!scene || !info) return fail_global(RT_ERR_INVALID, "null argument");
    *info = scene->info;
    return RT_OK;
}

int rt_scene_download(rt_scene scene, void *nodes, void *tris) {
    if (!scene) return fail_global(RT_ERR_INVALID, "null scene");
    rt_ctx ctx = scene->ctx;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    if (nodes && scene->info.nodes_total_bytes &&
        (e = hipMemcpy(nodes, scene->d_nodes, scene->info.nodes_total_bytes, hipMemcpyDeviceToHost)) !=
            hipSuccess)
        return hip_fail(ctx, e, "node download");
    if (tris && scene->info.tris_total_bytes &&
        (e = hipMemcpy(tris, scene->d_tris, scene->info.tris_total_bytes, hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(ctx, e, "triangle download");
    return RT_OK;
}

int rt_scene_set_analytic(rt_ctx ctx, rt_scene scene, const float *spheres, uint32_t ns, const float *cubes,
                          uint32_t nc, const float *planes, uint32_t np) {
    if (!ctx || !scene) return fail_global(RT_ERR_INVALID, "null context or scene");
    if (scene->ctx != ctx) return fail(ctx, RT_ERR_INVALID, "scene belongs to another context");
    if ((ns && !spheres) || (nc && !cubes) || (np && !planes)) return fail(ctx, RT_ERR_INVALID, "null array");
    // Object ids must stay below RT_MISS.
    const uint64_t total = static_cast<uint64_t>(ns) + nc + np;
    if (static_cast<uint64_t>(scene->info.num_tris) + total >= RT_MISS)
        return fail(ctx, RT_ERR_INVALID, "too many objects for 32-bit ids");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    // One buffer: spheres and planes as float4, then cubes as 6 floats.
    const size_t sph_bytes = static_cast<size_t>(ns) * 16, pl_bytes = static_cast<size_t>(np) * 16;
    const size_t cube_bytes = static_cast<size_t>(nc) * 24;
    const size_t bytes = sph_bytes + pl_bytes + cube_bytes;
    std::vector<char> staging(bytes);
    if (ns) std::memcpy(staging.data(), spheres, sph_bytes);
    if (np) std::memcpy(staging.data() + sph_bytes, planes, pl_bytes);
    if (nc) std::memcpy(staging.data() + sph_bytes + pl_bytes, cubes, cube_bytes);
    std::lock_guard<std::mutex> g(ctx->launch_mu);
    rtx::AnaDev a;
    int idx = -1;
    if (total) {
        // A buffer that is not published and whose launches have all run;
        // else a new one (launches queued on other streams may still read
        // every other buffer).
        for (size_t i = 0; i < scene->ana_bufs.size(); ++i) {
            rt_scene_s::AnaBuf &b = scene->ana_bufs[i];
            if (static_cast<int>(i) == scene->ana_cur || b.stuck) continue;
            for (size_t k = 0; k < b.uses.size();) {
                if (hipEventQuery(b.uses[k]) != hipErrorNotReady) {
                    ctx->ev_pool.push_back(b.uses[k]);
                    b.uses[k] = b.uses.back();
                    b.uses.pop_back();
                } else {
                    ++k;
                }
            }
            if (b.uses.empty() && b.cap >= bytes && (idx < 0 || b.cap < scene->ana_bufs[idx].cap))
                idx = static_cast<int>(i);
        }
        if (idx < 0) {
            rt_scene_s::AnaBuf b;
            b.cap = (bytes + 255) / 256 * 256;
            if ((e = rtx::dev_malloc(&b.d, b.cap)) != hipSuccess) return hip_fail(ctx, e, "analytic allocation");
            scene->ana_bufs.push_back(b);
            idx = static_cast<int>(scene->ana_bufs.size()) - 1;
        }
        if (!ctx->upload && (e = hipStreamCreateWithFlags(&ctx->upload, hipStreamNonBlocking)) != hipSuccess)
            return hip_fail(ctx, e, "upload stream");
        // Only the upload stream is waited for: no device-wide synchronisation.
        if ((e = hipMemcpyAsync(scene->ana_bufs[idx].d, staging.data(), bytes, hipMemcpyHostToDevice, ctx->upload)) !=
                hipSuccess ||
            (e = hipStreamSynchronize(ctx->upload)) != hipSuccess)
            return hip_fail(ctx, e, "analytic upload");
        char *p = static_cast<char *>(scene->ana_bufs[idx].d);
        a.spheres = ns ? reinterpret_cast<const float4 *>(p) : nullptr;
        a.planes = np ? reinterpret_cast<const float4 *>(p + sph_bytes) : nullptr;
        a.cubes = nc ? reinterpret_cast<const float *>(p + sph_bytes + pl_bytes) : nullptr;
        a.num_spheres = ns;
        a.num_cubes = nc;
        a.num_planes = np;
        a.count = static_cast<uint32_t>(total);
    }
    a.gen = ++ctx->ana_gen;
    scene->ana_cur = idx;
    scene->dev.ana = a;
    scene->info.num_spheres = ns;
    scene->info.num_cubes = nc;
    scene->info.num_planes = np;
    return RT_OK;
}

int rt_scene_destroy(rt_scene scene) {
    if (!scene) return RT_OK;
    (void)hipSetDevice(scene->ctx->device);
    (void)hipDeviceSynchronize();
    if (scene->d_nodes) (void)hipFree(scene->d_nodes);
    if (scene->d_tris) (void)hipFree(scene->d_tris);
    if (scene->d_src) (void)hipFree(scene->d_src);
    {
        std::lock_guard<std::mutex> g(scene->ctx->launch_mu);
        for (rt_scene_s::AnaBuf &b : scene->ana_bufs) {
            for (hipEvent_t ev : b.uses) scene->ctx->ev_pool.push_back(ev);
            if (b.d) (void)hipFree(b.d);
        }
        scene->ana_bufs.clear();
    }
    if (scene->d_normals) (void)hipFree(scene->d_normals);
    delete scene;
    return RT_OK;
}

int rt_trace_closest(rt_ctx ctx, rt_scene scene, const rt_ray *rays, uint32_t n, rt_hit *hits, void *stream) {
    return trace_common(ctx, scene, rtx::Query::Closest, rays, n, hits, nullptr, nullptr, nullptr, nullptr,
                        stream);
}

int rt_trace_occluded(rt_ctx ctx, rt_scene scene, const rt_ray *rays, uint32_t n, uint8_t *flags, void *stream) {
    if (!flags && n) return fail(ctx, RT_ERR_INVALID, "null flags");
    return trace_common(ctx, scene, rtx::Query::Occluded, rays, n, nullptr, flags, nullptr, nullptr, nullptr,
                        stream);
}

int rt_trace_occluded_mask(rt_ctx ctx, rt_scene scene, const rt_ray *rays, uint32_t n, uint64_t *masks,
                           void *stream) {
    if (!masks && n) return fail(ctx, RT_ERR_INVALID, "null masks");
    return trace_common(ctx, scene, rtx::Query::Occluded, rays, n, nullptr, nullptr, masks, nullptr, nullptr,
                        stream);
}

int rt_trace_closest_noalloc(rt_ctx ctx, rt_scene scene, const rt_ray *rays, uint32_t n, rt_hit *hits,
                             void *stream) {
    if (!ctx || !scene) return fail_global(RT_ERR_INVALID, "null context or scene");
    if (scene->ctx != ctx) return fail(ctx, RT_ERR_INVALID, "scene belongs to another context");
    if (n == 0) return RT_OK;
    if (!rays || !hits) return fail(ctx, RT_ERR_INVALID, "null rays or hits");
    unsigned char *base = static_cast<unsigned char *>(rtx::noalloc_scratch(ctx->device));
    if (!base) return fail(ctx, RT_ERR_INVALID, "no allocation-free scratch on device %d", ctx->device);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    hipStream_t st = static_cast<hipStream_t>(stream);
    FixedScratch fx;
    fx.work = reinterpret_cast<uint32_t *>(base);
    fx.stack = reinterpret_cast<uint2 *>(base + rtx::kNoallocHeaderBytes);
    fx.entries = rtx::kNoallocSpillEntries;
    rtx::TraceLaunch a;
    a.rays = rays;
    a.n = n;
    a.hits = hits;
    rtx::LaunchShape shape;
    std::lock_guard<std::mutex> g(ctx->launch_mu);
    int rc = prepare_launch(ctx, scene, rtx::Query::Closest, false, false, n, nullptr, st, shape, a, &fx);
    if (rc != RT_OK) return rc;
    e = static_cast<hipError_t>(rtx::noalloc_clear_work(fx.work, st));
    if (e != hipSuccess) return hip_fail(ctx, e, "scratch clear");
    e = static_cast<hipError_t>(rtx::launch_trace(scene->dev, rtx::Query::Closest, a, shape, st));
    if (e != hipSuccess) return hip_fail(ctx, e, "trace launch");
    rtx::ana_hold(ctx, scene, st);
    return RT_OK;
}

int rt_trace_closest_ex(rt_ctx ctx, rt_scene scene, const rt_ray *rays, uint32_t n, rt_hit *hits,
                        const rt_trace_opts *opts, uint32_t *counts, void *stream) {
    return trace_common(ctx, scene, rtx::Query::Closest, rays, n, hits, nullptr, nullptr, opts, counts, stream);
}

int rt_trace_occluded_ex(rt_ctx ctx, rt_scene scene, const rt_ray *rays, uint32_t n, uint8_t *flags,
                         uint64_t *masks, const rt_trace_opts *opts, uint32_t *counts, void *stream) {
    return trace_common(ctx, scene, rtx::Query::Occluded, rays, n, nullptr, flags, masks, opts, counts, stream);
}

int rt_trace_closest_brute(rt_ctx ctx, rt_scene scene, const rt_ray *rays, uint32_t n, rt_hit *hits,
                           void *stream) {
    if (!ctx || !scene) return fail_global(RT_ERR_INVALID, "null context or scene");
    if (n == 0) return RT_OK;
    if (!rays || !hits) return fail(ctx, RT_ERR_INVALID, "null rays or hits");
    if (!scene->dev.src && scene->dev.num_tris) return fail(ctx, RT_ERR_INVALID, "scene built without keep_source");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    std::lock_guard<std::mutex> g(ctx->launch_mu);
    e = static_cast<hipError_t>(
        rtx::launch_brute(scene->dev, rtx::Query::Closest, rays, n, hits, nullptr, static_cast<hipStream_t>(stream)));
    if (e != hipSuccess) return hip_fail(ctx, e, "brute launch");
    rtx::ana_hold(ctx, scene, static_cast<hipStream_t>(stream));
    return RT_OK;
}

int rt_trace_occluded_brute(rt_ctx ctx, rt_scene scene, const rt_ray *rays, uint32_t n, uint8_t *flags,
                            void *stream) {
    if (!ctx || !scene) return fail_global(RT_ERR_INVALID, "null context or scene");
    if (n == 0) return RT_OK;
    if (!rays || !flags) return fail(ctx, RT_ERR_INVALID, "null rays or flags");
    if (!scene->dev.src && scene->dev.num_tris) return fail(ctx, RT_ERR_INVALID, "scene built without keep_source");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    std::lock_guard<std::mutex> g(ctx->launch_mu);
    e = static_cast<hipError_t>(rtx::launch_brute(scene->dev, rtx::Query::Occluded, rays, n, nullptr, flags,
                                                  static_cast<hipStream_t>(stream)));
    if (e != hipSuccess) return hip_fail(ctx, e, "brute launch");
    rtx::ana_hold(ctx, scene, static_cast<hipStream_t>(stream));
    return RT_OK;
}

int rt_batch_create(rt_ctx ctx, const rt_scene *scenes, uint32_t nscenes, rt_batch *out) {
    if (!ctx || !out || (nscenes && !scenes)) return fail_global(RT_ERR_INVALID, "null argument");
    *out = nullptr;
    if (nscenes == 0) return fail(ctx, RT_ERR_INVALID, "a batch needs at least one scene");
    std::vector<rtx::BatchScene> table(nscenes);
    std::vector<rtx::BatchAna> ana(nscenes);
    bool has_ana = false;
    uint32_t width = 0, max_stack = 0;
    std::unique_lock<std::mutex> snap(ctx->launch_mu);  // the scenes' analytic sets as of now
    for (uint32_t i = 0; i < nscenes; ++i) {
        rt_scene sc = scenes[i];
        if (!sc || sc->ctx != ctx) return fail(ctx, RT_ERR_INVALID, "scene %u is null or from another context", i);
        if (width == 0) width = sc->dev.width;
        if (sc->dev.width != width) return fail(ctx, RT_ERR_INVALID, "all scenes of a batch need one node width");
        table[i].nodes = sc->dev.nodes;
        table[i].tris = sc->dev.tris;
        table[i].root = sc->dev.root;
        table[i].pad = 0;
        ana[i].ana = sc->dev.ana;
        ana[i].num_tris = sc->dev.num_tris;
        has_ana = has_ana || sc->dev.ana.count != 0;
        max_stack = std::max(max_stack, sc->dev.max_stack);
    }
    snap.unlock();
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    rt_batch b = new (std::nothrow) rt_batch_s;
    if (!b) return fail(ctx, RT_ERR_OOM, "host allocation failed");
    b->ctx = ctx;
    b->width = width;
    b->nscenes = nscenes;
    b->max_stack = max_stack;
    b->scenes.assign(scenes, scenes + nscenes);
    for (const rtx::BatchAna &x : ana) b->ana_snap.push_back(x.ana);
    if ((e = rtx::dev_malloc(&b->d_scenes, nscenes * sizeof(rtx::BatchScene))) != hipSuccess ||
        (e = hipMemcpy(b->d_scenes, table.data(), nscenes * sizeof(rtx::BatchScene), hipMemcpyHostToDevice)) !=
            hipSuccess ||
        (has_ana && ((e = rtx::dev_malloc(&b->d_ana, nscenes * sizeof(rtx::BatchAna))) != hipSuccess ||
                     (e = hipMemcpy(b->d_ana, ana.data(), nscenes * sizeof(rtx::BatchAna), hipMemcpyHostToDevice)) !=
                         hipSuccess))) {
        if (b->d_scenes) (void)hipFree(b->d_scenes);
        if (b->d_ana) (void)hipFree(b->d_ana);
        delete b;
        return hip_fail(ctx, e, "batch scene table");
    }
    *out = b;
    return RT_OK;
}

int rt_batch_trace(rt_ctx ctx, rt_batch batch, const rt_query *queries, uint32_t nq, const rt_trace_opts *opts,
                   void *stream) {
    if (!ctx || !batch) return fail_global(RT_ERR_INVALID, "null context or batch");
    if (batch->ctx != ctx) return fail(ctx, RT_ERR_INVALID, "batch belongs to another context");
    if (nq && !queries) return fail(ctx, RT_ERR_INVALID, "null queries");
    std::vector<rtx::BatchQuery> q;
    q.reserve(nq);
    uint64_t total = 0;
    for (uint32_t i = 0; i < nq; ++i) {
        const rt_query &in = queries[i];
        if (in.n == 0) continue;
        if (in.scene >= batch->nscenes) return fail(ctx, RT_ERR_INVALID, "query %u: scene %u out of range", i, in.scene);
        if (in.kind > 1) return fail(ctx, RT_ERR_INVALID, "query %u: unknown kind %u", i, in.kind);
        if (!in.rays || !in.out) return fail(ctx, RT_ERR_INVALID, "query %u: null buffer", i);
        rtx::BatchQuery b;
        b.scene = in.scene;
        b.kind = in.kind;
        b.n = in.n;
        b.rays = in.rays;
        b.out = in.out;
        q.push_back(b);
        total += in.n;
    }
    // The launch's waves sweep the index space from its start to its end, so
    // the rays placed last are the ones the launch drains on.  Expensive
    // queries go first: closest hit before occlusion (whose rays stop at the
    // first hit), larger scenes before smaller ones.  Each query writes only
    // its own output, so the order changes no result.
    std::stable_sort(q.begin(), q.end(), [&](const rtx::BatchQuery &x, const rtx::BatchQuery &y) {
        if (x.kind != y.kind) return x.kind < y.kind;
        return batch->scenes[x.scene]->dev.num_tris > batch->scenes[y.scene]->dev.num_tris;
    });
    {
        uint32_t start = 0;
        for (rtx::BatchQuery &b : q) {
            b.start = start;
            start += b.n;
        }
    }
    if (q.empty()) return RT_OK;
    if (total > 0x7FFFFFFFull) return fail(ctx, RT_ERR_INVALID, "at most 2^31-1 rays per batch launch");
    std::lock_guard<std::mutex> launch(ctx->launch_mu);
    for (uint32_t i = 0; i < batch->nscenes; ++i)
        if (!rtx::same_analytic(batch->scenes[i]->dev.ana, batch->ana_snap[i]))
            return fail(ctx, RT_ERR_INVALID,
                        "scene %u's analytic primitives changed after the batch was made; make a new batch", i);
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    hipStream_t st = static_cast<hipStream_t>(stream);
    rtx::BatchLaunch a;
    rtx::LaunchShape shape;
    int rc = rtx::batch_prepare(ctx, batch->width, batch->max_stack, static_cast<uint32_t>(total), opts, st, shape, a);
    if (rc != RT_OK) return rc;
    a.scenes = batch->d_scenes;
    a.nq = static_cast<uint32_t>(q.size());
    {
        std::lock_guard<std::mutex> g(batch->mu);
        rt_batch_s::Table &t = batch->tables[st];
        if (!t.ev[0]) {
            for (hipEvent_t &ev : t.ev)
                if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess)
                    return hip_fail(ctx, e, "query table event");
        }
        if (t.cap < q.size()) {
            if (t.cap && (e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(ctx, e, "query table release");
            free_table_buffers(t);
            const size_t cap = std::max<size_t>(q.size(), 64);
            const size_t bytes = cap * sizeof(rtx::BatchQuery);
            if ((e = rtx::dev_malloc(&t.d, bytes)) != hipSuccess ||
                (e = rtx::host_malloc_pinned(&t.h[0], bytes)) != hipSuccess ||
                (e = rtx::host_malloc_pinned(&t.h[1], bytes)) != hipSuccess) {
                free_table_buffers(t);
                return hip_fail(ctx, e, "query table");
            }
            t.cap = cap;
        }
        const int slot = t.next;
        t.next ^= 1;
        if (t.used[slot] && (e = hipEventSynchronize(t.ev[slot])) != hipSuccess)
            return hip_fail(ctx, e, "query table slot");
        std::memcpy(t.h[slot], q.data(), q.size() * sizeof(rtx::BatchQuery));
        if ((e = hipMemcpyAsync(t.d, t.h[slot], q.size() * sizeof(rtx::BatchQuery), hipMemcpyHostToDevice, st)) !=
                hipSuccess ||
            (e = hipEventRecord(t.ev[slot], st)) != hipSuccess)
            return hip_fail(ctx, e, "query table upload");
        t.used[slot] = true;
        a.queries = t.d;
    }
    e = static_cast<hipError_t>(rtx::launch_batch(batch->width, a, shape, st));
    if (e != hipSuccess) return hip_fail(ctx, e, "batch launch");
    if (batch->d_ana) {
        e = static_cast<hipError_t>(rtx::launch_batch_analytic(a, batch->d_ana, st));
        if (e != hipSuccess) return hip_fail(ctx, e, "batch analytic launch");
        for (rt_scene sc : batch->scenes) rtx::ana_hold(ctx, sc, st);
    }
    return RT_OK;
}

int rt_batch_destroy(rt_batch batch) {
    if (!batch) return RT_OK;
    (void)hipSetDevice(batch->ctx->device);
    (void)hipDeviceSynchronize();
    if (batch->d_scenes) (void)hipFree(batch->d_scenes);
    if (batch->d_ana) (void)hipFree(batch->d_ana);
    for (auto &kv : batch->tables) {
        free_table_buffers(kv.second);
        for (hipEvent_t ev : kv.second.ev)
            if (ev) (void)hipEventDestroy(ev);
    }
    delete batch;
    return RT_OK;
}

}  // extern "C"

// ---- autotuner ----------------------------------------------------------
// Launch shapes change speed only, so the tuner simply times candidates on a
// caller's ray sample with HIP events and keeps the fastest.
extern "C" int rt_autotune(rt_ctx ctx, rt_scene scene, uint32_t kind, const rt_ray *rays, uint32_t n, void *out,
                           uint32_t iters, rt_trace_opts *best, float *best_ms, void *stream) {
    if (!ctx || !scene || !best) return fail_global(RT_ERR_INVALID, "null argument");
    if (kind > 1) return fail(ctx, RT_ERR_INVALID, "unknown query kind %u", kind);
    if (n == 0 || !rays || !out) return fail(ctx, RT_ERR_INVALID, "the tuner needs a non-empty ray sample");
    if (iters == 0) iters = 5;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    hipStream_t st = static_cast<hipStream_t>(stream);
    std::vector<rt_trace_opts> cands;
    for (uint32_t block : {256u, 128u})
        for (uint32_t lds : {8u, 16u, 32u}) {
            rt_trace_opts o;
            rt_trace_opts_default(&o);
            o.block_size = block;
            o.lds_stack = lds;
            o.schedule = 1;
            cands.push_back(o);
            for (uint32_t chunk : {32u, 64u, 128u})
                for (uint32_t refill : {4u, 16u, 32u}) {
                    o.schedule = 0;
                    o.chunk = chunk;
                    o.refill = refill;
                    cands.push_back(o);
                }
        }
    hipEvent_t e0, e1;
    if ((e = hipEventCreate(&e0)) != hipSuccess) return hip_fail(ctx, e, "hipEventCreate");
    if ((e = hipEventCreate(&e1)) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_fail(ctx, e, "hipEventCreate");
    }
    float best_t = INFINITY;
    int rc = RT_OK;
    for (const rt_trace_opts &o : cands) {
        auto launch = [&]() {
            return kind == 0 ? rt_trace_closest_ex(ctx, scene, rays, n, static_cast<rt_hit *>(out), &o, nullptr, stream)
                             : rt_trace_occluded_ex(ctx, scene, rays, n, static_cast<uint8_t *>(out), nullptr, &o,
                                                    nullptr, stream);
        };
        if ((rc = launch()) != RT_OK) break;  // warm-up (allocates stream resources)
        if ((e = hipEventRecord(e0, st)) != hipSuccess) { rc = hip_fail(ctx, e, "hipEventRecord"); break; }
        for (uint32_t k = 0; k < iters && rc == RT_OK; ++k) rc = launch();
        if (rc != RT_OK) break;
        if ((e = hipEventRecord(e1, st)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess) {
            rc = hip_fail(ctx, e, "timing");
            break;
        }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms /= static_cast<float>(iters);
        if (ms < best_t) {
            best_t = ms;
            *best = o;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != RT_OK) return rc;
    if (best_ms) *best_ms = best_t;
    return RT_OK;
}

// Batch launches: block 128/256 x LDS ring 8/16/32 x chunk 64/128/256 x
// refill 16/24/32, timed on the caller's queries.  On the 64-scene batch the
// fastest shapes sit at large chunks and refills (chunk 256, refill 24);
// chunk 32 and refill 4 were never chosen.
extern "C" int rt_batch_autotune(rt_ctx ctx, rt_batch batch, const rt_query *queries, uint32_t nq, uint32_t iters,
                                 rt_trace_opts *best, float *best_ms, void *stream) {
    if (!ctx || !batch || !best) return fail_global(RT_ERR_INVALID, "null argument");
    if (iters == 0) iters = 3;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipEvent_t e0, e1;
    if ((e = hipEventCreate(&e0)) != hipSuccess) return hip_fail(ctx, e, "hipEventCreate");
    if ((e = hipEventCreate(&e1)) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_fail(ctx, e, "hipEventCreate");
    }
    float best_t = INFINITY;
    int rc = RT_OK;
    for (uint32_t block : {256u, 128u})
        for (uint32_t lds : {8u, 16u, 32u})
            for (uint32_t chunk : {64u, 128u, 256u})
                for (uint32_t refill : {16u, 24u, 32u}) {
                    if (rc != RT_OK) break;
                    rt_trace_opts o;
                    rt_trace_opts_default(&o);
                    o.block_size = block;
                    o.lds_stack = lds;
                    o.chunk = chunk;
                    o.refill = refill;
                    o.schedule = 0;
                    if ((rc = rt_batch_trace(ctx, batch, queries, nq, &o, stream)) != RT_OK) break;
                    if ((e = hipEventRecord(e0, st)) != hipSuccess) {
                        rc = hip_fail(ctx, e, "hipEventRecord");
                        break;
                    }
                    for (uint32_t k = 0; k < iters && rc == RT_OK; ++k)
                        rc = rt_batch_trace(ctx, batch, queries, nq, &o, stream);
                    if (rc != RT_OK) break;
                    if ((e = hipEventRecord(e1, st)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess) {
                        rc = hip_fail(ctx, e, "timing");
                        break;
                    }
                    float ms = 0.f;
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    ms /= static_cast<float>(iters);
                    if (ms < best_t) {
                        best_t = ms;
                        *best = o;
                    }
                }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != RT_OK) return rc;
    if (best_ms) *best_ms = best_t;
    return RT_OK;
}

// ---- callers ------------------------------------------------------------

namespace {

int caller_run(rt_ctx ctx, rt_scene scene, rtx::CallerKind kind, rtx::CallerLaunch &a, uint32_t units,
               hipStream_t st) {
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    rtx::CallerFn fn = rtx::caller_kernel(scene->dev.width, kind);
    if (!fn) return fail(ctx, RT_ERR_UNSUPPORTED, "node width not built");
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void *>(fn), 256, 0) !=
        hipSuccess || occ <= 0)
        occ = 1;
    const uint64_t want = (static_cast<uint64_t>(units) + 3) / 4;  // one wave per unit, 4 waves per block
    const uint32_t blocks =
        static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(want, static_cast<uint64_t>(ctx->num_cu) * occ)));
    const uint32_t S = rtx::caller_lds_stack(kind);
    a.gstack_depth = (scene->dev.max_stack > S ? scene->dev.max_stack - S : 0) + 1;
    a.err = ctx->d_err;
    uint32_t *work = nullptr;
    std::lock_guard<std::mutex> g(ctx->launch_mu);
    int rc = rtx::stream_resources(ctx, st, static_cast<size_t>(a.gstack_depth) * blocks * 256, &a.gstack, &work);
    if (rc != RT_OK) return rc;
    e = static_cast<hipError_t>(rtx::launch_caller(scene->dev.width, kind, scene->dev, a, blocks, st));
    if (e != hipSuccess) return hip_fail(ctx, e, "caller launch");
    rtx::ana_hold(ctx, scene, st);
    return RT_OK;
}

}  // namespace

namespace {

int raygen_run(rt_ctx ctx, rt_scene scene, rtx::CameraDev cam, uint32_t width, uint32_t height,
               const uint32_t *normals, float *dir_object, float *uv_normal, void *stream) {
    if (width == 0 || height == 0) return RT_OK;
    if (!dir_object || !uv_normal) return fail(ctx, RT_ERR_INVALID, "null output image");
    if (static_cast<uint64_t>(width) * height > 0x7FFFFFFFull) return fail(ctx, RT_ERR_INVALID, "image too large");
    rtx::CallerLaunch a;
    a.cam = cam;
    a.cam.inv_res[0] = 1.0f / static_cast<float>(width);
    a.cam.inv_res[1] = 1.0f / static_cast<float>(height);
    a.width = width;
    a.height = height;
    a.normals = normals ? normals : scene->d_normals;
    a.dir_object = reinterpret_cast<float4 *>(dir_object);
    a.uv_normal = reinterpret_cast<float4 *>(uv_normal);
    const uint32_t units = ((width + 15) / 16) * ((height + 3) / 4);
    return caller_run(ctx, scene, rtx::CallerKind::Raygen, a, units, static_cast<hipStream_t>(stream));
}

}  // namespace

extern "C" int rt_raygen_closest(rt_ctx ctx, rt_scene scene, const rt_camera *cam, uint32_t width, uint32_t height,
                                 const uint32_t *normals, float *dir_object, float *uv_normal, void *stream) {
    if (!ctx || !scene || !cam) return fail_global(RT_ERR_INVALID, "null argument");
    if (scene->ctx != ctx) return fail(ctx, RT_ERR_INVALID, "scene belongs to another context");
    rtx::CameraDev c{};
    for (int k = 0; k < 3; ++k) {
        c.eye[k] = cam->eye[k];
        c.p0[k] = cam->p0[k];
        c.p1[k] = cam->p1[k];
        c.p2[k] = cam->p2[k];
    }
    c.jitter[0] = cam->jitter[0];
    c.jitter[1] = cam->jitter[1];
    c.projection = RT_PROJ_DEFAULT;
    c.rand_jitter = 0;
    return raygen_run(ctx, scene, c, width, height, normals, dir_object, uv_normal, stream);
}

extern "C" int rt_raygen_primary(rt_ctx ctx, rt_scene scene, const rt_primary_camera *cam, uint32_t width,
                                 uint32_t height, const uint32_t *normals, float *dir_object, float *uv_normal,
                                 void *stream) {
    if (!ctx || !scene || !cam) return fail_global(RT_ERR_INVALID, "null argument");
    if (scene->ctx != ctx) return fail(ctx, RT_ERR_INVALID, "scene belongs to another context");
    if (cam->projection > RT_PROJ_STEREO_LR) return fail(ctx, RT_ERR_INVALID, "projection %u", cam->projection);
    if (cam->jitter_mode > 1) return fail(ctx, RT_ERR_INVALID, "jitter_mode %u", cam->jitter_mode);
    rtx::CameraDev c{};
    for (int k = 0; k < 3; ++k) {
        c.eye[k] = cam->eye[k];
        c.p0[k] = cam->p0[k];
        c.p1[k] = cam->p1[k];
        c.p2[k] = cam->p2[k];
        c.p3[k] = cam->p3[k];
        c.p4[k] = cam->p4[k];
        c.p5[k] = cam->p5[k];
    }
    c.ipd = cam->ipd;
    c.projection = cam->projection;
    c.rand_jitter = cam->jitter_mode == 0 ? 1u : 0u;
    c.seed[0] = cam->seed[0];
    c.seed[1] = cam->seed[1];
    c.jitter[0] = cam->jitter[0];
    c.jitter[1] = cam->jitter[1];
    return raygen_run(ctx, scene, c, width, height, normals, dir_object, uv_normal, stream);
}

extern "C" int rt_shadow_pass(rt_ctx ctx, rt_scene scene, const float *dir_object, const float eye[3],
                              const rt_light *light, const float seed[2], uint32_t width, uint32_t height,
                              uint32_t samples, uint64_t *tiles, void *stream) {
    if (!ctx || !scene || !eye || !light || !seed) return fail_global(RT_ERR_INVALID, "null argument");
    if (scene->ctx != ctx) return fail(ctx, RT_ERR_INVALID, "scene belongs to another context");
    if (width == 0 || height == 0 || samples == 0) return RT_OK;
    if (!dir_object || !tiles) return fail(ctx, RT_ERR_INVALID, "null buffer");
    const uint64_t words = static_cast<uint64_t>((width + 15) / 16) * ((height + 3) / 4) * samples;
    if (words > 0x7FFFFFFFull || static_cast<uint64_t>(width) * height > 0x7FFFFFFFull)
        return fail(ctx, RT_ERR_INVALID, "image too large");
    rtx::CallerLaunch a;
    for (int k = 0; k < 3; ++k) {
        a.cam.eye[k] = eye[k];
        a.light.pos[k] = light->pos[k];
    }
    a.cam.seed[0] = seed[0];
    a.cam.seed[1] = seed[1];
    a.light.rad_origin = light->rad_origin;
    a.light.dir[0] = light->dir[0];
    a.light.dir[1] = light->dir[1];
    a.light.color_type[0] = light->color_type[0];
    a.light.color_type[1] = light->color_type[1];
    a.width = width;
    a.height = height;
    a.samples = samples;
    a.dir_in = reinterpret_cast<const float4 *>(dir_object);
    a.tiles = reinterpret_cast<unsigned long long *>(tiles);
    return caller_run(ctx, scene, rtx::CallerKind::ShadowPass, a, static_cast<uint32_t>(words),
                      static_cast<hipStream_t>(stream));
}

extern "C" int rt_shadow_tiles(rt_ctx ctx, rt_scene scene, const rt_ray *rays, uint32_t width, uint32_t height,
                               uint32_t samples, uint64_t *tiles, void *stream) {
    if (!ctx || !scene) return fail_global(RT_ERR_INVALID, "null argument");
    if (scene->ctx != ctx) return fail(ctx, RT_ERR_INVALID, "scene belongs to another context");
    if (width == 0 || height == 0 || samples == 0) return RT_OK;
    if (!rays || !tiles) return fail(ctx, RT_ERR_INVALID, "null buffer");
    const uint64_t words = static_cast<uint64_t>((width + 15) / 16) * ((height + 3) / 4) * samples;
    if (words > 0x7FFFFFFFull || static_cast<uint64_t>(width) * height * samples > 0x7FFFFFFFull)
        return fail(ctx, RT_ERR_INVALID, "image too large");
    rtx::CallerLaunch a;
    a.width = width;
    a.height = height;
    a.samples = samples;
    a.rays = rays;
    a.tiles = reinterpret_cast<unsigned long long *>(tiles);
    return caller_run(ctx, scene, rtx::CallerKind::ShadowTiles, a, static_cast<uint32_t>(words),
                      static_cast<hipStream_t>(stream));
}

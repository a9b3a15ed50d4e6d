// C ABI of the wavefront path tracer: buffers, hipGraph capture and replay.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "alloc.hpp"
#include "internal.hpp"
#include "wavefront.hpp"

struct rt_pathtracer_s {
    rt_ctx ctx = nullptr;
    rt_scene scene = nullptr;
    rtx::PtLaunch a;
    uint32_t bounce_blocks = 0;
    hipStream_t stream = nullptr;  // private, non-blocking
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    float *graph_image = nullptr;
    void *bufs[8] = {};
    // Frame seeds are copied to the device from a pinned ring: slot k is
    // rewritten only once the copy that last read it has run (its event),
    // so a call returns without waiting for the device unless the host is
    // kSeedSlots frames ahead of it.
    static constexpr int kSeedSlots = 8;
    uint64_t *seed_ring = nullptr;  // pinned host memory
    hipEvent_t seed_ev[kSeedSlots] = {};
    bool seed_used[kSeedSlots] = {};
    int seed_next = 0;
};

namespace {

void release(rt_pathtracer p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    if (p->ev_in) (void)hipEventDestroy(p->ev_in);
    if (p->ev_out) (void)hipEventDestroy(p->ev_out);
    for (hipEvent_t &ev : p->seed_ev)
        if (ev) (void)hipEventDestroy(ev);
    if (p->seed_ring) (void)hipHostFree(p->seed_ring);
    for (void *b : p->bufs)
        if (b) (void)hipFree(b);
}

}  // namespace

extern "C" {

void rt_pt_opts_default(rt_pt_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->width = 1920;
    o->height = 1080;
    o->spp = 4;
    o->max_depth = 3;
    o->rr_survive = 0.8f;
    o->albedo = 0.7f;
}

int rt_pt_create(rt_ctx ctx, rt_scene scene, const rt_camera *cam, const rt_pt_opts *opts, rt_pathtracer *out) {
    if (!ctx || !scene || !cam || !out) return rtx::fail_global(RT_ERR_INVALID, "null argument");
    *out = nullptr;
    if (scene->ctx != ctx) return rtx::fail(ctx, RT_ERR_INVALID, "scene belongs to another context");
    if (scene->dev.ana.count)
        return rtx::fail(ctx, RT_ERR_UNSUPPORTED, "the path tracer takes triangle-only scenes");
    rt_pt_opts o;
    rt_pt_opts_default(&o);
    if (opts) o = *opts;
    const uint64_t npaths = static_cast<uint64_t>(o.width) * o.height * o.spp;
    if (npaths == 0 || npaths > 0x7FFFFFFFull) return rtx::fail(ctx, RT_ERR_INVALID, "bad image size or spp");
    if (o.max_depth > 64) return rtx::fail(ctx, RT_ERR_INVALID, "max_depth at most 64");
    if (!(o.rr_survive > 0.f && o.rr_survive <= 1.f)) return rtx::fail(ctx, RT_ERR_INVALID, "rr_survive in (0, 1]");
    // Measured on MI355X (1080p, 4 spp, 3 bounces, 2M-triangle split scene):
    // refill 4-8 is 1.3% faster per frame than 16, and chunks of 32, 48,
    // 96, 128 or 256 rays are 0.5-6% slower than 64.  Bounce rays are short,
    // so a wave gains from handing out new rays as soon as a few lanes idle.
    if (o.chunk == 0) o.chunk = 64;
    if (o.refill == 0) o.refill = 8;
    if (o.chunk < 16 || o.chunk > 4096 || o.refill > 64)
        return rtx::fail(ctx, RT_ERR_INVALID, "chunk in 16..4096 and refill in 1..64");
    rtx::PtBounceFn fn = rtx::pt_bounce_kernel(scene->dev.width);
    if (!fn) return rtx::fail(ctx, RT_ERR_UNSUPPORTED, "node width not built");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return rtx::hip_fail(ctx, e, "hipSetDevice");
    rt_pathtracer p = new (std::nothrow) rt_pathtracer_s;
    if (!p) return rtx::fail(ctx, RT_ERR_OOM, "host allocation failed");
    p->ctx = ctx;
    p->scene = scene;
    rtx::PtLaunch &a = p->a;
    for (int k = 0; k < 3; ++k) {
        a.cam.eye[k] = cam->eye[k];
        a.cam.p0[k] = cam->p0[k];
        a.cam.p1[k] = cam->p1[k];
        a.cam.p2[k] = cam->p2[k];
    }
    a.cam.jitter[0] = cam->jitter[0];
    a.cam.jitter[1] = cam->jitter[1];
    a.cam.inv_res[0] = 1.0f / static_cast<float>(o.width);
    a.cam.inv_res[1] = 1.0f / static_cast<float>(o.height);
    a.width = o.width;
    a.height = o.height;
    a.spp = o.spp;
    a.max_depth = o.max_depth;
    a.rr_survive = o.rr_survive;
    a.albedo = o.albedo;
    a.chunk = o.chunk;
    a.refill = o.refill;
    a.err = ctx->d_err;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void *>(fn), 256, 0) != hipSuccess ||
        occ <= 0)
        occ = 1;
    p->bounce_blocks = static_cast<uint32_t>(ctx->num_cu * (occ < 4 ? occ : 4));
    const rtx::DualStack ds = rtx::dual_stack(scene->dev.max_stack, 16, scene->dev.width, true);
    a.gstack_depth = ds.depth;
    a.dual_max = ds.dual_max;
    const size_t sizes[8] = {npaths * sizeof(rtx::PathState),
                             npaths * sizeof(rt_ray),
                             npaths * sizeof(rt_ray),
                             npaths * 4,
                             npaths * 4,
                             (o.max_depth + 2 + 4) * 4 + 64,
                             static_cast<size_t>(a.gstack_depth) * p->bounce_blocks * 256 * sizeof(uint2),
                             static_cast<size_t>(o.width) * o.height * 4};
    for (int i = 0; i < 8; ++i)
        if ((e = rtx::dev_malloc(&p->bufs[i], sizes[i])) != hipSuccess) {
            release(p);
            delete p;
            return rtx::hip_fail(ctx, e, "path tracer buffers");
        }
    a.paths = static_cast<rtx::PathState *>(p->bufs[0]);
    a.rays[0] = static_cast<rt_ray *>(p->bufs[1]);
    a.rays[1] = static_cast<rt_ray *>(p->bufs[2]);
    a.queue[0] = static_cast<uint32_t *>(p->bufs[3]);
    a.queue[1] = static_cast<uint32_t *>(p->bufs[4]);
    // Control words: the 64-bit frame seed first (8-byte aligned), then the
    // work counter (2 words, zero between launches), then the queue counts.
    a.frame_seed = static_cast<uint64_t *>(p->bufs[5]);
    a.work = static_cast<uint32_t *>(p->bufs[5]) + 2;
    a.counts = static_cast<uint32_t *>(p->bufs[5]) + 4;
    a.gstack = static_cast<uint2 *>(p->bufs[6]);
    // The control words are zeroed on the private stream that uses them.
    if ((e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipMemsetAsync(p->bufs[5], 0, sizes[5], p->stream)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&p->ev_in, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&p->ev_out, hipEventDisableTiming)) != hipSuccess ||
        (e = rtx::host_malloc_pinned(reinterpret_cast<void **>(&p->seed_ring), sizeof(uint64_t) * rt_pathtracer_s::kSeedSlots)) != hipSuccess) {
        release(p);
        delete p;
        return rtx::hip_fail(ctx, e, "path tracer setup");
    }
    for (hipEvent_t &ev : p->seed_ev)
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) {
            release(p);
            delete p;
            return rtx::hip_fail(ctx, e, "path tracer setup");
        }
    *out = p;
    return RT_OK;
}

int rt_pt_render(rt_pathtracer p, uint64_t frame_seed, float *image, int use_graph, void *stream) {
    if (!p || !image) return rtx::fail_global(RT_ERR_INVALID, "null argument");
    rt_ctx ctx = p->ctx;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return rtx::hip_fail(ctx, e, "hipSetDevice");
    hipStream_t user = static_cast<hipStream_t>(stream);
    // Order after the caller's stream, run on the private stream, and make
    // the caller's stream wait for the frame.
    if ((e = hipEventRecord(p->ev_in, user)) != hipSuccess || (e = hipStreamWaitEvent(p->stream, p->ev_in, 0)) != hipSuccess)
        return rtx::hip_fail(ctx, e, "stream order");
    // The seed goes through the next pinned ring slot; wait only if that
    // slot's previous copy is still queued.
    const int k = p->seed_next;
    p->seed_next = (k + 1) % rt_pathtracer_s::kSeedSlots;
    if (p->seed_used[k] && (e = hipEventSynchronize(p->seed_ev[k])) != hipSuccess)
        return rtx::hip_fail(ctx, e, "frame seed slot");
    p->seed_ring[k] = frame_seed;
    if ((e = hipMemcpyAsync(const_cast<uint64_t *>(p->a.frame_seed), p->seed_ring + k, 8, hipMemcpyHostToDevice,
                            p->stream)) != hipSuccess ||
        (e = hipEventRecord(p->seed_ev[k], p->stream)) != hipSuccess)
        return rtx::hip_fail(ctx, e, "frame seed");
    p->seed_used[k] = true;
    p->a.image = image;
    if (use_graph) {
        if (!p->exec || p->graph_image != image) {
            if (p->exec) (void)hipGraphExecDestroy(p->exec);
            if (p->graph) (void)hipGraphDestroy(p->graph);
            p->exec = nullptr;
            p->graph = nullptr;
            if ((e = hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
                return rtx::hip_fail(ctx, e, "graph capture");
            int rc = rtx::pt_enqueue_frame(p->scene->dev, p->a, p->bounce_blocks, p->stream);
            hipGraph_t g = nullptr;
            e = hipStreamEndCapture(p->stream, &g);
            if (rc != hipSuccess || e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return rtx::hip_fail(ctx, e != hipSuccess ? e : static_cast<hipError_t>(rc), "graph capture");
            }
            p->graph = g;
            if ((e = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0)) != hipSuccess)
                return rtx::hip_fail(ctx, e, "graph instantiate");
            p->graph_image = image;
        }
        if ((e = hipGraphLaunch(p->exec, p->stream)) != hipSuccess) return rtx::hip_fail(ctx, e, "graph launch");
    } else {
        e = static_cast<hipError_t>(rtx::pt_enqueue_frame(p->scene->dev, p->a, p->bounce_blocks, p->stream));
        if (e != hipSuccess) return rtx::hip_fail(ctx, e, "frame launch");
    }
    if ((e = hipEventRecord(p->ev_out, p->stream)) != hipSuccess || (e = hipStreamWaitEvent(user, p->ev_out, 0)) != hipSuccess)
        return rtx::hip_fail(ctx, e, "stream order");
    return RT_OK;
}

int rt_pt_counts(rt_pathtracer p, uint32_t *counts, uint32_t n) {
    if (!p || !counts) return rtx::fail_global(RT_ERR_INVALID, "null argument");
    const uint32_t m = std::min(n, p->a.max_depth + 1);
    hipError_t e = hipMemcpyAsync(counts, p->a.counts, m * 4, hipMemcpyDeviceToHost, p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) return rtx::hip_fail(p->ctx, e, "counts");
    return RT_OK;
}

int rt_pt_queue_rays(rt_pathtracer p, uint32_t depth, rt_ray *out, uint32_t n) {
    if (!p || (!out && n)) return rtx::fail_global(RT_ERR_INVALID, "null argument");
    // Depth d reads queue d & 1 and writes depth d + 1's rays into the other
    // one; the last depth writes none, so both queues still hold the last
    // two depths' rays.
    const uint32_t dmax = p->a.max_depth;
    if (depth > dmax || depth + 1u < dmax) return rtx::fail_global(RT_ERR_INVALID, "depth: only the last two are kept");
    if (static_cast<uint64_t>(n) > static_cast<uint64_t>(p->a.width) * p->a.height * p->a.spp)
        return rtx::fail_global(RT_ERR_INVALID, "n exceeds the queue");
    hipError_t e = hipSetDevice(p->ctx->device);
    if (e == hipSuccess && n)
        e = hipMemcpyAsync(out, p->a.rays[depth & 1u], static_cast<size_t>(n) * sizeof(rt_ray), hipMemcpyDefault,
                           p->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) return rtx::hip_fail(p->ctx, e, "queue rays");
    return RT_OK;
}

int rt_pt_destroy(rt_pathtracer p) {
    if (!p) return RT_OK;
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->stream);
    release(p);
    delete p;
    return RT_OK;
}

}  // extern "C"

// Internal definitions shared by the C ABI implementation files.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <vector>
#include <mutex>
#include <string>

#include "bvh_build.hpp"
#include "kernels.hpp"
#include "lbvh.hpp"
#include "rt/rt_c.h"

struct rt_ctx_s {
    int device = 0;
    int num_cu = 0;
    std::mutex mu;
    std::string err;
    uint32_t *d_err = nullptr;
    // Per-stream launch resources: the traversal spill stack and the work
    // counter of the dynamic schedule (two words, zero between launches).
    struct Spill {
        uint2 *buf = nullptr;
        size_t entries = 0;
        uint32_t *work = nullptr;
    };
    std::map<hipStream_t, Spill> spill;
    rtx::LbvhWorkspace *lbvh = nullptr;
    // Held from a launch's resources (spill stack, work counter, query
    // table) through its enqueue, and by rt_scene_set_analytic: launches
    // from several host threads, even on one stream, each get their own
    // consistent resources and analytic set.  Lock order: a server's mutex,
    // then launch_mu, then mu.
    std::mutex launch_mu;
    hipStream_t upload = nullptr;          // analytic uploads (non-blocking)
    uint64_t ana_gen = 0;                  // last analytic set number given out
    std::vector<hipEvent_t> ev_pool;       // idle events for analytic holds
};

struct rt_scene_s {
    rt_ctx ctx = nullptr;
    rtx::DevScene dev;
    rt_scene_info info{};
    void *d_nodes = nullptr;
    void *d_tris = nullptr;
    float *d_src = nullptr;
    // Analytic primitive sets (spheres | planes | cubes), one buffer per set
    // still in use: ana_cur is the published one (dev.ana points into it).
    // A launch that reads a set records an event in its buffer's `uses`; a
    // buffer is rewritten only when it is not published and all its uses
    // have completed, so launches queued before a replacement keep reading
    // the set they were issued with.  Guarded by ctx->launch_mu.
    struct AnaBuf {
        void *d = nullptr;
        size_t cap = 0;
        std::vector<hipEvent_t> uses;
        bool stuck = false;  // a use could not be recorded: never rewritten
    };
    std::vector<AnaBuf> ana_bufs;
    int ana_cur = -1;
    uint32_t *d_normals = nullptr;  // packed spheremap normals, 3 per triangle (optional)
};

struct rt_bvh_s {
    rtx::HostBvh h;
};

struct rt_batch_s {
    rt_ctx ctx = nullptr;
    uint32_t width = 4;
    uint32_t nscenes = 0;
    uint32_t max_stack = 0;
    rtx::BatchScene *d_scenes = nullptr;
    // The scenes' analytic primitives as the batch was made (d_ana, read by
    // the post-pass; null when no scene has any).  A trace checks that no
    // scene has replaced its set since: the table holds its buffers.
    rtx::BatchAna *d_ana = nullptr;
    std::vector<rt_scene> scenes;
    std::vector<rtx::AnaDev> ana_snap;
    // Per-stream query tables (grow-only): the device table the kernel
    // reads, and two pinned host slots its upload is staged from, so the
    // upload is asynchronous.  A slot is rewritten only after its previous
    // copy has run (its event): a call waits only when the host is two
    // calls ahead of the stream.
    struct Table {
        rtx::BatchQuery *d = nullptr;
        rtx::BatchQuery *h[2] = {nullptr, nullptr};
        hipEvent_t ev[2] = {nullptr, nullptr};
        bool used[2] = {false, false};
        int next = 0;
        size_t cap = 0;
    };
    std::mutex mu;
    std::map<hipStream_t, Table> tables;
};

namespace rtx {

// Records the message as the context's (and the thread's) last error and
// returns code.
int fail(rt_ctx ctx, int code, const char *fmt, ...);
int fail_global(int code, const char *fmt, ...);
int hip_fail(rt_ctx ctx, hipError_t e, const char *what);
void set_error(rt_ctx ctx, const std::string &msg);

// Whether two analytic sets are the same buffers and counts.
inline bool same_analytic(const AnaDev &x, const AnaDev &y) {
    return x.gen == y.gen && x.spheres == y.spheres && x.cubes == y.cubes && x.planes == y.planes &&
           x.num_spheres == y.num_spheres && x.num_cubes == y.num_cubes && x.num_planes == y.num_planes &&
           x.count == y.count;
}

// After enqueueing on `stream` a launch that reads the scene's published
// analytic set: records that use (caller holds ctx->launch_mu).
void ana_hold(rt_ctx ctx, rt_scene scene, hipStream_t stream);

// The stream's traversal spill stack (grown to at least `entries`) and its
// dynamic-schedule work counter.
int stream_resources(rt_ctx ctx, hipStream_t stream, size_t entries, uint2 **gstack, uint32_t **work);

// Launch shape, spill stack and work counter of one batch launch of `total`
// rays over scenes of width `width` (max_stack: largest stack bound of them).
int batch_prepare(rt_ctx ctx, uint32_t width, uint32_t max_stack, uint32_t total, const rt_trace_opts *opts,
                  hipStream_t stream, LaunchShape &shape, BatchLaunch &a);

}  // namespace rtx

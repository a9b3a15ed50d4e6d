// Ray-query serving loop with continuous batching.
//
// The reference renders one frame at a time from one recorded command list
// (RaytracingInterface::render, src/rt/raytracing_interface.cpp:194-255); a
// server instead keeps many scenes resident and answers independent query
// requests as they arrive:
//   * admission: FIFO; a request is admitted when the HBM budget reserved for
//     request buffers (rays + results) has room, and refused at submit if it
//     could never fit;
//   * continuous batching: whenever a stream slot is free, the admitted
//     requests form one tick = one multi-scene launch (k_trace_batch) with
//     the host->device and device->host copies around it on that slot's
//     stream; requests that arrive while ticks run go into the next tick on
//     another slot, so copies and kernels of several ticks overlap;
//   * plain mode: one request per launch and one launch at a time, the
//     per-query baseline on the same GPU;
//   * simulate mode: no device, ticks complete after a modelled time, so the
//     scheduler (admission, deadlines, cancellation, FIFO order) is tested on
//     a host without a GPU; simulate 2 makes every tick complete with a HIP
//     error, for the failure path.
// A ray's result depends only on (scene, ray), so a request's results are
// the same alone and in any mix.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "alloc.hpp"
#include "internal.hpp"

namespace {

using Clock = std::chrono::steady_clock;

double now_ms() {
    return std::chrono::duration<double, std::milli>(Clock::now().time_since_epoch()).count();
}

constexpr size_t kAlign = 256;
inline size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// First-fit allocator over [0, size) (offsets only).
class RangeAlloc {
   public:
    void reset(size_t size) {
        size_ = size;
        free_.clear();
        if (size) free_[0] = size;
        used_ = peak_ = 0;
    }
    bool alloc(size_t bytes, size_t *off) {
        bytes = align_up(std::max<size_t>(bytes, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes;
            const size_t at = it->first + bytes;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = align_up(std::max<size_t>(bytes, 1));
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t size() const { return size_; }
    size_t peak() const { return peak_; }

   private:
    std::map<size_t, size_t> free_;
    size_t size_ = 0, used_ = 0, peak_ = 0;
};

struct Req {
    rt_request r{};
    double t_submit = 0, t_start = 0, t_done = 0;
    int state = RT_TICKET_PENDING;
    int status = RT_OK;
    size_t off = 0, bytes = 0;
    // Device request: recorded on r.stream at submit; the tick waits for it.
    hipEvent_t ready = nullptr;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    uint32_t *err_word = nullptr;  // pinned: the device error word after the tick
    rtx::BatchQuery *host_q = nullptr;  // pinned
    rtx::BatchQuery *dev_q = nullptr;
    size_t cap = 0;
    bool busy = false;
    double sim_done = 0;
    std::vector<uint64_t> tickets;
};

struct SceneEntry {
    rt_scene scene = nullptr;
    bool owned = false;
    bool ready = false;
    uint32_t max_stack = 0;
    // The scene's analytic primitives as registered (its d_ana_table entry).
    rtx::AnaDev ana;
};

size_t request_bytes(const rt_request &r) {
    if (r.memory == RT_REQ_DEVICE) return 0;
    const size_t out = r.kind == RT_QUERY_CLOSEST ? sizeof(rt_hit) : 1;
    return align_up(static_cast<size_t>(r.n) * sizeof(rt_ray)) + align_up(static_cast<size_t>(r.n) * out);
}

}  // namespace

struct rt_server_s {
    rt_ctx ctx = nullptr;
    rt_server_opts o{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> pending;
    std::unordered_map<uint64_t, Req> reqs;
    uint64_t next_ticket = 1;
    RangeAlloc pool;
    uint8_t *d_pool = nullptr;
    std::vector<Slot> slots;
    std::vector<SceneEntry> scenes;
    rtx::BatchScene *d_table = nullptr;
    rtx::BatchAna *d_ana_table = nullptr;  // per scene, beside d_table
    hipStream_t build_stream = nullptr;
    std::mutex build_mu;
    std::thread worker;
    bool stop = false;
    std::string worker_error;
    // Events for device requests' hand-over (recorded at submit, back here
    // when the request ends).
    std::vector<hipEvent_t> ev_pool;
    // stats
    uint64_t submitted = 0, completed = 0, failed = 0, ticks = 0, rays = 0, tick_requests = 0;
    double busy_ms = 0, busy_since = -1;
};

namespace {

int server_fail(rt_server s, int code, const char *msg) {
    rtx::set_error(s ? s->ctx : nullptr, msg);
    return code;
}

void finish_req(rt_server s, uint64_t t, int status, double when) {
    Req &q = s->reqs[t];
    if (q.ready) {
        // Every wait on it was queued on a slot stream that has passed it
        // (finished or synchronized tick) or never queued (cancelled), so it
        // can be recorded again.
        s->ev_pool.push_back(q.ready);
        q.ready = nullptr;
    }
    q.state = status == RT_OK ? RT_TICKET_DONE : RT_TICKET_FAILED;
    q.status = status;
    q.t_done = when;
    if (status == RT_OK)
        ++s->completed;
    else
        ++s->failed;
}

// Records a hand-over event on the submitter's stream (caller holds s->mu).
// The event is made on the server's device whatever the calling thread's
// current device is.
int record_handover(rt_server s, hipStream_t stream, hipEvent_t *out) {
    hipEvent_t ev = nullptr;
    hipError_t e = hipSuccess;
    if (!s->ev_pool.empty()) {
        ev = s->ev_pool.back();
        s->ev_pool.pop_back();
    } else {
        int prev = -1;
        if ((e = hipGetDevice(&prev)) == hipSuccess && prev != s->ctx->device) e = hipSetDevice(s->ctx->device);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (prev >= 0 && prev != s->ctx->device) (void)hipSetDevice(prev);
        if (e != hipSuccess) return rtx::hip_fail(s->ctx, e, "request hand-over event");
    }
    if ((e = hipEventRecord(ev, stream)) != hipSuccess) {
        s->ev_pool.push_back(ev);
        return rtx::hip_fail(s->ctx, e, "request hand-over record (is the stream on the server's device?)");
    }
    *out = ev;
    return RT_OK;
}

// Enqueues one tick on a slot (caller holds s->mu).  Returns an RT_* code;
// on failure the tick's requests are failed by the caller, and nothing of
// the tick is left queued on the slot's stream (a failure after the first
// copy was queued waits for the stream), so the caller may release the
// requests' pool ranges at once.
int launch_tick(rt_server s, Slot &slot, const std::vector<uint64_t> &tick) {
    uint64_t total = 0;
    uint32_t max_stack = 0;
    for (uint64_t t : tick) {
        total += s->reqs[t].r.n;
        max_stack = std::max(max_stack, s->scenes[s->reqs[t].r.scene].max_stack);
    }
    if (s->o.simulate) {
        // Modelled service time: 20 us per tick plus 1 ns per ray.
        slot.sim_done = now_ms() + 0.02 + 1e-6 * static_cast<double>(total);
        return RT_OK;
    }
    hipError_t e;
    if (slot.cap < tick.size()) {
        if (slot.host_q) (void)hipHostFree(slot.host_q);
        if (slot.dev_q) (void)hipFree(slot.dev_q);
        slot.host_q = nullptr;
        slot.dev_q = nullptr;
        slot.cap = 0;
        const size_t cap = std::max<size_t>(tick.size(), 256);
        if ((e = rtx::host_malloc_pinned(&slot.host_q, cap * sizeof(rtx::BatchQuery))) != hipSuccess ||
            (e = rtx::dev_malloc(&slot.dev_q, cap * sizeof(rtx::BatchQuery))) != hipSuccess)
            return rtx::hip_fail(s->ctx, e, "tick query table");
        slot.cap = cap;
    }
    // The launch lock covers the analytic check, the launch resources and
    // the enqueue (its order after s->mu is the context's lock order).
    std::lock_guard<std::mutex> launch(s->ctx->launch_mu);
    bool ana = false;
    for (uint64_t t : tick) {
        const SceneEntry &en = s->scenes[s->reqs[t].r.scene];
        if (en.scene) {
            // The table holds the analytic set the scene had when it was
            // registered; a scene that replaced it since cannot be served.
            if (!rtx::same_analytic(en.scene->dev.ana, en.ana))
                return server_fail(s, RT_ERR_INVALID,
                                   "a scene's analytic primitives changed after it was added to the server");
            ana = ana || en.ana.count != 0;
        }
    }
    // Everything that can fail without queuing anything comes first.
    rtx::BatchLaunch a;
    rtx::LaunchShape shape;
    int rc = rtx::batch_prepare(s->ctx, s->o.node_width, max_stack, static_cast<uint32_t>(total), &s->o.launch,
                                slot.stream, shape, a);
    if (rc != RT_OK) return rc;
    // Expensive requests first, as rt_batch_trace orders its queries: the
    // launch drains on the rays placed last.
    std::vector<uint64_t> order(tick);
    std::stable_sort(order.begin(), order.end(), [&](uint64_t x, uint64_t y) {
        const rt_request &p = s->reqs[x].r, &q = s->reqs[y].r;
        if (p.kind != q.kind) return p.kind < q.kind;
        const rt_scene sa = s->scenes[p.scene].scene, sb = s->scenes[q.scene].scene;
        return (sa ? sa->dev.num_tris : 0u) > (sb ? sb->dev.num_tris : 0u);
    });
    uint32_t start = 0;
    for (size_t k = 0; k < order.size(); ++k) {
        Req &q = s->reqs[order[k]];
        rtx::BatchQuery &b = slot.host_q[k];
        b.scene = q.r.scene;
        b.kind = q.r.kind;
        b.n = q.r.n;
        b.start = start;
        start += q.r.n;
        if (q.r.memory == RT_REQ_DEVICE) {
            b.rays = q.r.rays;
            b.out = q.r.out;
        } else {
            uint8_t *base = s->d_pool + q.off;
            b.rays = reinterpret_cast<const rt_ray *>(base);
            b.out = base + align_up(static_cast<size_t>(q.r.n) * sizeof(rt_ray));
        }
    }
    // From the first queued copy on, a failure waits for the slot's stream.
    auto queued_fail = [&](hipError_t err, const char *what) {
        (void)hipStreamSynchronize(slot.stream);
        return rtx::hip_fail(s->ctx, err, what);
    };
    // Device buffers come from the submitters' streams: the tick waits for
    // what each had queued at submit.  Events on one stream complete in
    // order, so the newest submit per stream covers the others.
    std::vector<std::pair<hipStream_t, hipEvent_t>> handover;
    for (uint64_t t : tick) {  // tick is in submit (ticket) order
        const Req &q = s->reqs[t];
        if (!q.ready) continue;
        const hipStream_t st = static_cast<hipStream_t>(q.r.stream);
        auto it = std::find_if(handover.begin(), handover.end(), [&](const auto &h) { return h.first == st; });
        if (it == handover.end())
            handover.emplace_back(st, q.ready);
        else
            it->second = q.ready;
    }
    for (const auto &h : handover)
        if ((e = hipStreamWaitEvent(slot.stream, h.second, 0)) != hipSuccess)
            return queued_fail(e, "request hand-over");
    for (uint64_t t : tick) {
        Req &q = s->reqs[t];
        if (q.r.memory == RT_REQ_DEVICE) continue;
        if ((e = hipMemcpyAsync(s->d_pool + q.off, q.r.rays, static_cast<size_t>(q.r.n) * sizeof(rt_ray),
                                hipMemcpyHostToDevice, slot.stream)) != hipSuccess)
            return queued_fail(e, "request upload");
    }
    if ((e = hipMemcpyAsync(slot.dev_q, slot.host_q, tick.size() * sizeof(rtx::BatchQuery), hipMemcpyHostToDevice,
                            slot.stream)) != hipSuccess)
        return queued_fail(e, "tick table upload");
    a.scenes = s->d_table;
    a.queries = slot.dev_q;
    a.nq = static_cast<uint32_t>(tick.size());
    if ((e = static_cast<hipError_t>(rtx::launch_batch(s->o.node_width, a, shape, slot.stream))) != hipSuccess)
        return queued_fail(e, "tick launch");
    if (ana) {
        if ((e = static_cast<hipError_t>(rtx::launch_batch_analytic(a, s->d_ana_table, slot.stream))) != hipSuccess)
            return queued_fail(e, "tick analytic launch");
        for (uint64_t t : tick) {
            const SceneEntry &en = s->scenes[s->reqs[t].r.scene];
            if (en.scene) rtx::ana_hold(s->ctx, en.scene, slot.stream);
        }
    }
    for (size_t k = 0; k < tick.size(); ++k) {
        Req &q = s->reqs[tick[k]];
        if (q.r.memory == RT_REQ_DEVICE) continue;
        const size_t out = q.r.kind == RT_QUERY_CLOSEST ? sizeof(rt_hit) : 1;
        const uint8_t *src = s->d_pool + q.off + align_up(static_cast<size_t>(q.r.n) * sizeof(rt_ray));
        if ((e = hipMemcpyAsync(q.r.out, src, static_cast<size_t>(q.r.n) * out, hipMemcpyDeviceToHost, slot.stream)) !=
            hipSuccess)
            return queued_fail(e, "result download");
    }
    // The device error word as the tick ends: a traversal that overflowed
    // its bounded stack fails the tick's requests.
    if ((e = hipMemcpyAsync(slot.err_word, s->ctx->d_err, sizeof(uint32_t), hipMemcpyDeviceToHost, slot.stream)) !=
            hipSuccess ||
        (e = hipEventRecord(slot.done, slot.stream)) != hipSuccess)
        return queued_fail(e, "tick event");
    return RT_OK;
}

// Whether a tick is still running or has ended (its status comes beside).
enum class TickState { Running, Done };

// Whether the slot's tick has ended, and the status of its requests: RT_OK,
// RT_ERR_OVERFLOW (the device error word was set), or RT_ERR_HIP (the tick's
// completion reported a HIP error, e.g. after a device fault; the slot is
// retired with it instead of polling forever).
TickState slot_finished(rt_server s, Slot &slot, int *status) {
    *status = RT_OK;
    if (s->o.simulate) {
        if (now_ms() < slot.sim_done) return TickState::Running;
        if (s->o.simulate == 2) *status = RT_ERR_HIP;  // injected failure (host tests)
        return TickState::Done;
    }
    const hipError_t e = hipEventQuery(slot.done);
    if (e == hipErrorNotReady) return TickState::Running;
    if (e != hipSuccess) {
        rtx::set_error(s->ctx, std::string("tick completion: ") + hipGetErrorString(e));
        *status = RT_ERR_HIP;
    } else if (*slot.err_word != 0u) {
        rtx::set_error(s->ctx, "a tick's traversal overflowed its stack (device error word set)");
        *status = RT_ERR_OVERFLOW;
    }
    return TickState::Done;
}

void worker_main(rt_server s) {
    if (!s->o.simulate) (void)hipSetDevice(s->ctx->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        bool progressed = false;
        const double now = now_ms();
        // 1. completions
        size_t busy = 0;
        for (Slot &slot : s->slots) {
            if (!slot.busy) continue;
            int status = RT_OK;
            if (slot_finished(s, slot, &status) == TickState::Running) {
                ++busy;
                continue;
            }
            const double t = now_ms();
            for (uint64_t tk : slot.tickets) {
                Req &q = s->reqs[tk];
                if (q.bytes) s->pool.release(q.off, q.bytes);
                finish_req(s, tk, status, t);
            }
            slot.tickets.clear();
            slot.busy = false;
            progressed = true;
        }
        if (busy == 0 && s->busy_since >= 0) {
            s->busy_ms += now - s->busy_since;
            s->busy_since = -1;
        }
        if (progressed) s->cv_done.notify_all();
        // 2. deadlines of requests that have not started
        for (auto it = s->pending.begin(); it != s->pending.end();) {
            Req &q = s->reqs[*it];
            if (q.r.deadline_ms > 0 && now - q.t_submit > q.r.deadline_ms) {
                finish_req(s, *it, RT_ERR_CANCELLED, now);
                it = s->pending.erase(it);
                progressed = true;
                s->cv_done.notify_all();
            } else {
                ++it;
            }
        }
        // 3. admission into a free slot
        Slot *free_slot = nullptr;
        for (Slot &slot : s->slots)
            if (!slot.busy) {
                free_slot = &slot;
                break;
            }
        if (s->o.plain && busy) free_slot = nullptr;  // one launch at a time
        if (free_slot && !s->pending.empty()) {
            std::vector<uint64_t> tick;
            uint64_t rays = 0;
            for (auto it = s->pending.begin(); it != s->pending.end();) {
                Req &q = s->reqs[*it];
                if (!s->scenes[q.r.scene].ready) {  // scene still building: later requests may go first
                    ++it;
                    continue;
                }
                if (!tick.empty() && rays + q.r.n > s->o.max_tick_rays) break;
                size_t off = 0;
                const size_t bytes = request_bytes(q.r);
                if (bytes && !s->pool.alloc(bytes, &off)) break;  // FIFO: wait for memory
                q.off = off;
                q.bytes = bytes;
                q.state = RT_TICKET_RUNNING;
                q.t_start = now;
                tick.push_back(*it);
                rays += q.r.n;
                it = s->pending.erase(it);
                if (s->o.plain) break;
            }
            if (!tick.empty()) {
                int rc = launch_tick(s, *free_slot, tick);
                if (rc != RT_OK) {
                    for (uint64_t tk : tick) {
                        Req &q = s->reqs[tk];
                        if (q.bytes) s->pool.release(q.off, q.bytes);
                        finish_req(s, tk, rc, now_ms());
                    }
                    s->cv_done.notify_all();
                } else {
                    free_slot->busy = true;
                    free_slot->tickets = tick;
                    ++s->ticks;
                    s->tick_requests += tick.size();
                    s->rays += rays;
                    if (s->busy_since < 0) s->busy_since = now;
                }
                progressed = true;
            }
        }
        bool any_busy = false;
        for (Slot &slot : s->slots) any_busy |= slot.busy;
        if (s->stop && s->pending.empty() && !any_busy) break;
        if (!progressed) {
            if (any_busy) {
                // Poll in-flight ticks at a fine grain (latency resolution).
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(5));
                lk.lock();
            } else {
                s->cv_work.wait_for(lk, std::chrono::milliseconds(2));
            }
        }
    }
}

}  // namespace

extern "C" {

void rt_server_opts_default(rt_server_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->streams = 4;
    o->max_tick_rays = 8u << 20;
    o->node_width = 4;
    o->max_scenes = 1024;
    rt_trace_opts_default(&o->launch);
    o->launch.refill = 0;  // chosen per mode at creation
    o->launch.chunk = 0;
}

int rt_server_create(rt_ctx ctx, const rt_server_opts *opts, rt_server *out) {
    if (!out) return rtx::fail_global(RT_ERR_INVALID, "null output");
    *out = nullptr;
    rt_server_opts o;
    rt_server_opts_default(&o);
    if (opts) o = *opts;
    if (!ctx && !o.simulate) return rtx::fail_global(RT_ERR_INVALID, "null context");
    if (o.streams == 0) o.streams = 4;
    if (o.streams > 16) return rtx::fail(ctx, RT_ERR_INVALID, "at most 16 streams");
    if (o.max_tick_rays == 0) o.max_tick_rays = 8u << 20;
    if (o.max_tick_rays > 0x7FFFFFFFu) return rtx::fail(ctx, RT_ERR_INVALID, "max_tick_rays too large");
    if (o.node_width == 0) o.node_width = 4;
    if (o.node_width != 2 && o.node_width != 4 && o.node_width != 8)
        return rtx::fail(ctx, RT_ERR_INVALID, "node_width must be 2, 4 or 8");
    if (o.max_scenes == 0) o.max_scenes = 1024;
    if (o.plain) o.streams = 1;
    // Tick shapes measured on MI355X (32 resident scenes, requests of
    // 256-4M rays with their rays in HBM): a multi-request tick is fastest
    // with chunks of 256 and refill 24 (+5% peak over 64/16, as for the
    // 64-scene batch); a one-request launch with 64/16 (chunks of 256 lost
    // 20%: one query's rays spread over fewer waves).
    if (o.launch.chunk == 0) o.launch.chunk = o.plain ? 64u : 256u;
    if (o.launch.refill == 0) o.launch.refill = o.plain ? 16u : 24u;
    rt_server s = new (std::nothrow) rt_server_s;
    if (!s) return rtx::fail(ctx, RT_ERR_OOM, "host allocation failed");
    s->ctx = ctx;
    s->o = o;
    s->slots.resize(o.streams);
    hipError_t e = hipSuccess;
    if (!o.simulate) {
        if ((e = hipSetDevice(ctx->device)) != hipSuccess) goto hip_err;
        if (o.mem_budget == 0) {
            size_t fr = 0, tot = 0;
            if ((e = hipMemGetInfo(&fr, &tot)) != hipSuccess) goto hip_err;
            o.mem_budget = std::min<uint64_t>(fr / 2, 64ull << 30);
            s->o.mem_budget = o.mem_budget;
        }
        if ((e = rtx::dev_malloc(&s->d_pool, o.mem_budget)) != hipSuccess) goto hip_err;
        if ((e = rtx::dev_malloc(&s->d_table, o.max_scenes * sizeof(rtx::BatchScene))) != hipSuccess) goto hip_err;
        if ((e = rtx::dev_malloc(&s->d_ana_table, o.max_scenes * sizeof(rtx::BatchAna))) != hipSuccess) goto hip_err;
        // Non-blocking streams: ticks, builds and the null stream never
        // serialise against each other.
        if ((e = hipStreamCreateWithFlags(&s->build_stream, hipStreamNonBlocking)) != hipSuccess) goto hip_err;
        for (Slot &slot : s->slots) {
            if ((e = hipStreamCreateWithFlags(&slot.stream, hipStreamNonBlocking)) != hipSuccess) goto hip_err;
            if ((e = hipEventCreateWithFlags(&slot.done, hipEventDisableTiming)) != hipSuccess) goto hip_err;
            if ((e = rtx::host_malloc_pinned(reinterpret_cast<void **>(&slot.err_word), sizeof(uint32_t))) != hipSuccess)
                goto hip_err;
            *slot.err_word = 0u;
        }
    } else if (o.mem_budget == 0) {
        s->o.mem_budget = o.mem_budget = 1ull << 30;
    }
    s->pool.reset(o.mem_budget);
    s->worker = std::thread(worker_main, s);
    *out = s;
    return RT_OK;
hip_err:
    for (Slot &slot : s->slots) {
        if (slot.stream) (void)hipStreamDestroy(slot.stream);
        if (slot.done) (void)hipEventDestroy(slot.done);
        if (slot.err_word) (void)hipHostFree(slot.err_word);
    }
    if (s->build_stream) (void)hipStreamDestroy(s->build_stream);
    if (s->d_pool) (void)hipFree(s->d_pool);
    if (s->d_table) (void)hipFree(s->d_table);
    if (s->d_ana_table) (void)hipFree(s->d_ana_table);
    delete s;
    return rtx::hip_fail(ctx, e, "server resources");
}

// Adds a scene table entry (simulate mode accepts a null scene).
static int register_scene(rt_server s, rt_scene scene, bool owned, uint32_t *id) {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->scenes.size() >= s->o.max_scenes) return server_fail(s, RT_ERR_INVALID, "scene table is full");
    SceneEntry en;
    en.scene = scene;
    en.owned = owned;
    en.ready = true;
    if (scene) {
        rt_scene_info info;
        rt_scene_info_get(scene, &info);
        if (info.node_width != s->o.node_width) return server_fail(s, RT_ERR_INVALID, "scene width differs from the server's");
        en.max_stack = info.max_stack;
        std::unique_lock<std::mutex> snap(s->ctx->launch_mu);  // the analytic set as of now
        en.ana = scene->dev.ana;
        snap.unlock();
        rtx::BatchScene b;
        b.nodes = scene->dev.nodes;
        b.tris = scene->dev.tris;
        b.root = scene->dev.root;
        b.pad = 0;
        rtx::BatchAna an;
        an.ana = en.ana;
        an.num_tris = scene->dev.num_tris;
        // A new slot of the tables; in-flight ticks only read older slots.
        hipError_t e = hipMemcpyAsync(s->d_table + s->scenes.size(), &b, sizeof b, hipMemcpyHostToDevice, s->build_stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(s->d_ana_table + s->scenes.size(), &an, sizeof an, hipMemcpyHostToDevice,
                               s->build_stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->build_stream);
        if (e != hipSuccess) return rtx::hip_fail(s->ctx, e, "scene table");
    }
    *id = static_cast<uint32_t>(s->scenes.size());
    s->scenes.push_back(en);
    s->cv_work.notify_all();
    return RT_OK;
}

int rt_server_add_scene(rt_server s, rt_scene scene, uint32_t *id) {
    if (!s || !id) return rtx::fail_global(RT_ERR_INVALID, "null argument");
    if (!scene && !s->o.simulate) return server_fail(s, RT_ERR_INVALID, "null scene");
    if (scene && scene->ctx != s->ctx) return server_fail(s, RT_ERR_INVALID, "scene from another context");
    return register_scene(s, scene, false, id);
}

int rt_server_upload_scene(rt_server s, const float *tri_xyz, uint32_t n, uint32_t *id) {
    if (!s || !id || (n && !tri_xyz)) return rtx::fail_global(RT_ERR_INVALID, "null argument");
    if (s->o.simulate) return register_scene(s, nullptr, false, id);
    std::lock_guard<std::mutex> bg(s->build_mu);
    hipError_t e = hipSetDevice(s->ctx->device);
    if (e != hipSuccess) return rtx::hip_fail(s->ctx, e, "hipSetDevice");
    float *d = nullptr;
    const size_t bytes = static_cast<size_t>(n) * 9 * sizeof(float);
    if (n) {
        if ((e = hipMallocAsync(reinterpret_cast<void **>(&d), bytes, s->build_stream)) != hipSuccess)
            return rtx::hip_fail(s->ctx, e, "staging");
        if ((e = hipMemcpyAsync(d, tri_xyz, bytes, hipMemcpyHostToDevice, s->build_stream)) != hipSuccess) {
            (void)hipFreeAsync(d, s->build_stream);
            return rtx::hip_fail(s->ctx, e, "staging copy");
        }
    }
    rt_build_opts bo;
    rt_build_opts_default(&bo);
    bo.builder = RT_BUILD_LBVH_GPU;
    bo.node_width = s->o.node_width;
    bo.keep_source = 0;
    rt_scene scene = nullptr;
    int rc = rt_scene_create_device(s->ctx, d, n, &bo, s->build_stream, &scene);
    if (d) (void)hipFreeAsync(d, s->build_stream);
    (void)hipStreamSynchronize(s->build_stream);
    if (rc != RT_OK) return rc;
    rc = register_scene(s, scene, true, id);
    if (rc != RT_OK) rt_scene_destroy(scene);
    return rc;
}

int rt_server_submit(rt_server s, const rt_request *req, uint64_t *ticket) {
    if (!s || !req || !ticket) return rtx::fail_global(RT_ERR_INVALID, "null argument");
    if (req->kind > 1) return server_fail(s, RT_ERR_INVALID, "unknown query kind");
    if (req->memory > 1) return server_fail(s, RT_ERR_INVALID, "unknown memory kind");
    if (req->n && (!req->rays || !req->out) && !s->o.simulate) return server_fail(s, RT_ERR_INVALID, "null buffers");
    if (req->n > s->o.max_tick_rays) return server_fail(s, RT_ERR_ADMISSION, "request larger than a tick");
    if (request_bytes(*req) > s->pool.size())
        return server_fail(s, RT_ERR_ADMISSION, "request can never fit the server's memory budget");
    std::lock_guard<std::mutex> g(s->mu);
    if (req->scene >= s->scenes.size()) return server_fail(s, RT_ERR_INVALID, "unknown scene id");
    hipEvent_t ready = nullptr;
    if (req->memory == RT_REQ_DEVICE && req->n && !s->o.simulate) {
        const int rc = record_handover(s, static_cast<hipStream_t>(req->stream), &ready);
        if (rc != RT_OK) return rc;
    }
    const uint64_t t = s->next_ticket++;
    Req &q = s->reqs[t];
    q.r = *req;
    q.ready = ready;
    q.t_submit = now_ms();
    ++s->submitted;
    if (req->n == 0) {
        finish_req(s, t, RT_OK, q.t_submit);
    } else {
        s->pending.push_back(t);
        s->cv_work.notify_all();
    }
    *ticket = t;
    return RT_OK;
}

int rt_server_cancel(rt_server s, uint64_t ticket) {
    if (!s) return rtx::fail_global(RT_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return server_fail(s, RT_ERR_INVALID, "unknown ticket");
    if (it->second.state != RT_TICKET_PENDING) return server_fail(s, RT_ERR_INVALID, "request already started");
    s->pending.erase(std::find(s->pending.begin(), s->pending.end(), ticket));
    finish_req(s, ticket, RT_ERR_CANCELLED, now_ms());
    s->cv_done.notify_all();
    return RT_OK;
}

static void read_ticket(const Req &q, int *state, int *status, double *latency_ms) {
    if (state) *state = q.state;
    if (status) *status = q.status;
    if (latency_ms) *latency_ms = q.state >= RT_TICKET_DONE ? q.t_done - q.t_submit : -1.0;
}

int rt_server_poll(rt_server s, uint64_t ticket, int *state, int *status, double *latency_ms) {
    if (!s) return rtx::fail_global(RT_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return server_fail(s, RT_ERR_INVALID, "unknown ticket");
    read_ticket(it->second, state, status, latency_ms);
    return RT_OK;
}

int rt_server_wait(rt_server s, uint64_t ticket, double timeout_ms, int *state, int *status, double *latency_ms) {
    if (!s) return rtx::fail_global(RT_ERR_INVALID, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return server_fail(s, RT_ERR_INVALID, "unknown ticket");
    const auto until = Clock::now() + std::chrono::duration<double, std::milli>(timeout_ms > 0 ? timeout_ms : 1e12);
    s->cv_done.wait_until(lk, until, [&] { return s->reqs[ticket].state >= RT_TICKET_DONE; });
    read_ticket(s->reqs[ticket], state, status, latency_ms);
    return RT_OK;
}

int rt_server_forget(rt_server s, uint64_t ticket) {
    if (!s) return rtx::fail_global(RT_ERR_INVALID, "null server");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return server_fail(s, RT_ERR_INVALID, "unknown ticket");
    if (it->second.state < RT_TICKET_DONE) return server_fail(s, RT_ERR_INVALID, "request not finished");
    s->reqs.erase(it);
    return RT_OK;
}

int rt_server_stats_get(rt_server s, rt_server_stats *st) {
    if (!s || !st) return rtx::fail_global(RT_ERR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    std::memset(st, 0, sizeof *st);
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->failed = s->failed;
    st->ticks = s->ticks;
    st->rays = s->rays;
    st->mem_budget = s->pool.size();
    st->mem_peak = s->pool.peak();
    st->busy_ms = s->busy_ms + (s->busy_since >= 0 ? now_ms() - s->busy_since : 0.0);
    st->mean_requests_per_tick = s->ticks ? static_cast<double>(s->tick_requests) / s->ticks : 0.0;
    return RT_OK;
}

int rt_server_destroy(rt_server s) {
    if (!s) return RT_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    if (!s->o.simulate) {
        (void)hipSetDevice(s->ctx->device);
        for (Slot &slot : s->slots) {
            if (slot.stream) (void)hipStreamSynchronize(slot.stream);
            if (slot.stream) (void)hipStreamDestroy(slot.stream);
            if (slot.done) (void)hipEventDestroy(slot.done);
            if (slot.host_q) (void)hipHostFree(slot.host_q);
            if (slot.dev_q) (void)hipFree(slot.dev_q);
            if (slot.err_word) (void)hipHostFree(slot.err_word);
        }
        if (s->build_stream) (void)hipStreamDestroy(s->build_stream);
        for (auto &kv : s->reqs)
            if (kv.second.ready) (void)hipEventDestroy(kv.second.ready);
        for (hipEvent_t ev : s->ev_pool) (void)hipEventDestroy(ev);
        for (SceneEntry &en : s->scenes)
            if (en.owned && en.scene) rt_scene_destroy(en.scene);
        if (s->d_pool) (void)hipFree(s->d_pool);
        if (s->d_table) (void)hipFree(s->d_table);
        if (s->d_ana_table) (void)hipFree(s->d_ana_table);
    }
    delete s;
    return RT_OK;
}

}  // extern "C"

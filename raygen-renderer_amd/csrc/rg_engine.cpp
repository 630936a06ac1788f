// rg_engine.cpp — host side of the streaming engine: job admission by HBM
// footprint into one device arena, the issue table, hipGraph replay of the
// captured iterations, completion tracking from a device status snapshot,
// batch rendering (rg_render_batch), the request server (rg_server_*) and
// the autotuner (rg_autotune).
//
// The reference renders one job per RayRenderer::render call
// (rayrenderer.cpp:478-596); these entry points run many such jobs at once
// on one GPU, each producing what that call would for its scene and
// settings.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "rg_internal.h"
#include "rg_pt.h"
#include "rg_stream.h"

namespace rg {
namespace {

double now_s() {
    static const auto t0 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int hip_ok(hipError_t e, const char *what) {
    if (e == hipSuccess) return RG_OK;
    return set_error(RG_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

constexpr uint64_t kAlign = 256;
uint64_t align_up(uint64_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over one device allocation (offsets only).
class Arena {
  public:
    void reset(uint64_t size) {
        size_ = size;
        used_ = 0;
        free_.clear();
        if (size) free_[0] = size;
    }
    bool alloc(uint64_t bytes, uint64_t &off) {
        bytes = align_up(bytes);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            off = it->first;
            const uint64_t rem = it->second - bytes;
            free_.erase(it);
            if (rem) free_[off + bytes] = rem;
            used_ += bytes;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes) {
        bytes = align_up(bytes);
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
    uint64_t size() const { return size_; }
    uint64_t used() const { return used_; }

  private:
    uint64_t size_ = 0, used_ = 0;
    std::map<uint64_t, uint64_t> free_;
};

struct Request {
    uint64_t ticket = 0;
    uint32_t scene = 0;
    rg_settings st{};
    float *out_hdr = nullptr, *out_sum = nullptr, *out_sumsq = nullptr;
    uint64_t samples = 0;  // npix * spp
    uint32_t npix = 0;
    uint64_t bytes = 0;    // arena footprint
    uint64_t off = 0;
    uint64_t begin = 0;
    uint32_t slot = 0;
    std::atomic<int> state{0};  // 0 pending, 1 issued, 2 finishing, 3 done
    int status = RG_OK;
    double t_submit = 0.0, t_admit = 0.0, t_done = 0.0;
    hipEvent_t ev = nullptr;
    int pin = -1;
    uint32_t bounces = 0;     // max_depth + 2: the longest a path of the job lives, in iterations
    uint64_t issued_iter = ~0ull;  // iteration count when the cursor was first seen past the job
    float *d_hdr = nullptr, *d_sum = nullptr, *d_sumsq = nullptr;
};

struct PinnedBuf {
    void *p = nullptr;
    size_t cap = 0;
    bool busy = false;
};

// Footprint of one job in the arena: per-sample buffer, sums, HDR image.
uint64_t job_bytes(uint64_t npix, uint64_t spp) {
    return align_up(npix * spp * 12) + 2 * align_up(npix * 12) + align_up(npix * 16);
}

// The single-job path's samples in flight, which fixes its accumulation
// grouping (rg_render).
uint32_t plain_group(const rg_settings &st) {
    const uint32_t npix = st.width * st.height;
    const uint32_t spp = std::max<uint32_t>(1, st.spp);
    uint32_t S = st.samples_in_flight;
    if (S == 0) S = std::max<uint32_t>(1, (1u << 21) / npix);
    return std::min(S, spp);
}

}  // namespace

class Engine {
  public:
    ~Engine() { shutdown(); }

    int init(Ctx *c, rg_scene *const *sc, uint32_t n, uint64_t budget, const rg_tune *t) {
        ctx = c;
        if (!c || (n && !sc)) return set_error(RG_E_INVALID, "null argument");
        if (c->device < 0) return set_error(RG_E_NODEVICE, "host-only context has no device");
        if (n == 0) return set_error(RG_E_INVALID, "no scenes");
        if (int rc = dev_set_device(c->device)) return rc;
        if (int rc = hip_ok(hipStreamCreate(&stream), "stream create")) return rc;
        uint32_t need = 0;
        int width = 0;
        for (uint32_t i = 0; i < n; i++) {
            Scene *s = reinterpret_cast<Scene *>(sc[i]);
            if (!s) return set_error(RG_E_INVALID, "null scene");
            if (int rc = prepare_scene(s, stream)) return rc;
            need = std::max(need, s->stack_need);
            if (width && s->width != width)
                return set_error(RG_E_UNSUPPORTED, "all scenes of one engine need the same BVH width");
            width = s->width;
            scenes.push_back(s);
        }
        tune = {1u << 21, 0, 8, 8, 4, (uint32_t)width};
        if (t) {
            if (t->pool_paths) tune.pool = t->pool_paths;
            if (t->stack) tune.stack = t->stack;
            if (t->tile) tune.tile = t->tile;
            if (t->blocks_per_cu) tune.blocks_per_cu = t->blocks_per_cu;
            if (t->iters_per_graph) tune.iters = t->iters_per_graph;
        }
        if (tune.stack == 0)
            for (uint32_t s : {32u, 64u, 128u, 256u})
                if (stream_stack_ok(s, need)) {
                    tune.stack = s;
                    break;
                }
        if (!stream_stack_ok(tune.stack, need))
            return set_error(RG_E_UNSUPPORTED, "traversal stack too small for the deepest scene BVH");
        if (tune.pool < 1024 || tune.pool > (1u << 26)) return set_error(RG_E_INVALID, "pool_paths out of range");
        if (tune.tile > 64 || tune.blocks_per_cu == 0 || tune.blocks_per_cu > 32)
            return set_error(RG_E_INVALID, "tile or blocks_per_cu out of range");
        if (tune.iters == 0 || (tune.iters & 1) || tune.iters > 64)
            return set_error(RG_E_INVALID, "iters_per_graph must be even, 2..64");
        n_cu = stream_cu_count(c->device);
        if (n_cu <= 0) return set_error(RG_E_HIP, "cannot query compute units");

        const uint64_t P = tune.pool;
        struct A {
            void **p;
            uint64_t bytes;
        } allocs[] = {
            {(void **)&sp.st.ray, P * 32},    {(void **)&sp.st.beta, P * 16},  {(void **)&sp.st.tp, P * 16},
            {(void **)&sp.st.L, P * 16},      {(void **)&sp.st.meta, P * 16},  {(void **)&sp.st.sid, P * 16},
            {(void **)&sp.hit, P * 16},       {(void **)&sp.sray, P * 64},     {(void **)&sp.scontrib, P * 32},
            {(void **)&sp.socc, P * 2},       {(void **)&sp.ssrc, P * 8},      {(void **)&sp.nshadow, P * 4},
            {(void **)&sp.holes, P * 4},      {&d_jobs, (uint64_t)kMaxJobs * kDevJobBytes},
            {&d_status, status_bytes()},      {(void **)&d_issue, (uint64_t)kMaxJobs * sizeof(IssueEntry)},
        };
        for (auto &a : allocs) {
            if (int rc = dev_malloc(a.p, a.bytes)) return rc;
            dev_allocs.push_back(*a.p);
        }
        sp.cap = tune.pool;
        sp.ctr = (StreamCounters *)d_status;
        sp.issue = d_issue;
        if (budget == 0) budget = 16ull << 30;
        if (int rc = dev_malloc((void **)&d_arena, budget)) return rc;
        dev_allocs.push_back(d_arena);
        arena.reset(budget);

        if (int rc = hip_ok(hipHostMalloc((void **)&h_jobs, (size_t)kMaxJobs * kDevJobBytes), "pinned alloc")) return rc;
        for (int k = 0; k < 2; k++) {
            if (int rc = hip_ok(hipHostMalloc((void **)&h_status[k], status_bytes()), "pinned alloc")) return rc;
            if (int rc = hip_ok(hipHostMalloc((void **)&h_issue[k], (size_t)kMaxJobs * sizeof(IssueEntry) + 64),
                                "pinned alloc"))
                return rc;
            if (int rc = hip_ok(hipEventCreateWithFlags(&status_ev[k], hipEventDisableTiming), "event")) return rc;
            if (int rc = hip_ok(hipEventCreateWithFlags(&issue_ev[k], hipEventDisableTiming), "event")) return rc;
        }
        if (int rc = hip_ok(hipHostMalloc((void **)&h_zero, 64), "pinned alloc")) return rc;
        std::memset(h_zero, 0, 64);
        for (uint32_t s = kMaxJobs; s-- > 0;) free_slots.push_back(s);

        if (int rc = stream_launch_reset(sp, stream)) return rc;
        // capture the iterations once; every replay reuses the same arguments
        hipGraph_t g = nullptr;
        if (int rc = hip_ok(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed), "begin capture")) return rc;
        int lrc = RG_OK;
        for (uint32_t it = 0; it < tune.iters && lrc == RG_OK; it++)
            lrc = stream_launch_iteration(sp, d_jobs, (int)(it & 1), tune, n_cu, stream);
        const hipError_t ce = hipStreamEndCapture(stream, &g);
        if (lrc) return lrc;
        if (int rc = hip_ok(ce, "end capture")) return rc;
        const hipError_t ie = hipGraphInstantiate(&graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (int rc = hip_ok(ie, "graph instantiate")) return rc;
        return hip_ok(hipStreamSynchronize(stream), "engine init");
    }

    void shutdown() {
        if (worker.joinable()) {
            {
                std::lock_guard<std::mutex> lk(mu);
                stop = true;
            }
            cv_work.notify_all();
            worker.join();
        }
        if (stream) (void)hipStreamSynchronize(stream);
        {
            std::lock_guard<std::mutex> lk(mu);
            for (auto &kv : requests) {
                Request *r = kv.second.get();
                if (r->state != 3) {
                    r->state = 3;
                    r->status = RG_E_CANCELLED;
                }
                if (r->ev) (void)hipEventDestroy(r->ev);
                r->ev = nullptr;
            }
        }
        cv_done.notify_all();
        if (graph) (void)hipGraphExecDestroy(graph);
        graph = nullptr;
        for (void *p : dev_allocs) dev_free(p);
        dev_allocs.clear();
        for (char *p : {h_jobs, h_status[0], h_status[1], (char *)h_issue[0], (char *)h_issue[1], (char *)h_zero})
            if (p) (void)hipHostFree(p);
        h_jobs = h_status[0] = h_status[1] = nullptr;
        h_issue[0] = h_issue[1] = nullptr;
        h_zero = nullptr;
        for (PinnedBuf &b : pins)
            if (b.p) (void)hipHostFree(b.p);
        pins.clear();
        for (int k = 0; k < 2; k++) {
            if (status_ev[k]) (void)hipEventDestroy(status_ev[k]);
            if (issue_ev[k]) (void)hipEventDestroy(issue_ev[k]);
            status_ev[k] = issue_ev[k] = nullptr;
        }
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    // Validates and queues one job; thread-safe.
    int submit(const rg_job &job, float *hdr, float *sum, float *sumsq, uint64_t *ticket) {
        const rg_settings &st = job.settings;
        if (job.scene >= scenes.size()) return set_error(RG_E_INVALID, "scene index out of range");
        if (st.width == 0 || st.height == 0 || st.width > 16384 || st.height > 16384)
            return set_error(RG_E_INVALID, "resolution must be 1..16384 per axis");
        if (st.max_depth < 1 || st.max_depth > 64) return set_error(RG_E_INVALID, "max_depth must be 1..64");
        if (st.spp == 0) return set_error(RG_E_INVALID, "spp must be >= 1");
        if (st.adaptive || st.denoise)
            return set_error(RG_E_UNSUPPORTED, "batched jobs render a fixed sample count without denoise");
        const uint64_t npix = (uint64_t)st.width * st.height;
        if (npix * st.spp >= (1ull << 32)) return set_error(RG_E_BUDGET, "job has 2^32 or more samples");
        const uint64_t bytes = job_bytes(npix, st.spp);
        if (bytes > arena.size()) return set_error(RG_E_BUDGET, "job footprint exceeds the engine's HBM budget");
        auto r = std::make_unique<Request>();
        r->scene = job.scene;
        r->st = st;
        r->out_hdr = hdr;
        r->out_sum = sum;
        r->out_sumsq = sumsq;
        r->npix = (uint32_t)npix;
        r->samples = npix * st.spp;
        r->bytes = bytes;
        r->t_submit = now_s();
        {
            std::lock_guard<std::mutex> lk(mu);
            r->ticket = next_ticket++;
            if (ticket) *ticket = r->ticket;
            inbox.push_back(r.get());
            requests[r->ticket] = std::move(r);
        }
        cv_work.notify_all();
        return RG_OK;
    }

    // One round of the loop: admit, replay, read status, finish, deliver.
    int step() {
        if (ctx->cancel.load()) return fail_all(RG_E_CANCELLED, "cancelled");
        {
            std::lock_guard<std::mutex> lk(mu);
            while (!inbox.empty()) {
                pending.push_back(inbox.front());
                inbox.pop_front();
            }
        }
        if (int rc = admit()) return fail_all(rc, nullptr);
        if (int rc = upload_issue()) return fail_all(rc, nullptr);
        bool running = false;
        for (Request *r : active) running |= r->state == 1;
        if (running) {
            if (int rc = hip_ok(hipGraphLaunch(graph, stream), "graph launch")) return fail_all(rc, nullptr);
            graphs++;
            const int s = status_cur;
            status_cur ^= 1;
            if (int rc = hip_ok(hipMemcpyAsync(h_status[s], d_status, status_bytes(), hipMemcpyDeviceToHost, stream),
                                "status copy"))
                return fail_all(rc, nullptr);
            if (int rc = hip_ok(hipEventRecord(status_ev[s], stream), "event record")) return fail_all(rc, nullptr);
            status_pending[s] = true;
            if (status_pending[s ^ 1])
                if (int rc = process_status(s ^ 1)) return fail_all(rc, nullptr);
        } else {
            for (int k : {status_cur, status_cur ^ 1})
                if (status_pending[k])
                    if (int rc = process_status(k)) return fail_all(rc, nullptr);
        }
        return deliver(!running);
    }

    // True while anything is queued, admitted or finishing.
    bool busy() {
        std::lock_guard<std::mutex> lk(mu);
        return !inbox.empty() || !pending.empty() || !active.empty();
    }

    int run_until_idle() {
        while (busy())
            if (int rc = step()) return rc;
        return RG_OK;
    }

    void start_thread() {
        const int dev = ctx->device;
        worker = std::thread([this, dev]() {
            dev_set_device(dev);
            for (;;) {
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv_work.wait(lk, [&] { return stop || !inbox.empty() || !pending.empty() || !active.empty(); });
                    if (stop) return;
                }
                if (step()) {
                    // failures are recorded on the requests; keep serving
                }
            }
        });
    }

    int wait(uint64_t ticket, double timeout_s, rg_request_info *info) {
        std::unique_lock<std::mutex> lk(mu);
        auto it = requests.find(ticket);
        if (it == requests.end()) return set_error(RG_E_INVALID, "unknown ticket");
        Request *r = it->second.get();
        auto ready = [&] { return r->state == 3; };
        if (timeout_s < 0) cv_done.wait(lk, ready);
        else if (!cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) return RG_TIMEOUT;
        if (info) {
            info->status = r->status;
            info->pad = 0;
            info->samples = r->samples;
            info->submit_s = r->t_submit;
            info->admit_s = r->t_admit;
            info->done_s = r->t_done;
        }
        const int st = r->status;
        requests.erase(it);
        if (st != RG_OK) return set_error(st, "request failed: " + failure);
        return RG_OK;
    }

    void get_stats(uint64_t out[8]) {
        std::lock_guard<std::mutex> lk(mu);
        out[0] = n_done;
        out[1] = samples_done;
        out[2] = dev_stats[0];
        out[3] = dev_stats[1];
        out[4] = dev_stats[3];
        out[5] = arena.used();
        out[6] = arena.size();
        out[7] = graphs;
    }

    StreamTune tune{};
    uint64_t dev_stats[4] = {0, 0, 0, 0};

  private:
    static size_t status_bytes() { return sizeof(StreamCounters); }

    int admit() {
        while (!pending.empty() && !free_slots.empty()) {
            Request *r = pending.front();
            uint64_t off = 0;
            if (!arena.alloc(r->bytes, off)) break;  // first come, first served
            pending.pop_front();
            r->off = off;
            r->slot = free_slots.back();
            free_slots.pop_back();
            r->begin = seq_end;
            seq_end += r->samples;
            char *base = d_arena + off;
            const uint64_t np = r->npix;
            JobDesc d;
            d.scene = scenes[r->scene];
            d.st = r->st;
            d.d_samples = (float *)base;
            d.d_sum = (float *)(base + align_up(np * r->st.spp * 12));
            d.d_sumsq = (float *)((char *)d.d_sum + align_up(np * 12));
            d.d_hdr = (float *)((char *)d.d_sumsq + align_up(np * 12));
            d.group = plain_group(r->st);
            d.tile = tune.tile;
            char *rec = h_jobs + (size_t)r->slot * kDevJobBytes;
            stream_job_record(d, rec);
            if (int rc = hip_ok(hipMemcpyAsync((char *)d_jobs + (size_t)r->slot * kDevJobBytes, rec, kDevJobBytes,
                                               hipMemcpyHostToDevice, stream),
                                "job upload"))
                return rc;
            r->d_hdr = d.d_hdr;
            r->d_sum = d.d_sum;
            r->d_sumsq = d.d_sumsq;
            r->bounces = r->st.max_depth + 2;
            r->state = 1;
            r->t_admit = now_s();
            active.push_back(r);
            issue_dirty = true;
        }
        return RG_OK;
    }

    int upload_issue() {
        if (!issue_dirty) return RG_OK;
        const int b = issue_cur;
        issue_cur ^= 1;
        if (int rc = hip_ok(hipEventSynchronize(issue_ev[b]), "issue staging")) return rc;
        IssueEntry *e = h_issue[b];
        uint32_t n = 0;
        for (Request *r : active)
            if (r->begin + r->samples > cursor_seen) e[n++] = {r->begin, r->slot, 0};
        uint64_t *hdr = (uint64_t *)(e + kMaxJobs);
        hdr[0] = seq_end;
        ((uint32_t *)(hdr + 1))[0] = n;
        if (n)
            if (int rc = hip_ok(hipMemcpyAsync(d_issue, e, n * sizeof(IssueEntry), hipMemcpyHostToDevice, stream),
                                "issue upload"))
                return rc;
        if (int rc = hip_ok(hipMemcpyAsync(&sp.ctr->seq_end, hdr, 8, hipMemcpyHostToDevice, stream), "issue upload"))
            return rc;
        if (int rc = hip_ok(hipMemcpyAsync(&sp.ctr->n_issue, hdr + 1, 4, hipMemcpyHostToDevice, stream),
                            "issue upload"))
            return rc;
        if (int rc = hip_ok(hipEventRecord(issue_ev[b], stream), "event record")) return rc;
        issue_dirty = false;
        return RG_OK;
    }

    int process_status(int k) {
        if (int rc = hip_ok(hipEventSynchronize(status_ev[k]), "status wait")) return rc;
        status_pending[k] = false;
        const StreamCounters *C = (const StreamCounters *)h_status[k];
        // after a replay (an even number of iterations) the live cursor is slot 0
        cursor_seen = C->cursor[0];
        const uint64_t iters = C->stats[3];
        for (int q = 0; q < 4; q++) dev_stats[q] = C->stats[q];
        for (Request *r : active) {
            if (r->state != 1) continue;
            // all samples issued by iteration issued_iter; none lives longer
            // than `bounces` iterations
            if (r->issued_iter == ~0ull && cursor_seen >= r->begin + r->samples) r->issued_iter = iters;
            if (r->issued_iter == ~0ull || iters + 1 < r->issued_iter + r->bounces) continue;
            if (int rc = stream_launch_finish(d_jobs, r->slot, r->npix, stream)) return rc;
            const size_t hb = (size_t)r->npix * 16, sb = (size_t)r->npix * 12;
            const size_t need = (r->out_hdr ? hb : 0) + (r->out_sum ? sb : 0) + (r->out_sumsq ? sb : 0);
            if (need) {
                r->pin = get_pinned(need);
                if (r->pin < 0) return set_error(RG_E_NOMEM, "pinned staging allocation failed");
                char *p = (char *)pins[r->pin].p;
                if (r->out_hdr) {
                    if (int rc = hip_ok(hipMemcpyAsync(p, r->d_hdr, hb, hipMemcpyDeviceToHost, stream), "result copy"))
                        return rc;
                    p += hb;
                }
                if (r->out_sum) {
                    if (int rc = hip_ok(hipMemcpyAsync(p, r->d_sum, sb, hipMemcpyDeviceToHost, stream), "result copy"))
                        return rc;
                    p += sb;
                }
                if (r->out_sumsq)
                    if (int rc = hip_ok(hipMemcpyAsync(p, r->d_sumsq, sb, hipMemcpyDeviceToHost, stream),
                                        "result copy"))
                        return rc;
            }
            if (!r->ev)
                if (int rc = hip_ok(hipEventCreateWithFlags(&r->ev, hipEventDisableTiming), "event")) return rc;
            if (int rc = hip_ok(hipEventRecord(r->ev, stream), "event record")) return rc;
            r->state = 2;
        }
        return RG_OK;
    }

    int get_pinned(size_t need) {
        int best = -1;
        for (size_t i = 0; i < pins.size(); i++)
            if (!pins[i].busy && pins[i].cap >= need && (best < 0 || pins[i].cap < pins[best].cap)) best = (int)i;
        if (best < 0) {
            PinnedBuf b;
            b.cap = (need + (1u << 20) - 1) & ~(size_t)((1u << 20) - 1);
            if (hipHostMalloc(&b.p, b.cap) != hipSuccess) return -1;
            pins.push_back(b);
            best = (int)pins.size() - 1;
        }
        pins[best].busy = true;
        return best;
    }

    // Hands finished requests to their waiters; blocks on the oldest
    // finishing request when nothing else is running.
    int deliver(bool may_block) {
        bool any = false;
        for (size_t i = 0; i < active.size();) {
            Request *r = active[i];
            if (r->state != 2) {
                i++;
                continue;
            }
            hipError_t q = hipEventQuery(r->ev);
            if (q == hipErrorNotReady && may_block && !any) q = hipEventSynchronize(r->ev);
            if (q == hipErrorNotReady) {
                i++;
                continue;
            }
            if (int rc = hip_ok(q, "result wait")) return fail_all(rc, nullptr);
            if (r->pin >= 0) {
                const char *p = (const char *)pins[r->pin].p;
                const size_t hb = (size_t)r->npix * 16, sb = (size_t)r->npix * 12;
                if (r->out_hdr) {
                    std::memcpy(r->out_hdr, p, hb);
                    p += hb;
                }
                if (r->out_sum) {
                    std::memcpy(r->out_sum, p, sb);
                    p += sb;
                }
                if (r->out_sumsq) std::memcpy(r->out_sumsq, p, sb);
                pins[r->pin].busy = false;
                r->pin = -1;
            }
            arena.release(r->off, r->bytes);
            free_slots.push_back(r->slot);
            active.erase(active.begin() + (long)i);
            any = true;
            {
                std::lock_guard<std::mutex> lk(mu);
                r->t_done = now_s();
                r->state = 3;
                n_done++;
                samples_done += r->samples;
            }
            cv_done.notify_all();
        }
        return RG_OK;
    }

    int fail_all(int rc, const char *msg) {
        std::string text = msg ? std::string(msg) : std::string(rg_last_error());
        if (stream) (void)hipStreamSynchronize(stream);
        {
            std::lock_guard<std::mutex> lk(mu);
            failure = text;
            for (auto &kv : requests) {
                Request *r = kv.second.get();
                if (r->state == 3) continue;
                r->state = 3;
                r->status = rc;
                r->t_done = now_s();
                if (r->pin >= 0) pins[r->pin].busy = false;
                r->pin = -1;
            }
            inbox.clear();
        }
        for (Request *r : active) {
            arena.release(r->off, r->bytes);
            free_slots.push_back(r->slot);
        }
        active.clear();
        pending.clear();
        cv_done.notify_all();
        return set_error(rc, text);
    }

    Ctx *ctx = nullptr;
    std::vector<Scene *> scenes;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
    StreamPool sp{};
    void *d_jobs = nullptr;
    void *d_status = nullptr;
    IssueEntry *d_issue = nullptr;
    char *d_arena = nullptr;
    std::vector<void *> dev_allocs;
    Arena arena;
    char *h_jobs = nullptr;
    char *h_status[2] = {nullptr, nullptr};
    hipEvent_t status_ev[2] = {nullptr, nullptr};
    bool status_pending[2] = {false, false};
    int status_cur = 0;
    IssueEntry *h_issue[2] = {nullptr, nullptr};
    hipEvent_t issue_ev[2] = {nullptr, nullptr};
    int issue_cur = 0;
    uint32_t *h_zero = nullptr;
    std::vector<PinnedBuf> pins;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> inbox;    // submitted, not yet seen by the loop
    std::deque<Request *> pending;  // waiting for a slot and arena space
    std::vector<Request *> active;  // admitted, in issue order
    std::map<uint64_t, std::unique_ptr<Request>> requests;
    std::vector<uint32_t> free_slots;
    uint64_t seq_end = 0, cursor_seen = 0, next_ticket = 1;
    bool issue_dirty = false;
    bool stop = false;
    std::thread worker;
    std::string failure;
    uint64_t n_done = 0, samples_done = 0, graphs = 0;

};

}  // namespace rg

using namespace rg;

struct rg_server {
    std::unique_ptr<Engine> eng;
};

namespace {

// Runs n jobs to completion on an initialised engine.
int batch_on(Engine &e, const rg_job *jobs, uint32_t n, float *const *hdr_out, float *const *sum_out) {
    uint64_t first = 0;
    for (uint32_t i = 0; i < n; i++) {
        uint64_t t = 0;
        if (int rc = e.submit(jobs[i], hdr_out ? hdr_out[i] : nullptr, sum_out ? sum_out[i] : nullptr, nullptr, &t))
            return rc;
        if (i == 0) first = t;
    }
    if (int rc = e.run_until_idle()) return rc;
    for (uint32_t i = 0; i < n; i++)
        if (int rc = e.wait(first + i, 0.0, nullptr)) return rc;
    return RG_OK;
}

uint64_t batch_budget(const rg_job *jobs, uint32_t n) {
    uint64_t need = 0, big = 0;
    for (uint32_t i = 0; i < n; i++) {
        const rg_settings &st = jobs[i].settings;
        const uint64_t b = align_up(job_bytes((uint64_t)st.width * st.height, st.spp));
        need += b;
        big = std::max(big, b);
    }
    // everything at once up to 64 GiB; beyond that jobs wait for space
    return std::max<uint64_t>(std::max(std::min<uint64_t>(need, 64ull << 30), big), 1ull << 20);
}

}  // namespace

extern "C" {

int rg_render_batch(rg_ctx *ctx, rg_scene *const *scenes, uint32_t n_scenes, const rg_job *jobs, uint32_t n_jobs,
                    const rg_tune *tune, float *const *hdr_out, float *const *sum_out, uint64_t stats[4]) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || (n_jobs && !jobs)) return set_error(RG_E_INVALID, "null argument");
    Engine e;
    if (int rc = e.init(c, scenes, n_scenes, batch_budget(jobs, n_jobs), tune)) return rc;
    if (int rc = batch_on(e, jobs, n_jobs, hdr_out, sum_out)) return rc;
    if (stats)
        for (int k = 0; k < 4; k++) stats[k] = e.dev_stats[k];
    return RG_OK;
}

int rg_server_create(rg_ctx *ctx, rg_scene *const *scenes, uint32_t n_scenes, uint64_t hbm_budget,
                     const rg_tune *tune, rg_server **out) {
    if (!out) return set_error(RG_E_INVALID, "null argument");
    *out = nullptr;
    auto s = std::make_unique<rg_server>();
    s->eng = std::make_unique<Engine>();
    if (int rc = s->eng->init(reinterpret_cast<Ctx *>(ctx), scenes, n_scenes, hbm_budget, tune)) return rc;
    s->eng->start_thread();
    *out = s.release();
    return RG_OK;
}

int rg_server_submit(rg_server *srv, const rg_job *job, float *hdr_out, float *sum_out, uint64_t *ticket) {
    if (!srv || !job) return set_error(RG_E_INVALID, "null argument");
    return srv->eng->submit(*job, hdr_out, sum_out, nullptr, ticket);
}

int rg_server_wait(rg_server *srv, uint64_t ticket, double timeout_s, rg_request_info *info) {
    if (!srv) return set_error(RG_E_INVALID, "null server");
    return srv->eng->wait(ticket, timeout_s, info);
}

int rg_server_stats(rg_server *srv, uint64_t out[8]) {
    if (!srv || !out) return set_error(RG_E_INVALID, "null argument");
    srv->eng->get_stats(out);
    return RG_OK;
}

int rg_server_get_tune(rg_server *srv, rg_tune *out) {
    if (!srv || !out) return set_error(RG_E_INVALID, "null argument");
    const StreamTune &t = srv->eng->tune;
    *out = {t.pool, t.stack, t.tile, t.blocks_per_cu, t.iters, 0};
    return RG_OK;
}

int rg_server_destroy(rg_server *srv) {
    if (!srv) return RG_OK;
    srv->eng->shutdown();
    delete srv;
    return RG_OK;
}

int rg_autotune(rg_ctx *ctx, rg_scene *const *scenes, uint32_t n_scenes, const rg_job *probe, uint32_t n_probe,
                rg_tune *out) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !out || (n_scenes && !scenes)) return set_error(RG_E_INVALID, "null argument");
    std::vector<rg_job> jobs;
    if (probe && n_probe) {
        jobs.assign(probe, probe + n_probe);
    } else {
        for (uint32_t i = 0; i < n_scenes; i++) {
            rg_job j;
            std::memset(&j, 0, sizeof j);
            j.scene = i;
            rg_settings_default(&j.settings);
            j.settings.width = 512;
            j.settings.height = 384;
            j.settings.spp = 8;
            j.settings.max_depth = 6;
            jobs.push_back(j);
        }
    }
    uint32_t need = 0;
    for (uint32_t i = 0; i < n_scenes; i++) {
        const Scene *s = reinterpret_cast<const Scene *>(scenes[i]);
        if (!s) return set_error(RG_E_INVALID, "null scene");
        need = std::max(need, s->stack_need);
    }
    std::vector<uint32_t> stacks;
    for (uint32_t s : {32u, 64u, 128u, 256u})
        if (stream_stack_ok(s, need) && stacks.size() < 2) stacks.push_back(s);
    if (stacks.empty()) return set_error(RG_E_UNSUPPORTED, "traversal stack too small for the deepest scene BVH");
    const uint64_t budget = batch_budget(jobs.data(), (uint32_t)jobs.size());
    // seconds for the probe jobs under one setting: two runs on one engine,
    // the second timed (graph, pinned buffers and caches warm)
    auto measure = [&](const rg_tune &t, double &dt) -> int {
        Engine e;
        if (int rc = e.init(c, scenes, n_scenes, budget, &t)) return rc;
        for (int rep = 0; rep < 2; rep++) {
            const double t0 = now_s();
            if (int rc = batch_on(e, jobs.data(), (uint32_t)jobs.size(), nullptr, nullptr)) return rc;
            dt = now_s() - t0;
        }
        return RG_OK;
    };
    // coordinate descent: one knob at a time over its candidates, keeping
    // the best setting found so far for the others
    rg_tune best_t{1u << 21, stacks[0], 8, 8, 4, 0};
    double best = 0.0;
    if (int rc = measure(best_t, best)) return rc;
    const std::vector<std::vector<uint32_t>> options = {
        {1u << 20, 1u << 21, 1u << 22}, stacks, {1u, 8u, 16u}, {4u, 8u, 16u}, {2u, 4u, 8u}};
    for (size_t knob = 0; knob < options.size(); knob++) {
        for (uint32_t v : options[knob]) {
            rg_tune t = best_t;
            uint32_t *field[] = {&t.pool_paths, &t.stack, &t.tile, &t.blocks_per_cu, &t.iters_per_graph};
            if (*field[knob] == v) continue;
            *field[knob] = v;
            double dt = 0.0;
            if (int rc = measure(t, dt)) return rc;
            if (dt < best) {
                best = dt;
                best_t = t;
            }
        }
    }
    *out = best_t;
    return RG_OK;
}

}  // extern "C"

// rg_engine.cpp — host side of the batched engine: job admission by HBM
// footprint into one device arena, packing of the admitted jobs' chunks into
// shared generations, hipGraph replay of the bounce loop, completion
// tracking by generation events, batch rendering (rg_render_batch), the
// request server (rg_server_*) and the autotuner (rg_autotune).
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
    uint32_t group = 1;    // samples per pixel per chunk (the single-job path's samples in flight)
    uint32_t issued = 0;   // samples per pixel put into generations so far
    uint64_t bytes = 0;    // arena footprint
    uint64_t off = 0;
    uint32_t slot = 0;
    std::atomic<int> state{0};  // 0 pending, 1 admitted, 2 all samples issued, 3 done
    int status = RG_OK;
    double t_submit = 0.0, t_admit = 0.0, t_done = 0.0;
    int pin = -1;           // pinned staging slab of the results (-1: none)
    uint64_t pin_off = 0, pin_bytes = 0;
    float *d_hdr = nullptr, *d_sum = nullptr, *d_sumsq = nullptr;
};

// Pinned host memory for result copies, carved out of a few large slabs:
// one hipHostMalloc per job cost more than a small job's whole render.
struct PinnedSlab {
    char *p = nullptr;
    Arena arena;
};

// Footprint of one job in the arena: sums, sums of squares, HDR image.
uint64_t job_bytes(uint64_t npix, uint64_t) { return 2 * align_up(npix * 12) + align_up(npix * 16); }

// The single-job path's samples in flight, which fixes its accumulation
// grouping (rg_render).
uint32_t plain_group(const rg_settings &st) {
    const uint32_t npix = st.width * st.height;
    const uint32_t spp = std::max<uint32_t>(1, st.spp);
    uint32_t S = st.samples_in_flight;
    if (S == 0) S = std::max<uint32_t>(1, kPathsInFlight / npix);
    return std::min(S, spp);
}

// A generation launched and not yet known to be finished: its event, its
// lane and the jobs whose last samples it carried.
struct Inflight {
    hipEvent_t ev = nullptr;
    int lane = 0;
    std::vector<Request *> finishing;
};

// One stream with its own path state: consecutive generations rotate over
// the lanes, so the tail bounces of one overlap the next one's work.
struct Lane {
    hipStream_t stream = nullptr;
    BatchArgs args{};
    void *d_chunks = nullptr;
    GenChunk *h_chunks = nullptr;          // pinned chunk-table staging
    AccGroup *h_groups = nullptr;          // pinned accumulation-group staging
    AccGroup *d_groups = nullptr;
    unsigned long long *h_stats = nullptr;  // pinned counter snapshot
    hipEvent_t stage_ev = nullptr;          // staging of the last generation consumed
    hipEvent_t acc_ev = nullptr;            // sums and results of the last generation queued
    unsigned long long stats[3] = {0, 0, 0};
};

}  // namespace

class Engine {
  public:
    ~Engine() { shutdown(); }

    // paths_hint (0: unknown): the most path slots a generation can use, the
    // total samples of a known job list.
    // pinned_hint (0: unknown): bytes of results the jobs will copy out.
    // lanes_hint (0: unknown): the most generations a known job list needs.
    int init(Ctx *c, rg_scene *const *sc, uint32_t n, uint64_t budget, const rg_tune *t, uint64_t paths_hint = 0,
             uint64_t pinned_hint = 0, uint32_t lanes_hint = 0) {
        ctx = c;
        if (pinned_hint) next_slab = std::min<uint64_t>(std::max<uint64_t>(pinned_hint, 1u << 20), 2ull << 30);
        if (!c || (n && !sc)) return set_error(RG_E_INVALID, "null argument");
        if (c->device < 0) return set_error(RG_E_NODEVICE, "host-only context has no device");
        if (n == 0) return set_error(RG_E_INVALID, "no scenes");
        if (int rc = dev_set_device(c->device)) return rc;
        // the lanes' streams belong to the context: creating a stream costs
        // milliseconds, more than a small batch's render
        for (int l = 0; l < kMaxLanes; l++) {
            if (!c->lane_streams[l])
                if (int rc = hip_ok(hipStreamCreate((hipStream_t *)&c->lane_streams[l]), "stream create")) return rc;
            lanes[l].stream = (hipStream_t)c->lane_streams[l];
        }
        hipStream_t stream = lanes[0].stream;
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
            scene_bvh.push_back(stream_job_bvh(*s));
        }
        tune = {kPathsInFlight, 0, 0, 0, (uint32_t)width};
        n_lanes = 3;
        if (t) {
            if (t->lanes) n_lanes = (int)t->lanes;
            if (t->pool_paths) tune.budget = t->pool_paths;
            if (t->stack) tune.stack = t->stack;
            if (t->blocks_per_cu) tune.blocks_per_cu = t->blocks_per_cu;
            tune.graphs = t->iters_per_graph ? 1 : 0;
        }
        if (tune.stack == 0)
            for (uint32_t s : {32u, 64u, 128u, 256u})
                if (stream_stack_ok(s, need)) {
                    tune.stack = s;
                    break;
                }
        if (!stream_stack_ok(tune.stack, need))
            return set_error(RG_E_UNSUPPORTED, "traversal stack too small for the deepest scene BVH");
        if (tune.budget < 1024 || tune.budget > (1u << 26)) return set_error(RG_E_INVALID, "pool_paths out of range");
        if (tune.blocks_per_cu > 32) return set_error(RG_E_INVALID, "blocks_per_cu out of range");
        if (n_lanes < 1 || n_lanes > kMaxLanes) return set_error(RG_E_INVALID, "lanes must be 1 to 4");
        if (lanes_hint) n_lanes = std::min<int>(n_lanes, (int)lanes_hint);
        n_cu = stream_cu_count(c->device);
        if (n_cu <= 0) return set_error(RG_E_HIP, "cannot query compute units");

        // path slots: a generation holds up to `budget` paths, and always at
        // least one whole chunk of the default grouping
        cap = std::max(tune.budget, kPathsInFlight);
        if (paths_hint) cap = (uint32_t)std::min<uint64_t>(cap, std::max<uint64_t>(paths_hint, 1024));
        const uint64_t P = cap;
        // one device block for every lane's path state, chunk table and the
        // job records, and one pinned block for their host staging (each
        // allocation call costs about as much as a small job's render)
        static const uint64_t kLaneArrays[] = {32, 16, 16, 16, 16, 16, 32 * kNeeSlots, 4 * kNeeSlots, kNeeSlots,
                                               16 * kNeeSlots, 4, 4, 4, 4, 4};
        uint64_t lane_bytes = 2 * align_up(64) + align_up((uint64_t)kMaxChunks * sizeof(GenChunk)) +
                              align_up((uint64_t)kMaxChunks * sizeof(AccGroup));
        for (uint64_t b : kLaneArrays) lane_bytes += align_up(P * b);
        const uint64_t dev_bytes = n_lanes * lane_bytes + align_up((uint64_t)kMaxJobs * kDevJobBytes);
        char *blk = nullptr;
        if (int rc = dev_malloc((void **)&blk, dev_bytes)) return rc;
        dev_allocs.push_back(blk);
        auto carve = [&](uint64_t bytes) {
            char *q = blk;
            blk += align_up(bytes);
            return q;
        };
        const uint64_t host_lane = align_up((uint64_t)kMaxChunks * sizeof(GenChunk)) +
                                   align_up((uint64_t)kMaxChunks * sizeof(AccGroup)) + align_up(64);
        char *hblk = nullptr;
        if (int rc = hip_ok(hipHostMalloc((void **)&hblk, align_up((uint64_t)kMaxJobs * kDevJobBytes) +
                                                             n_lanes * host_lane),
                            "pinned alloc"))
            return rc;
        h_jobs = hblk;
        hblk += align_up((uint64_t)kMaxJobs * kDevJobBytes);
        for (int l = 0; l < n_lanes; l++) {
            Lane &ln = lanes[l];
            PTState &st = ln.args.st;
            std::memset(&st, 0, sizeof st);
            st.ray = (rg_ray *)carve(P * 32);
            st.hit = (rg_hit *)carve(P * 16);
            st.beta = (float4 *)carve(P * 16);
            st.tp = (float4 *)carve(P * 16);
            st.L = (float4 *)carve(P * 16);
            st.meta = (int4 *)carve(P * 16);
            st.sray = (rg_ray *)carve(P * 32 * kNeeSlots);
            st.slot_stride = (uint32_t)P;
            st.ssrc = (uint32_t *)carve(P * 4 * kNeeSlots);
            st.socc = (uint8_t *)carve(P * kNeeSlots);
            st.scontrib = (float4 *)carve(P * 16 * kNeeSlots);
            st.nshadow = (uint32_t *)carve(P * 4);
            st.q[0] = (uint32_t *)carve(P * 4);
            st.q[1] = (uint32_t *)carve(P * 4);
            ln.args.chunk_of = (uint32_t *)carve(P * 4);
            ln.args.mixed = (uint32_t *)carve(P * 4);
            st.count = (uint32_t *)carve(64);
            st.stats = (unsigned long long *)carve(64);
            ln.d_chunks = carve((uint64_t)kMaxChunks * sizeof(GenChunk));
            ln.d_groups = (AccGroup *)carve((uint64_t)kMaxChunks * sizeof(AccGroup));
            ln.args.chunks = (const GenChunk *)ln.d_chunks;
            ln.args.cap = cap;
            if (int rc = launch_fill_u32(st.stats, 0u, 16, stream)) return rc;
            ln.h_chunks = (GenChunk *)hblk;
            ln.h_groups = (AccGroup *)(hblk + align_up((uint64_t)kMaxChunks * sizeof(GenChunk)));
            ln.h_stats = (unsigned long long *)((char *)ln.h_groups + align_up((uint64_t)kMaxChunks * sizeof(AccGroup)));
            hblk += host_lane;
            for (hipEvent_t *e : {&ln.stage_ev, &ln.acc_ev}) {
                if (int rc = hip_ok(hipEventCreateWithFlags(e, hipEventDisableTiming), "event")) return rc;
                if (int rc = hip_ok(hipEventRecord(*e, stream), "event record")) return rc;
            }
        }
        d_jobs = carve((uint64_t)kMaxJobs * kDevJobBytes);
        if (budget == 0) budget = 16ull << 30;
        if (int rc = dev_malloc((void **)&d_arena, budget)) return rc;
        dev_allocs.push_back(d_arena);
        arena.reset(budget);
        if (int rc = hip_ok(hipEventCreateWithFlags(&adm_ev, hipEventDisableTiming), "event")) return rc;
        for (uint32_t s = kMaxJobs; s-- > 0;) free_slots.push_back(s);
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
        sync_lanes();
        {
            std::lock_guard<std::mutex> lk(mu);
            for (auto &kv : requests) {
                Request *r = kv.second.get();
                if (r->state != 3) {
                    r->state = 3;
                    r->status = RG_E_CANCELLED;
                }
            }
        }
        cv_done.notify_all();
        for (auto &g : graphs)
            if (g.second) (void)hipGraphExecDestroy(g.second);
        graphs.clear();
        for (Inflight &f : inflight)
            if (f.ev) (void)hipEventDestroy(f.ev);
        inflight.clear();
        for (hipEvent_t e : spare_events) (void)hipEventDestroy(e);
        spare_events.clear();
        for (void *p : dev_allocs) dev_free(p);
        dev_allocs.clear();
        if (h_jobs) (void)hipHostFree(h_jobs);
        h_jobs = nullptr;
        for (PinnedSlab &b : slabs)
            if (b.p) (void)hipHostFree(b.p);
        slabs.clear();
        if (adm_ev) (void)hipEventDestroy(adm_ev);
        adm_ev = nullptr;
        for (Lane &ln : lanes) {
            // h_chunks / h_stats are parts of the h_jobs block
            for (hipEvent_t e : {ln.stage_ev, ln.acc_ev})
                if (e) (void)hipEventDestroy(e);
            ln = Lane{};  // the stream stays with the context
        }
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
        if (int rc = check_shader(st)) return rc;
        if (st.shader != 5)
            return set_error(RG_E_UNSUPPORTED, "batched jobs use shading system 5; render systems 0-3 with rg_render");
        const uint64_t npix = (uint64_t)st.width * st.height;
        if (npix * st.spp >= (1ull << 32)) return set_error(RG_E_BUDGET, "job has 2^32 or more samples");
        const uint32_t group = plain_group(st);
        if (npix * group > cap) return set_error(RG_E_BUDGET, "job's samples in flight exceed the engine's path slots");
        const uint64_t bytes = job_bytes(npix, st.spp);
        if (bytes > arena.size()) return set_error(RG_E_BUDGET, "job footprint exceeds the engine's HBM budget");
        auto r = std::make_unique<Request>();
        r->scene = job.scene;
        r->st = st;
        r->out_hdr = hdr;
        r->out_sum = sum;
        r->out_sumsq = sumsq;
        r->npix = (uint32_t)npix;
        r->group = group;
        r->samples = npix * st.spp;
        r->bytes = bytes;
        r->t_submit = now_s();
        {
            std::lock_guard<std::mutex> lk(mu);
            // overload policy: with a limit set, a request whose samples would
            // take the work not yet done past it is refused (a request always
            // runs when nothing else is outstanding)
            if (queue_limit && outstanding > 0 && outstanding + r->samples > queue_limit) {
                n_rejected++;
                return set_error(RG_E_BUSY, "server busy: " + std::to_string(outstanding) +
                                                " samples outstanding, limit " + std::to_string(queue_limit));
            }
            outstanding += r->samples;
            r->ticket = next_ticket++;
            if (ticket) *ticket = r->ticket;
            inbox.push_back(r.get());
            requests[r->ticket] = std::move(r);
        }
        cv_work.notify_all();
        return RG_OK;
    }

    // One round of the loop: admit, launch the next generation, deliver.
    int step() {
        // rg_cancel fails what is queued and running; later requests run
        if (ctx->cancel.exchange(false)) return fail_all(RG_E_CANCELLED, "cancelled");
        {
            std::lock_guard<std::mutex> lk(mu);
            while (!inbox.empty()) {
                pending.push_back(inbox.front());
                inbox.pop_front();
            }
        }
        if (int rc = admit()) return fail_all(rc, nullptr);
        // one generation running per lane and one queued: later requests
        // join the next generation instead of queueing behind everything
        // already launched
        while (inflight.size() >= (size_t)n_lanes + 1)
            if (int rc = retire(true)) return fail_all(rc, nullptr);
        bool launched = false;
        if (int rc = launch_generation(launched)) return fail_all(rc, nullptr);
        if (int rc = retire(!launched)) return fail_all(rc, nullptr);
        return RG_OK;
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
        out[2] = dev_stats[0].load();
        out[3] = dev_stats[1].load();
        out[4] = n_generations.load();
        out[5] = fb_used.load();
        out[6] = arena.size();
        out[7] = n_replays.load();
    }

    void set_queue_limit(uint64_t limit) {
        std::lock_guard<std::mutex> lk(mu);
        queue_limit = limit;
    }
    void get_queue(uint64_t out[3]) {
        std::lock_guard<std::mutex> lk(mu);
        out[0] = outstanding;
        out[1] = queue_limit;
        out[2] = n_rejected;
    }

    int lanes_used() const { return n_lanes; }

    StreamTune tune{};
    // written by the engine loop, read by rg_server_stats from any thread
    std::atomic<uint64_t> dev_stats[4] = {{0}, {0}, {0}, {0}};

  private:
    // Uploads go to the stream of the lane the next generation runs on;
    // adm_ev, recorded after them, orders the other lane's next generation.
    int admit() {
        hipStream_t stream = lanes[next_lane].stream;
        bool any = false;
        while (!pending.empty() && !free_slots.empty()) {
            Request *r = pending.front();
            uint64_t off = 0;
            if (!arena.alloc(r->bytes, off)) break;  // first come, first served
            fb_used = arena.used();
            pending.pop_front();
            r->off = off;
            r->slot = free_slots.back();
            free_slots.pop_back();
            char *base = d_arena + off;
            const uint64_t np = r->npix;
            JobDesc d;
            d.scene = scenes[r->scene];
            d.st = r->st;
            d.d_sum = (float *)base;
            d.d_sumsq = (float *)(base + align_up(np * 12));
            d.d_hdr = (float *)((char *)d.d_sumsq + align_up(np * 12));
            // the slot's pinned record is free: its last job was delivered,
            // so the upload of that job's record has completed
            char *rec = h_jobs + (size_t)r->slot * kDevJobBytes;
            stream_job_record(d, rec);
            if (int rc = hip_ok(hipMemcpyAsync((char *)d_jobs + (size_t)r->slot * kDevJobBytes, rec, kDevJobBytes,
                                               hipMemcpyHostToDevice, stream),
                                "job upload"))
                return rc;
            // sums and sums of squares are adjacent in the arena: one fill
            if (int rc = launch_fill_u32(d.d_sum, 0u, (align_up(np * 12) + np * 12) / 4, stream)) return rc;
            r->d_hdr = d.d_hdr;
            r->d_sum = d.d_sum;
            r->d_sumsq = d.d_sumsq;
            r->state = 1;
            r->t_admit = now_s();
            active.push_back(r);
            any = true;
        }
        if (any) {
            // earlier admissions this lane has not waited for are folded in,
            // so the new event covers every upload so far
            if (int rc = wait_admissions(next_lane)) return rc;
            if (int rc = hip_ok(hipEventRecord(adm_ev, stream), "event record")) return rc;
            lane_adm_seq[next_lane] = ++adm_seq;
        }
        return RG_OK;
    }

    // Packs the next chunks of the admitted jobs of one scene (the scene of
    // the oldest job with samples left; first come, first served; a job's
    // chunks in sample order) into one generation and launches it.
    int launch_generation(bool &launched) {
        launched = false;
        std::vector<GenChunk> ch;
        std::vector<Request *> last;  // jobs whose final chunk is in this generation
        uint64_t total = 0;
        uint32_t bounces = 0;
        uint32_t scene = ~0u;
        for (Request *r : active)
            if (r->state == 1) {
                scene = r->scene;
                break;
            }
        for (Request *r : active) {
            if (r->state != 1 || r->scene != scene) continue;
            while (r->issued < r->st.spp && ch.size() < kMaxChunks) {
                const uint32_t S = std::min(r->group, r->st.spp - r->issued);
                const uint64_t paths = (uint64_t)r->npix * S;
                if (!ch.empty() && total + paths > tune.budget) break;
                if (total + paths > cap) break;
                GenChunk c;
                std::memset(&c, 0, sizeof c);
                c.job = r->slot;
                c.first = (uint32_t)total;
                c.npix = r->npix;
                c.width = r->st.width;
                c.sample_base = r->issued;
                c.S = S;
                c.bounces = (uint32_t)path_events(*scenes[r->scene], r->st.max_depth);
                c.clamp = r->st.firefly_clamp;
                ch.push_back(c);
                total += paths;
                r->issued += S;
                bounces = std::max(bounces, c.bounces);
                chunk_jobs.push_back(r);
            }
            if (r->issued >= r->st.spp) {
                r->state = 2;
                last.push_back(r);
            }
            if (ch.size() >= kMaxChunks || total >= tune.budget) break;
        }
        if (ch.empty()) {
            chunk_jobs.clear();
            return RG_OK;
        }
        const int L = next_lane;
        Lane &ln = lanes[L];
        hipStream_t stream = ln.stream;
        // the lane's pinned chunk table and counter snapshot are free again
        if (int rc = hip_ok(hipEventSynchronize(ln.stage_ev), "chunk staging")) return rc;
        std::memcpy(ln.h_chunks, ch.data(), ch.size() * sizeof(GenChunk));
        // one accumulation group per job (its chunks are consecutive)
        uint32_t ngroups = 0, total_pix = 0;
        for (size_t k = 0; k < ch.size(); k++) {
            if (k == 0 || chunk_jobs[k] != chunk_jobs[k - 1]) {
                AccGroup &g = ln.h_groups[ngroups++];
                std::memset(&g, 0, sizeof g);
                g.c0 = (uint32_t)k;
                g.pix_first = total_pix;
                g.npix = ch[k].npix;
                g.finish = chunk_jobs[k]->state == 2 ? 1u : 0u;
                total_pix += ch[k].npix;
            }
            ln.h_groups[ngroups - 1].c1 = (uint32_t)k + 1;
        }
        if (int rc = wait_admissions(L)) return rc;
        if (int rc = hip_ok(hipMemcpyAsync(ln.d_chunks, ln.h_chunks, ch.size() * sizeof(GenChunk),
                                           hipMemcpyHostToDevice, stream),
                            "chunk upload"))
            return rc;
        if (int rc = hip_ok(hipMemcpyAsync(ln.d_groups, ln.h_groups, ngroups * sizeof(AccGroup),
                                           hipMemcpyHostToDevice, stream),
                            "group upload"))
            return rc;
        BatchArgs a = ln.args;
        a.nchunks = (uint32_t)ch.size();
        if (int rc = stream_launch_camera(a, d_jobs, (uint32_t)total, stream)) return rc;
        if (int rc = launch_bounces(L, a, scene, (uint32_t)total, bounces)) return rc;
        // a job's chunks are summed in generation order across the lanes
        if (last_lane >= 0 && last_lane != L)
            if (int rc = hip_ok(hipStreamWaitEvent(stream, lanes[last_lane].acc_ev, 0), "stream wait")) return rc;
        if (int rc = stream_launch_accumulate(a, d_jobs, ln.d_groups, ngroups, total_pix, stream)) return rc;
        chunk_jobs.clear();
        for (Request *r : last)
            if (int rc = finish(r, stream)) return rc;
        if (int rc = hip_ok(hipEventRecord(ln.acc_ev, stream), "event record")) return rc;
        if (int rc = hip_ok(hipMemcpyAsync(ln.h_stats, ln.args.st.stats, 32, hipMemcpyDeviceToHost, stream),
                            "stats copy"))
            return rc;
        if (int rc = hip_ok(hipEventRecord(ln.stage_ev, stream), "event record")) return rc;
        Inflight f;
        if (!spare_events.empty()) {
            f.ev = spare_events.back();
            spare_events.pop_back();
        } else if (int rc = hip_ok(hipEventCreateWithFlags(&f.ev, hipEventDisableTiming), "event")) {
            return rc;
        }
        if (int rc = hip_ok(hipEventRecord(f.ev, stream), "event record")) return rc;
        f.lane = L;
        f.finishing = std::move(last);
        inflight.push_back(std::move(f));
        n_generations++;
        last_lane = L;
        next_lane = (L + 1) % n_lanes;
        launched = true;
        return RG_OK;
    }

    // The bounce loop, replayed from a captured graph when tune.graphs is
    // set.  Its kernels read the queue lengths from device counters and the
    // chunks through the per-path chunk index, so one capture serves every
    // generation of a scene with the same bounce count and slot count
    // rounded up to a power of two (the extra blocks find no work).
    int launch_bounces(int L, const BatchArgs &a, uint32_t scene, uint32_t paths, uint32_t bounces) {
        hipStream_t stream = lanes[L].stream;
        const DevBvh &bv = scene_bvh[scene];
        const bool ext = scene_uses_ext(*scenes[scene]);
        if (!tune.graphs) return stream_launch_bounces(a, d_jobs, bv, paths, bounces, tune, n_cu, ext, stream);
        uint32_t slots = 1024;
        while (slots < paths && slots < cap) slots <<= 1;
        slots = std::min(slots, cap);
        const uint64_t key = ((uint64_t)L << 56) | ((uint64_t)scene << 40) | ((uint64_t)bounces << 32) | slots;
        auto it = graphs.find(key);
        if (it == graphs.end()) {
            hipGraph_t g = nullptr;
            if (int rc = hip_ok(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed), "begin capture"))
                return rc;
            const int lrc = stream_launch_bounces(a, d_jobs, bv, slots, bounces, tune, n_cu, ext, stream);
            const hipError_t ce = hipStreamEndCapture(stream, &g);
            if (lrc) return lrc;
            if (int rc = hip_ok(ce, "end capture")) return rc;
            hipGraphExec_t ex = nullptr;
            const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (int rc = hip_ok(ie, "graph instantiate")) return rc;
            it = graphs.emplace(key, ex).first;
        }
        n_replays++;
        return hip_ok(hipGraphLaunch(it->second, stream), "graph launch");
    }

    // Queues the copies of a finished job's results (its HDR image is
    // written by the generation's accumulation).
    int finish(Request *r, hipStream_t stream) {
        const size_t hb = (size_t)r->npix * 16, sb = (size_t)r->npix * 12;
        const size_t need = (r->out_hdr ? hb : 0) + (r->out_sum ? sb : 0) + (r->out_sumsq ? sb : 0);
        if (!need) return RG_OK;
        if (!get_pinned(need, r->pin, r->pin_off)) return set_error(RG_E_NOMEM, "pinned staging allocation failed");
        r->pin_bytes = need;
        char *p = slabs[r->pin].p + r->pin_off;
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
            if (int rc = hip_ok(hipMemcpyAsync(p, r->d_sumsq, sb, hipMemcpyDeviceToHost, stream), "result copy"))
                return rc;
        return RG_OK;
    }

    // First fit in the existing slabs, else a new slab (the first sized by
    // the hint, then doubling up to 1 GiB, at least `need`).
    bool get_pinned(size_t need, int &slab, uint64_t &off) {
        for (size_t i = 0; i < slabs.size(); i++)
            if (slabs[i].arena.alloc(need, off)) {
                slab = (int)i;
                return true;
            }
        const uint64_t size = std::max<uint64_t>(align_up(need), next_slab);
        next_slab = std::min<uint64_t>(2 * next_slab, 1ull << 30);
        PinnedSlab b;
        if (hipHostMalloc((void **)&b.p, size) != hipSuccess) return false;
        b.arena.reset(size);
        slabs.push_back(std::move(b));
        slab = (int)slabs.size() - 1;
        return slabs[slab].arena.alloc(need, off);
    }

    // Retires finished generations in launch order and hands their finished
    // jobs to the waiters; with `block`, waits for the oldest one.
    int retire(bool block) {
        while (!inflight.empty()) {
            Inflight &f = inflight.front();
            hipError_t q = hipEventQuery(f.ev);
            if (q == hipErrorNotReady && block) q = hipEventSynchronize(f.ev);
            if (q == hipErrorNotReady) return RG_OK;
            if (int rc = hip_ok(q, "generation wait")) return rc;
            Lane &ln = lanes[f.lane];
            for (int k = 0; k < 3; k++) ln.stats[k] = ln.h_stats[k];
            for (int k = 0; k < 3; k++) {
                uint64_t v = 0;
                for (const Lane &x : lanes) v += x.stats[k];
                dev_stats[k] = v;
            }
            dev_stats[3] = n_generations.load();
            for (Request *r : f.finishing) deliver(r);
            spare_events.push_back(f.ev);
            inflight.pop_front();
            block = false;
        }
        return RG_OK;
    }

    void deliver(Request *r) {
        if (r->pin >= 0) {
            const char *p = slabs[r->pin].p + r->pin_off;
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
            slabs[r->pin].arena.release(r->pin_off, r->pin_bytes);
            r->pin = -1;
        }
        arena.release(r->off, r->bytes);
        fb_used = arena.used();
        free_slots.push_back(r->slot);
        active.erase(std::find(active.begin(), active.end(), r));
        {
            std::lock_guard<std::mutex> lk(mu);
            r->t_done = now_s();
            r->state = 3;
            n_done++;
            samples_done += r->samples;
            outstanding -= r->samples;
        }
        cv_done.notify_all();
    }

    int fail_all(int rc, const char *msg) {
        std::string text = msg ? std::string(msg) : std::string(rg_last_error());
        sync_lanes();
        for (Inflight &f : inflight) spare_events.push_back(f.ev);
        inflight.clear();
        chunk_jobs.clear();
        // release before the requests are marked done: a waiter may free a
        // request as soon as it is
        for (Request *r : active) {
            arena.release(r->off, r->bytes);
            free_slots.push_back(r->slot);
        }
        fb_used = arena.used();
        active.clear();
        {
            std::lock_guard<std::mutex> lk(mu);
            failure = text;
            for (auto &kv : requests) {
                Request *r = kv.second.get();
                if (r->state == 3) continue;
                r->state = 3;
                r->status = rc;
                outstanding -= r->samples;
                r->t_done = now_s();
                if (r->pin >= 0) slabs[r->pin].arena.release(r->pin_off, r->pin_bytes);
                r->pin = -1;
            }
            inbox.clear();
        }
        pending.clear();
        cv_done.notify_all();
        return set_error(rc, text);
    }

    Ctx *ctx = nullptr;
    std::vector<Scene *> scenes;
    std::vector<DevBvh> scene_bvh;
    // Orders lane L after every admission upload queued so far.
    int wait_admissions(int L) {
        if (lane_adm_seq[L] == adm_seq) return RG_OK;
        lane_adm_seq[L] = adm_seq;
        return hip_ok(hipStreamWaitEvent(lanes[L].stream, adm_ev, 0), "stream wait");
    }

    void sync_lanes() {
        for (Lane &ln : lanes)
            if (ln.stream) (void)hipStreamSynchronize(ln.stream);
    }

    int n_cu = 0;
    uint32_t cap = 0;
    static constexpr int kMaxLanes = 4;
    Lane lanes[kMaxLanes];
    int n_lanes = 3, next_lane = 0, last_lane = -1;
    hipEvent_t adm_ev = nullptr;  // after every admission upload so far
    uint64_t adm_seq = 0;         // admissions recorded into adm_ev
    uint64_t lane_adm_seq[kMaxLanes] = {0, 0, 0, 0};  // the last of them each lane is ordered after
    std::map<uint64_t, hipGraphExec_t> graphs;  // bounce loops by lane, scene, bounces, slots
    void *d_jobs = nullptr;
    char *d_arena = nullptr;
    std::vector<void *> dev_allocs;
    Arena arena;
    char *h_jobs = nullptr;
    std::deque<Inflight> inflight;
    std::vector<hipEvent_t> spare_events;
    std::vector<Request *> chunk_jobs;  // job of each chunk of the generation being built
    std::vector<PinnedSlab> slabs;
    uint64_t next_slab = 64ull << 20;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> inbox;    // submitted, not yet seen by the loop
    std::deque<Request *> pending;  // waiting for a slot and arena space
    std::vector<Request *> active;  // admitted, in admission order
    std::map<uint64_t, std::unique_ptr<Request>> requests;
    std::vector<uint32_t> free_slots;
    uint64_t next_ticket = 1;
    bool stop = false;
    std::thread worker;
    std::string failure;
    uint64_t n_done = 0, samples_done = 0;
    // samples of the requests submitted and not done, the limit on them
    // (0: none) and the requests refused by it; guarded by mu
    uint64_t outstanding = 0, queue_limit = 0, n_rejected = 0;
    std::atomic<uint64_t> n_generations{0}, n_replays{0}, fb_used{0};
};

}  // namespace rg

namespace rg {

// Destroys the lane streams a context's engines used (rg_destroy).
void release_engine(Ctx *c) {
    for (void *&s : c->lane_streams) {
        if (s) (void)hipStreamDestroy((hipStream_t)s);
        s = nullptr;
    }
}

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

// A job list's samples per scene: a generation (chunks of one scene) holds
// at most the largest of them, and the list needs at most the sum over the
// scenes of ceil(samples / generation budget) generations.
void batch_shape(const rg_job *jobs, uint32_t n, const rg_tune *tune, uint64_t &paths, uint32_t &gens) {
    std::map<uint32_t, uint64_t> per;
    for (uint32_t i = 0; i < n; i++)
        per[jobs[i].scene] += (uint64_t)jobs[i].settings.width * jobs[i].settings.height * jobs[i].settings.spp;
    const uint64_t budget = (tune && tune->pool_paths) ? tune->pool_paths : kPathsInFlight;
    paths = 1;
    uint64_t g = 0;
    for (auto &kv : per) {
        paths = std::max(paths, kv.second);
        g += (kv.second + budget - 1) / budget;
    }
    gens = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, 1u << 30));
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
    c->cancel.store(false);  // as rg_render: a batch clears the flag when it starts
    Engine e;
    uint64_t out_bytes = 0;
    for (uint32_t i = 0; i < n_jobs; i++) {
        const uint64_t np = (uint64_t)jobs[i].settings.width * jobs[i].settings.height;
        out_bytes += align_up((hdr_out && hdr_out[i] ? np * 16 : 0) + (sum_out && sum_out[i] ? np * 12 : 0));
    }
    uint64_t paths = 0;
    uint32_t gens = 0;
    batch_shape(jobs, n_jobs, tune, paths, gens);
    if (int rc = e.init(c, scenes, n_scenes, batch_budget(jobs, n_jobs), tune, paths, out_bytes, gens)) return rc;
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

int rg_server_set_queue_limit(rg_server *srv, uint64_t max_outstanding_samples) {
    if (!srv) return set_error(RG_E_INVALID, "null server");
    srv->eng->set_queue_limit(max_outstanding_samples);
    return RG_OK;
}

int rg_server_queue(rg_server *srv, uint64_t out[3]) {
    if (!srv || !out) return set_error(RG_E_INVALID, "null argument");
    srv->eng->get_queue(out);
    return RG_OK;
}

int rg_server_get_tune(rg_server *srv, rg_tune *out) {
    if (!srv || !out) return set_error(RG_E_INVALID, "null argument");
    const StreamTune &t = srv->eng->tune;
    *out = {t.budget, t.stack, (uint32_t)srv->eng->lanes_used(), t.blocks_per_cu, t.graphs ? 1u : 0u, 0};
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
        uint64_t paths = 0;
        uint32_t gens = 0;
        batch_shape(jobs.data(), (uint32_t)jobs.size(), &t, paths, gens);
        if (int rc = e.init(c, scenes, n_scenes, budget, &t, paths, 0, gens)) return rc;
        for (int rep = 0; rep < 2; rep++) {
            const double t0 = now_s();
            if (int rc = batch_on(e, jobs.data(), (uint32_t)jobs.size(), nullptr, nullptr)) return rc;
            dt = now_s() - t0;
        }
        return RG_OK;
    };
    // coordinate descent: one knob at a time over its candidates, keeping
    // the best setting found so far for the others
    rg_tune best_t{kPathsInFlight, stacks[0], 3, 0, 0, 0};
    double best = 0.0;
    if (int rc = measure(best_t, best)) return rc;
    const std::vector<std::vector<uint32_t>> options = {
        {1u << 23, 1u << 24, 1u << 25}, stacks, {1u, 2u, 3u, 4u}, {0u, 4u, 8u, 16u}, {0u, 1u}};
    for (size_t knob = 0; knob < options.size(); knob++) {
        for (uint32_t v : options[knob]) {
            rg_tune t = best_t;
            uint32_t *field[] = {&t.pool_paths, &t.stack, &t.lanes, &t.blocks_per_cu, &t.iters_per_graph};
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

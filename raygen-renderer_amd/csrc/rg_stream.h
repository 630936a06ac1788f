// rg_stream.h — the streaming (batched) wavefront engine: many render jobs
// share one pool of paths.  Every iteration puts the next samples of the
// issue sequence (the concatenation of the admitted jobs' samples,
// pass-major within a job) into the positions freed by finished paths, so
// the extend / shade / shadow kernels always see a full pool whatever the
// mix of resolutions, sample counts and depths.  Finished samples land in
// their job's per-sample buffer; a job is resolved once all its samples are
// in, summing them per pixel in the same order and grouping as the
// single-job path, so a job's image does not depend on what ran beside it.
// A job is complete max_depth + 1 iterations after the issue cursor passed
// its last sample (no path lives longer), which the host reads from the
// counter snapshot it takes after every graph replay.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rg_internal.h"
#include "rg_pt.h"

namespace rg {

constexpr uint32_t kMaxJobs = 1024;    // job slots of one engine
constexpr size_t kDevJobBytes = 512;   // device job record

// Knobs picked by the autotuner (mirror of rg_tune).
struct StreamTune {
    uint32_t pool;           // path slots in flight (queue size)
    uint32_t stack;          // traversal stack entries: 32, 64, 128 or 256
    uint32_t tile;           // pixel tile edge of the issue order (1 = scanline)
    uint32_t blocks_per_cu;  // cap on persistent traversal blocks per compute unit
    uint32_t iters;          // iterations per captured graph (even)
    uint32_t width;          // node width of every resident scene's BVH (4 or 8)
};

// Device-side counters; iteration parity qi reads slot [qi] and writes [qi^1].
struct StreamCounters {
    uint64_t cursor[2];            // next sample of the issue sequence
    uint64_t seq_end;              // end of the issue sequence (host-written)
    uint32_t nholes[2];            // free positions in the hole list
    uint32_t nsh[2];               // shadow-ray queue length
    uint32_t fetch[2];             // any-hit dynamic-fetch cursor
    uint32_t n_issue;              // issue-table entries (host-written)
    uint32_t pad[3];
    unsigned long long stats[4];   // extension rays, shadow rays, samples issued, iterations
};
static_assert(sizeof(StreamCounters) % 16 == 0, "counter block keeps 16-byte alignment");

// Issue-table entry: the job whose samples start at `begin` in the issue
// sequence.  Entries are sorted by begin.
struct IssueEntry {
    uint64_t begin;
    uint32_t job;
    uint32_t pad;
};

// Path state, structure of arrays indexed by pool position.  A path keeps
// its position from its camera ray to its last event, so every kernel reads
// and writes element i in thread i; a finished path's position goes to the
// hole list and the next iteration's regen puts a new sample there.
struct PathSoA {
    rg_ray *ray;
    float4 *beta;   // estimator weight, w = bsdfSampledPdf
    float4 *tp;     // reference throughput (Russian roulette)
    float4 *L;      // radiance so far, w = alive after the event
    int4 *meta;     // passes, medium, next LDS dim, next Philox draw
    int4 *sid;      // job slot (-1: free position), pass, pixel, bounces taken
};

// Device pointers of the path pool, passed to every kernel by value.
struct StreamPool {
    uint32_t cap;
    PathSoA st;
    rg_hit *hit;
    rg_ray *sray;       // NEE rays per (position, k), k < 2 (tmin = predicate)
    float4 *scontrib;   // unoccluded contribution per (position, k)
    uint8_t *socc;      // occluded flag per (position, k)
    uint32_t *ssrc;     // queue of the (position, k) pairs with a ray: 2 * position + k
    uint32_t *nshadow;  // shadow rays per position this event
    uint32_t *holes;    // free positions
    StreamCounters *ctr;
    const IssueEntry *issue;
};

// What the host knows of one job when it writes the device record.
struct JobDesc {
    const Scene *scene;
    rg_settings st;
    float *d_samples;  // spp * npix * 3 floats
    float *d_sum, *d_sumsq, *d_hdr;
    uint32_t group;    // samples per pixel per accumulation group
    uint32_t tile;
};

// Writes the device job record (kDevJobBytes) for `d` into `out`.
void stream_job_record(const JobDesc &d, void *out);
// Grid geometry of the persistent kernels.
int stream_cu_count(int device);
// Puts the pool in its initial state: every slot free, queues empty.
int stream_launch_reset(const StreamPool &sp, void *stream);
// One iteration (regenerate, extend, shade, shadow) at parity qi.
int stream_launch_iteration(const StreamPool &sp, const void *d_jobs, int qi, const StreamTune &tune, int n_cu,
                            void *stream);
// Sums a finished job's samples per pixel and writes its HDR image.
int stream_launch_finish(const void *d_jobs, uint32_t slot, uint32_t npix, void *stream);
// Largest stack_need the traversal templates cover for a stack knob value.
bool stream_stack_ok(uint32_t stack, uint32_t need);

}  // namespace rg

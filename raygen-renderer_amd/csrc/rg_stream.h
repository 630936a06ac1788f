// rg_stream.h — the batched engine: many render jobs packed into shared
// wavefront generations.
//
// A generation is a list of chunks; a chunk is one job's samples
// [base, base + S) of every pixel, S being the single-job path's samples in
// flight for that job (its accumulation group, capped by what is left).
// Chunks of any number of jobs of one scene share the generation's path
// slots up to the budget, and the bounce loop (extend, shade, shadow rays)
// runs over one compacted queue of all their paths on the scene's
// tree, each event with its own job's scene record.  After the loop every chunk is
// summed per pixel into its job's sums exactly as the single-job path sums
// one of its generations, so a job's image does not depend on what ran
// beside it.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rg_internal.h"
#include "rg_pt.h"

namespace rg {

constexpr uint32_t kMaxJobs = 1024;    // job slots of one engine
constexpr size_t kDevJobBytes = 512;   // device job record
constexpr uint32_t kMaxChunks = 4096;  // chunks of one generation

// Knobs picked by the autotuner (mirror of rg_tune).
struct StreamTune {
    uint32_t budget;         // path slots of a generation
    uint32_t stack;          // traversal stack entries: 32, 64, 128 or 256
    uint32_t blocks_per_cu;  // cap on persistent traversal blocks per compute unit (0: none)
    uint32_t graphs;         // replay each generation's bounce loop as a captured hipGraph
    uint32_t width;          // node width of every resident scene's BVH (4 or 8)
};

// One chunk of a generation.  Chunks are sorted by `first`.
struct GenChunk {
    uint32_t job;          // job slot
    uint32_t first;        // first path slot of the chunk in the generation
    uint32_t npix, width;  // the job's image
    uint32_t sample_base;  // first sample of every pixel in the chunk
    uint32_t S;            // samples per pixel in the chunk
    uint32_t bounces;      // the job's path events (max_depth + 2): its paths
                           // still alive after that many are closed
    float clamp;           // the job's firefly clamp (0: none)
};

// One job's chunks in a generation, [c0, c1) of the chunk table (a job's
// chunks are consecutive), and its pixels' place in the generation's
// accumulation grid.  finish: the job's last samples are in this
// generation, so its HDR image is written too.
struct AccGroup {
    uint32_t c0, c1;
    uint32_t pix_first, npix;
    uint32_t finish;
    uint32_t pad[3];
};

// Device pointers of a generation, passed to every kernel by value: the
// single-job path's path state, the chunk of each path slot, the chunk table.
struct BatchArgs {
    PTState st;            // count[3]: length of the mixed queue
    uint32_t *chunk_of;
    uint32_t *mixed;       // paths shaded by the mixed-wave shade kernel
    const GenChunk *chunks;
    uint32_t nchunks;
    uint32_t cap;  // path slots of the buffers
};

// What the host knows of one job when it writes the device record.
struct JobDesc {
    const Scene *scene;
    rg_settings st;
    float *d_sum, *d_sumsq, *d_hdr;
};

// Writes the device job record (kDevJobBytes) for `d` into `out`.
void stream_job_record(const JobDesc &d, void *out);
// A scene's tree as the traversal kernels take it.
DevBvh stream_job_bvh(const Scene &s);
// Grid geometry of the persistent kernels.
int stream_cu_count(int device);
// Camera rays of every chunk's samples (queue 0 = all `total` paths).
int stream_launch_camera(const BatchArgs &a, const void *d_jobs, uint32_t total, void *stream);
// The bounce loop (max_bounces events, the largest of the chunks' bounce
// counts; each path stops after its own chunk's); the per-path kernels are sized for `paths` slots
// (at least the generation's paths, at most the buffers' slots).
// ext: the scene needs the extended-media shade kernels (scene_uses_ext)
int stream_launch_bounces(const BatchArgs &a, const void *d_jobs, const DevBvh &bv, uint32_t paths,
                          uint32_t max_bounces, const StreamTune &tune, int n_cu, bool ext, void *stream);
// One launch for the generation: every chunk summed per pixel into its
// job's sums (the single-job path's accumulation of one generation, a job's
// chunks in order), and the HDR image (mean * exposure, clamped at 0) of
// every job that finishes in it.  total_pix = sum of the groups' npix.
int stream_launch_accumulate(const BatchArgs &a, const void *d_jobs, const AccGroup *d_groups, uint32_t ngroups,
                             uint32_t total_pix, void *stream);
// Largest stack_need the traversal templates cover for a stack knob value.
bool stream_stack_ok(uint32_t stack, uint32_t need);

}  // namespace rg

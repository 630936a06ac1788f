// rg_stream.hip — kernels of the batched engine: packed wavefront
// generations of many jobs (see rg_stream.h).
//
// Per generation:
//   camera  one path per (chunk, sample, pixel): the chunk is found by a
//           binary search over the sorted chunk table, the camera ray built
//           with the job's scene record (rayrenderer.cpp:860-934); the queue
//           is the identity over all paths
//   then max_bounces times, over the compacted queue of live paths:
//   extend  closest hit on the generation's tree (every chunk of a
//           generation renders one scene; persistent waves, leaf
//           postponement, lane refill)
//   shade   the previous event's unoccluded NEE contributions, then one
//           path event with the job's scene record, wave-uniform for the
//           lanes of the wave's first job (the same step() as the
//           single-job path, at its register budget); paths of other jobs in
//           the wave are shaded by a second kernel that loops over the
//           distinct jobs; the event's 0-2 NEE rays and the paths still
//           live (not finished, not at their job's last event) are queued
//           with block-aggregated appends
//   anyhit  the queued NEE rays, one per lane
//   then one launch adds each finished path's last NEE contributions,
//   clamps with the job's firefly clamp, sums each chunk per pixel into its
//   job's sums and writes the HDR image of every job that finishes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>

#include "rg_pt_dev.h"
#include "rg_stream.h"

namespace rg {

// Device job record.
struct DevJob {
    PTScene sc;
    float *sum;      // [npix][3]
    float *sumsq;    // [npix][3]
    float *hdr;      // [npix][4]
    uint32_t npix, spp;
    float clamp, exposure;
};
struct DevJobSlot {
    DevJob j;
    uint8_t pad[kDevJobBytes - sizeof(DevJob)];
};
static_assert(sizeof(DevJobSlot) == kDevJobBytes, "job record size");

DevBvh dev_bvh(const Scene &s);

DevBvh stream_job_bvh(const Scene &s) { return dev_bvh(s); }

void stream_job_record(const JobDesc &d, void *out) {
    DevJobSlot rec;
    std::memset(&rec, 0, sizeof rec);
    DevJob &J = rec.j;
    J.sc = make_ptscene(*d.scene, d.st);
    J.sum = d.d_sum;
    J.sumsq = d.d_sumsq;
    J.hdr = d.d_hdr;
    J.npix = d.st.width * d.st.height;
    J.spp = d.st.spp;
    J.clamp = d.st.firefly_clamp;
    J.exposure = d.scene->camera.exposure;
    std::memcpy(out, &rec, sizeof rec);
}

// Runs f with the job record of the calling lane, one distinct job at a time
// across the wave, so the record's fields are wave-uniform inside f.
template <class F>
__device__ __forceinline__ void with_job(const DevJobSlot *__restrict__ jobs, uint32_t job, F &&f) {
    for (;;) {
        const uint32_t j = __builtin_amdgcn_readfirstlane(job);
        if (j == job) {
#if defined(__HIP_DEVICE_COMPILE__)
            // read through the constant address space: the record is uniform
            // across the lanes in here, so it loads with scalar loads
            typedef __attribute__((address_space(4))) const DevJob *ConstJob;
            const DevJob J = *(ConstJob)&jobs[j].j;
#else
            const DevJob &J = jobs[j].j;
#endif
            f(J);
            break;
        }
    }
}

// The last chunk whose first path slot is <= p.
__device__ __forceinline__ uint32_t find_chunk(const GenChunk *__restrict__ ch, uint32_t n, uint32_t p) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (ch[mid].first <= p) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void k_bset_u32(uint32_t *p, uint32_t v) { *p = v; }

__global__ __launch_bounds__(kPtBlock) void k_bcamera(BatchArgs a, const DevJobSlot *__restrict__ jobs,
                                                       uint32_t total) {
    const uint32_t i = blockIdx.x * kPtBlock + threadIdx.x;
    if (i >= total) return;
    const PTState &st = a.st;
    const uint32_t c = find_chunk(a.chunks, a.nchunks, i);
    const GenChunk &C = a.chunks[c];
    const uint32_t local = i - C.first;
    const uint32_t pix = local % C.npix;
    const int sample = (int)(C.sample_base + local / C.npix);
    with_job(jobs, C.job, [&](const DevJob &J) {
        Rng R = rng_for(J.sc, pix, sample, 0, 0);
        v3 o, d;
        camera_ray(J.sc, R, pix % J.sc.width, pix / J.sc.width, o, d);
        st.ray[i] = {{o.x, o.y, o.z}, 0.0f, {d.x, d.y, d.z}, MAX_RAY_DISTANCE};
        st.beta[i] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        st.tp[i] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        st.L[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        st.meta[i] = make_int4(-1, pack_medium(J.sc.global_medium, -1, false), R.dim, (int)R.k);
        st.nshadow[i] = 0;
    });
    st.q[0][i] = i;
    a.chunk_of[i] = c;
    if (i == 0) atomicAdd(st.stats + 2, (unsigned long long)total);
}

// Closest hit of every live path (queue qi) on the generation's tree (every
// chunk of a generation renders the same scene).
struct BatchExtendSrc {
    const rg_ray *ray;
    const uint32_t *q;
    rg_hit *hit;
    __device__ __forceinline__ bool load(uint64_t i, const DevBvh &, PsRay &r) const {
        const rg_ray x = ray[q[i]];
        r = {x.o[0], x.o[1], x.o[2], x.d[0], x.d[1], x.d[2], x.tmin, x.tmax, 0u};
        return true;
    }
    __device__ __forceinline__ void store_closest(uint64_t i, uint32_t prim, float t, float u, float v) const {
        hit[q[i]] = {prim, t, u, v};
    }
    __device__ __forceinline__ void store_any(uint64_t, bool) const {}
};

// Block 0 also clears the shadow-ray queue, the mixed queue and the next
// path queue, which nothing reads while this kernel runs.
template <int kStack, int kW>
__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(7))) void k_bextend(DevBvh bv, BatchArgs a, int qi) {
    RG_LDS_COLUMN(lc, kPtBlock);
    const PTState &st = a.st;
    const uint32_t count = st.count[qi];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st.count[2] = 0;
        st.count[3] = 0;
        st.count[qi ^ 1] = 0;
        if (count) atomicAdd(st.stats + 0, (unsigned long long)count);
    }
    trace_ps_loop<kStack, kW, false, kPtBlock>(bv, BatchExtendSrc{st.ray, st.q[qi], st.hit}, count, lc);
}

// One path event of live path p, one of chunk C's paths, with its job's
// scene record, the bounce-th of the generation.  The previous event's
// unoccluded NEE contributions are added first (in event order); the path
// stays live (alive set) unless the event ended it or it was its job's last
// event (bounce + 1 = max_depth + 2): a closed path keeps its radiance and
// its last event's NEE rays for the accumulation, as the single-job path
// closes its paths.
template <bool kExt>
__device__ __forceinline__ uint32_t shade_one(const PTState &st, const GenChunk &C, const PTScene &sc, uint32_t p,
                                              uint32_t bounce, bool &alive) {
    const uint32_t local = p - C.first;
    const uint32_t pix = local % C.npix;
    const int sample = (int)(C.sample_base + local / C.npix);
    const int4 meta = st.meta[p];
    Rng R = rng_for(sc, pix, sample, meta.z, (uint32_t)meta.w);
    const rg_ray r = st.ray[p];
    const float4 b = st.beta[p];
    Path P;
    P.o = mk(r.o[0], r.o[1], r.o[2]);
    P.d = mk(r.d[0], r.d[1], r.d[2]);
    P.beta = f4v(b);
    P.bsdf_pdf = b.w;
    P.tp = f4v(st.tp[p]);
    P.L = with_last_nee(st, p);
    P.passes = meta.x;
    P.medium = medium_of(meta.y);
    P.chroma = chroma_of(meta.y);
    P.bg = bg_of(meta.y);
    P.alive = true;
    Shadows S = shadows_of(st, p);
    const rg_hit h = st.hit[p];
    step<kExt>(sc, R, P, S, h);
    st.ray[p] = {{P.o.x, P.o.y, P.o.z}, 0.0f, {P.d.x, P.d.y, P.d.z}, MAX_RAY_DISTANCE};
    st.beta[p] = v4(P.beta, P.bsdf_pdf);
    st.tp[p] = v4(P.tp, 0.0f);
    st.L[p] = v4(P.L, P.alive ? 1.0f : 0.0f);
    st.meta[p] = make_int4(P.passes, pack_medium(P.medium, P.chroma, P.bg), R.dim, (int)R.k);
    st.nshadow[p] = (uint32_t)S.n;
    alive = P.alive && bounce + 1 < C.bounces;
    return (uint32_t)S.n;
}

// The event's NEE rays go to the shadow-ray queue (one ray per lane there),
// a live path to the next path queue.  Every thread of the block calls it.
__device__ __forceinline__ void queue_results(const PTState &st, uint32_t p, uint32_t ns, bool alive, int qi,
                                              uint32_t *lds) {
    for (uint32_t k = 0; k < kNeeSlots; k++) {
        const uint32_t pos = block_slot(k < ns, st.count + 2, lds);
        if (k < ns) st.ssrc[pos] = k * st.slot_stride + p;
    }
    const uint32_t pos = block_slot(alive, st.count + (qi ^ 1), lds);
    if (alive) st.q[qi ^ 1][pos] = p;
}

// Copies a record whose address is the same in every lane into scalar
// registers.  Called in wave-uniform control flow: loads from a uniform
// address inside a divergent branch are vector loads, and the copy keeps the
// shade body's reads of the record out of vector registers.
template <class T>
__device__ __forceinline__ void uniform_copy(T &dst, const T *src) {
    static_assert(sizeof(T) % 4 == 0, "record of whole words");
    const uint32_t *s = reinterpret_cast<const uint32_t *>(src);
    uint32_t *d = reinterpret_cast<uint32_t *>(&dst);
#pragma unroll
    for (uint32_t k = 0; k < sizeof(T) / 4; k++) d[k] = __builtin_amdgcn_readfirstlane(s[k]);
}

// One path event per live path (queue qi).  The lanes whose path is in the
// chunk of the wave's first lane run with that chunk and its job's scene
// record in scalar registers (the single-job shade kernel's register budget
// and occupancy); the other paths of the wave (waves that straddle two
// chunks) go to the mixed queue for k_bshade_mixed.
template <bool kExt>
__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_bshade(
    BatchArgs a, const DevJobSlot *__restrict__ jobs, int qi, uint32_t bounce) {
    __shared__ uint32_t lds[kPtBlock / 64 + 1];
    const PTState &st = a.st;
    const uint32_t i = blockIdx.x * kPtBlock + threadIdx.x;
    const uint32_t count = st.count[qi];
    if (blockIdx.x * kPtBlock >= count) return;  // whole block past the end
    uint32_t p = 0, ns = 0, c = 0xFFFFFFFFu;
    bool alive = false;
    if (i < count) {
        p = st.q[qi][i];
        c = a.chunk_of[p];
    }
    // the live lanes of a wave are a prefix of it: lane 0 is live when any is
    const uint32_t c0 = __builtin_amdgcn_readfirstlane(c);
    const bool mine = i < count && c == c0;
    wave_append(i < count && c != c0, p, a.mixed, st.count + 3);
    GenChunk C;
    PTScene sc;
    if (c0 != 0xFFFFFFFFu) {  // wave-uniform
        uniform_copy(C, &a.chunks[c0]);
        uniform_copy(sc, &jobs[C.job].j.sc);
    }
    if (mine) ns = shade_one<kExt>(st, C, sc, p, bounce, alive);
    queue_results(st, p, ns, alive, qi, lds);
}

// The mixed queue's path events, each with its own job's record (one
// distinct job at a time across the wave).  A fixed grid walks the queue in
// block-sized steps.
template <bool kExt>
__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_bshade_mixed(
    BatchArgs a, const DevJobSlot *__restrict__ jobs, int qi, uint32_t bounce) {
    __shared__ uint32_t lds[kPtBlock / 64 + 1];
    const PTState &st = a.st;
    const uint32_t count = st.count[3];
    for (uint32_t base = blockIdx.x * kPtBlock; base < count; base += gridDim.x * kPtBlock) {
        const uint32_t i = base + threadIdx.x;
        uint32_t p = 0, ns = 0;
        bool alive = false;
        if (i < count) {
            p = a.mixed[i];
            const GenChunk &C = a.chunks[a.chunk_of[p]];
            with_job(jobs, C.job, [&](const DevJob &J) { ns = shade_one<kExt>(st, C, J.sc, p, bounce, alive); });
        }
        queue_results(st, p, ns, alive, qi, lds);
    }
}

// Any-hit test of every queued NEE ray on the generation's tree.
struct BatchShadowSrc {
    const rg_ray *sray;
    const uint32_t *ssrc;
    uint8_t *socc;
    __device__ __forceinline__ bool load(uint64_t i, const DevBvh &, PsRay &r) const {
        const rg_ray x = sray[ssrc[i]];
        r = {x.o[0], x.o[1], x.o[2], x.d[0], x.d[1], x.d[2], 0.0f, x.tmax, 1u << (int)x.tmin};
        return true;
    }
    __device__ __forceinline__ void store_closest(uint64_t, uint32_t, float, float, float) const {}
    __device__ __forceinline__ void store_any(uint64_t i, bool o) const { socc[ssrc[i]] = o ? 1 : 0; }
};

template <int kStack, int kW>
__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(7))) void k_banyhit(DevBvh bv, BatchArgs a) {
    RG_LDS_COLUMN(lc, kPtBlock);
    const PTState &st = a.st;
    const uint32_t n = st.count[2];
    if (blockIdx.x == 0 && threadIdx.x == 0 && n) atomicAdd(st.stats + 1, (unsigned long long)n);
    trace_ps_loop<kStack, kW, true, kPtBlock>(bv, BatchShadowSrc{st.sray, st.ssrc, st.socc}, n, lc);
}

// Per-pixel sums of each chunk's S samples in sample order (a sample's value:
// its path's radiance plus its last event's unoccluded NEE contributions,
// clamped by the job's firefly clamp), added to its job's sums chunk after
// chunk (the single-job path's accumulation of one generation, the same
// float operations), then, for a job that finishes,
// its HDR pixel (rayrenderer.cpp:990-1009): mean * exposure, clamped at 0.
__global__ __launch_bounds__(kPtBlock) void k_baccumulate(BatchArgs a, const DevJobSlot *__restrict__ jobs,
                                                           const AccGroup *__restrict__ groups, uint32_t ngroups,
                                                           uint32_t total_pix) {
    const uint32_t i = blockIdx.x * kPtBlock + threadIdx.x;
    if (i >= total_pix) return;
    uint32_t lo = 0, hi = ngroups;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (groups[mid].pix_first <= i) lo = mid;
        else hi = mid;
    }
    const AccGroup g = groups[lo];
    const uint32_t pix = i - g.pix_first;
    const DevJob &J = jobs[a.chunks[g.c0].job].j;
    float ax = J.sum[3 * pix], ay = J.sum[3 * pix + 1], az = J.sum[3 * pix + 2];
    float bx = J.sumsq[3 * pix], by = J.sumsq[3 * pix + 1], bz = J.sumsq[3 * pix + 2];
    for (uint32_t c = g.c0; c < g.c1; c++) {
        const GenChunk &C = a.chunks[c];
        float sx = 0.0f, sy = 0.0f, sz = 0.0f, qx = 0.0f, qy = 0.0f, qz = 0.0f;
        for (uint32_t s = 0; s < C.S; s++) {
            const float4 v = sample_value(a.st, C.first + s * g.npix + pix, C.clamp, J.sc.back);
            sx = sx + v.x;
            sy = sy + v.y;
            sz = sz + v.z;
            qx = qx + v.x * v.x;
            qy = qy + v.y * v.y;
            qz = qz + v.z * v.z;
        }
        ax = ax + sx;
        ay = ay + sy;
        az = az + sz;
        bx = bx + qx;
        by = by + qy;
        bz = bz + qz;
    }
    J.sum[3 * pix] = ax;
    J.sum[3 * pix + 1] = ay;
    J.sum[3 * pix + 2] = az;
    J.sumsq[3 * pix] = bx;
    J.sumsq[3 * pix + 1] = by;
    J.sumsq[3 * pix + 2] = bz;
    if (g.finish) {
        const float inv = J.spp > 0 ? 1.0f / (float)J.spp : 0.0f;
        J.hdr[4 * pix] = fmaxf(ax * inv * J.exposure, 0.0f);
        J.hdr[4 * pix + 1] = fmaxf(ay * inv * J.exposure, 0.0f);
        J.hdr[4 * pix + 2] = fmaxf(az * inv * J.exposure, 0.0f);
        J.hdr[4 * pix + 3] = 1.0f;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int chk_s(hipError_t e, const char *what) {
    if (e == hipSuccess) return RG_OK;
    return set_error(RG_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

int stream_cu_count(int device) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return 0;
    return p.multiProcessorCount;
}

bool stream_stack_ok(uint32_t stack, uint32_t need) {
    return (stack == 32 || stack == 64 || stack == 128 || stack == 256) && need <= stack;
}

int stream_launch_camera(const BatchArgs &a, const void *d_jobs, uint32_t total, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_bset_u32, dim3(1), dim3(1), 0, s, a.st.count + 0, total);
    if (total)
        hipLaunchKernelGGL(k_bcamera, dim3((total + kPtBlock - 1) / kPtBlock), dim3(kPtBlock), 0, s, a,
                           (const DevJobSlot *)d_jobs, total);
    return chk_s(hipGetLastError(), "batch camera launch");
}

template <int kStack, int kW>
static void bounces_impl(const BatchArgs &a, const DevJobSlot *jobs, const DevBvh &bv, uint32_t paths,
                         uint32_t max_bounces, const StreamTune &tune, int n_cu, bool ext, hipStream_t s) {
    const uint32_t flat = (std::min(paths, a.cap) + kPtBlock - 1) / kPtBlock;
    // persistent traversal grids: what the device holds at once, or fewer
    // when blocks_per_cu caps it
    uint32_t ext_blocks = std::min<uint32_t>(flat, resident_blocks<k_bextend<kStack, kW>>(kPtBlock));
    uint32_t any_blocks = std::min<uint32_t>(2 * flat, resident_blocks<k_banyhit<kStack, kW>>(kPtBlock));
    if (tune.blocks_per_cu) {
        ext_blocks = std::min<uint32_t>(ext_blocks, (uint32_t)n_cu * tune.blocks_per_cu);
        any_blocks = std::min<uint32_t>(any_blocks, (uint32_t)n_cu * tune.blocks_per_cu);
    }
    // the mixed queue holds a few waves' worth of paths per chunk boundary
    const uint32_t mixed_blocks = std::min<uint32_t>(flat, 1024);
    int qi = 0;
    for (uint32_t b = 0; b < max_bounces; b++) {
        hipLaunchKernelGGL((k_bextend<kStack, kW>), dim3(ext_blocks), dim3(kPtBlock), 0, s, bv, a, qi);
        if (ext) {
            hipLaunchKernelGGL(k_bshade<true>, dim3(flat), dim3(kPtBlock), 0, s, a, jobs, qi, b);
            hipLaunchKernelGGL(k_bshade_mixed<true>, dim3(mixed_blocks), dim3(kPtBlock), 0, s, a, jobs, qi, b);
        } else {
            hipLaunchKernelGGL(k_bshade<false>, dim3(flat), dim3(kPtBlock), 0, s, a, jobs, qi, b);
            hipLaunchKernelGGL(k_bshade_mixed<false>, dim3(mixed_blocks), dim3(kPtBlock), 0, s, a, jobs, qi, b);
        }
        hipLaunchKernelGGL((k_banyhit<kStack, kW>), dim3(any_blocks), dim3(kPtBlock), 0, s, bv, a);
        qi ^= 1;
    }
}

int stream_launch_bounces(const BatchArgs &a, const void *d_jobs, const DevBvh &bv, uint32_t paths,
                          uint32_t max_bounces, const StreamTune &tune, int n_cu, bool ext, void *stream) {
    const DevJobSlot *jobs = (const DevJobSlot *)d_jobs;
    hipStream_t s = (hipStream_t)stream;
    const bool w8 = tune.width == 8;
    switch (tune.stack) {
        case 32: w8 ? bounces_impl<32, 8>(a, jobs, bv, paths, max_bounces, tune, n_cu, ext, s) : bounces_impl<32, 4>(a, jobs, bv, paths, max_bounces, tune, n_cu, ext, s); break;
        case 64: w8 ? bounces_impl<64, 8>(a, jobs, bv, paths, max_bounces, tune, n_cu, ext, s) : bounces_impl<64, 4>(a, jobs, bv, paths, max_bounces, tune, n_cu, ext, s); break;
        case 128: w8 ? bounces_impl<128, 8>(a, jobs, bv, paths, max_bounces, tune, n_cu, ext, s) : bounces_impl<128, 4>(a, jobs, bv, paths, max_bounces, tune, n_cu, ext, s); break;
        case 256: w8 ? bounces_impl<256, 8>(a, jobs, bv, paths, max_bounces, tune, n_cu, ext, s) : bounces_impl<256, 4>(a, jobs, bv, paths, max_bounces, tune, n_cu, ext, s); break;
        default: return set_error(RG_E_INVALID, "stack must be 32, 64, 128 or 256");
    }
    return chk_s(hipGetLastError(), "batch bounce launch");
}

int stream_launch_accumulate(const BatchArgs &a, const void *d_jobs, const AccGroup *d_groups, uint32_t ngroups,
                             uint32_t total_pix, void *stream) {
    if (!ngroups || !total_pix) return RG_OK;
    hipLaunchKernelGGL(k_baccumulate, dim3((total_pix + kPtBlock - 1) / kPtBlock), dim3(kPtBlock), 0,
                       (hipStream_t)stream, a, (const DevJobSlot *)d_jobs, d_groups, ngroups, total_pix);
    return chk_s(hipGetLastError(), "batch accumulate launch");
}

}  // namespace rg

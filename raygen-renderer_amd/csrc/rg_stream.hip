// rg_stream.hip — kernels of the streaming (batched) wavefront engine.
//
// Each iteration at parity qi runs five kernels over one pool of paths
// shared by every admitted job.  A path keeps its pool position for life;
// kernels touch element i in thread i.
//   regen   puts the next samples of the issue sequence (camera rays,
//           rayrenderer.cpp:860-934) into positions freed last iteration
//   extend  closest-hit traversal of every live position (persistent waves,
//           leaf postponement and lane refill; each ray on its job's tree)
//   shade   one path event per live position (the same step() as the
//           single-job path, with the job's scene record loaded
//           wave-uniform); the event's 0-2 NEE rays are queued with
//           block-aggregated appends (one atomic per block)
//   anyhit  the queued NEE rays, one per lane, persistent waves
//   resolve adds the unoccluded contributions in event order; a finished
//           path writes its clamped radiance to the job's sample buffer and
//           frees its position
// All bookkeeping lives in device counters, so a run of iterations needs no
// host round trip and is captured once as a hipGraph.
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
    DevBvh bv;
    float *samples;  // [spp][npix][3] clamped radiance
    float *sum;      // [npix][3]
    float *sumsq;    // [npix][3]
    float *hdr;      // [npix][4]
    uint32_t npix, spp, group, max_bounces;
    float clamp, exposure;
    uint32_t tile, pad;
};
struct DevJobSlot {
    DevJob j;
    uint8_t pad[kDevJobBytes - sizeof(DevJob)];
};
static_assert(sizeof(DevJobSlot) == kDevJobBytes, "job record size");

DevBvh dev_bvh(const Scene &s);

void stream_job_record(const JobDesc &d, void *out) {
    DevJobSlot rec;
    std::memset(&rec, 0, sizeof rec);
    DevJob &J = rec.j;
    J.sc = make_ptscene(*d.scene, d.st);
    J.bv = dev_bvh(*d.scene);
    J.samples = d.d_samples;
    J.sum = d.d_sum;
    J.sumsq = d.d_sumsq;
    J.hdr = d.d_hdr;
    J.npix = d.st.width * d.st.height;
    J.spp = d.st.spp;
    J.group = d.group;
    J.max_bounces = d.st.max_depth + 2;
    J.clamp = d.st.firefly_clamp;
    J.exposure = d.scene->camera.exposure;
    J.tile = d.tile;
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

// Issue order of the pixels of one pass: bands of T rows, each walked in
// T-column chunks (partial bands and chunks at the image edge), so the 64
// lanes of a wave cover a compact block of the image.
__device__ __forceinline__ uint32_t tiled_pixel(uint32_t k, uint32_t W, uint32_t H, uint32_t T) {
    if (T <= 1) return k;
    const uint32_t band = k / (W * T);
    const uint32_t y0 = band * T;
    const uint32_t hb = min(T, H - y0);
    const uint32_t idx = k - y0 * W;
    const uint32_t c = idx / (hb * T);
    const uint32_t x0 = c * T;
    const uint32_t wc = min(T, W - x0);
    const uint32_t r = idx - c * hb * T;
    return (y0 + r / wc) * W + x0 + r % wc;
}

// Last issue-table entry whose begin <= g.
__device__ __forceinline__ uint32_t find_issue(const IssueEntry *__restrict__ t, uint32_t n, uint64_t g) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (t[mid].begin <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void k_sreset(StreamPool sp) {
    const uint32_t i = blockIdx.x * kPtBlock + threadIdx.x;
    if (i < sp.cap) {
        sp.st.sid[i] = make_int4(-1, 0, 0, 0);
        sp.holes[i] = sp.cap - 1 - i;
    }
    if (i == 0) {
        StreamCounters *C = sp.ctr;
        C->cursor[0] = C->cursor[1] = 0;
        C->seq_end = 0;
        C->nholes[0] = sp.cap;
        C->nholes[1] = 0;
        C->nsh[0] = C->nsh[1] = 0;
        C->fetch[0] = C->fetch[1] = 0;
        C->n_issue = 0;
        for (int k = 0; k < 4; k++) C->stats[k] = 0;
    }
}

// Puts the next samples of the issue sequence into free positions.
__global__ __launch_bounds__(kPtBlock) void k_sregen(StreamPool sp, const DevJobSlot *__restrict__ jobs, int qi) {
    StreamCounters *C = sp.ctr;
    const uint32_t h = C->nholes[qi];
    const uint64_t cur = C->cursor[qi];
    const uint64_t end = C->seq_end;
    const uint32_t n_issue = C->n_issue;
    // nothing is issued without an issue table covering [cur, end)
    const uint64_t left = (end > cur && n_issue > 0) ? end - cur : 0;
    const uint32_t issue = (uint32_t)(left < (uint64_t)h ? left : (uint64_t)h);
    const uint32_t gid = blockIdx.x * kPtBlock + threadIdx.x;
    if (gid == 0) {
        C->cursor[qi ^ 1] = cur + issue;
        C->nholes[qi ^ 1] = h - issue;
        C->nsh[qi] = 0;
        C->fetch[qi] = 0;
        C->stats[0] += sp.cap - h + issue;
        C->stats[2] += issue;
        C->stats[3] += 1;
    }
    const PathSoA S = sp.st;
    const uint32_t stride = gridDim.x * kPtBlock;
    for (uint32_t k = gid; k < issue; k += stride) {
        const uint64_t g = cur + k;
        const IssueEntry e = sp.issue[find_issue(sp.issue, n_issue, g)];
        const uint32_t i = sp.holes[h - 1 - k];
        const uint32_t local = (uint32_t)(g - e.begin);
        with_job(jobs, e.job, [&](const DevJob &J) {
            const uint32_t pass = local / J.npix;
            const uint32_t pix = tiled_pixel(local - pass * J.npix, J.sc.width, J.sc.height, J.tile);
            Rng R = rng_for(J.sc, pix, (int)pass, 0, 0);
            v3 o, d;
            camera_ray(J.sc, R, pix % J.sc.width, pix / J.sc.width, o, d);
            S.ray[i] = {{o.x, o.y, o.z}, 0.0f, {d.x, d.y, d.z}, MAX_RAY_DISTANCE};
            S.beta[i] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
            S.tp[i] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
            S.L[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            S.meta[i] = make_int4(-1, J.sc.global_medium, R.dim, (int)R.k);
            S.sid[i] = make_int4((int)e.job, (int)pass, (int)pix, 0);
        });
    }
}

// Closest hit of every live position; each ray traverses its own job's tree.
struct PoolExtendSrc {
    const int4 *sid;
    const rg_ray *ray;
    rg_hit *hit;
    const DevJobSlot *jobs;
    __device__ __forceinline__ bool load(uint64_t i, DevBvh &bv, PsRay &r) const {
        const int job = sid[i].x;
        if (job < 0) return false;
        bv = jobs[job].j.bv;
        const rg_ray x = ray[i];
        r = {x.o[0], x.o[1], x.o[2], x.d[0], x.d[1], x.d[2], x.tmin, x.tmax, 0u};
        return true;
    }
    __device__ __forceinline__ void store_closest(uint64_t i, uint32_t prim, float t, float u, float v) const {
        hit[i] = {prim, t, u, v};
    }
    __device__ __forceinline__ void store_any(uint64_t, bool) const {}
};

template <int kStack, int kW>
__global__ __launch_bounds__(kPtBlock) void k_sextend(StreamPool sp, const DevJobSlot *__restrict__ jobs) {
    RG_LDS_COLUMN(lc, kPtBlock);
    const DevBvh none{};
    trace_ps_loop<kStack, kW, false, kPtBlock, PoolExtendSrc, true>(none, PoolExtendSrc{sp.st.sid, sp.st.ray, sp.hit, jobs},
                                                                  sp.cap, lc);
}

// One path event per live position, state updated in place; the event's
// NEE rays are queued for the any-hit kernel.
__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_sshade(StreamPool sp, const DevJobSlot *__restrict__ jobs, int qi) {
    __shared__ uint32_t lds[kPtBlock / 64 + 1];
    const uint32_t i = blockIdx.x * kPtBlock + threadIdx.x;
    const PathSoA S = sp.st;
    uint32_t ns = 0;
    int4 sid = make_int4(-1, 0, 0, 0);
    if (i < sp.cap) sid = S.sid[i];
    if (sid.x >= 0) {
        const int4 meta = S.meta[i];
        with_job(jobs, (uint32_t)sid.x, [&](const DevJob &J) {
            const PTScene &sc = J.sc;
            Rng R = rng_for(sc, (uint32_t)sid.z, sid.y, meta.z, (uint32_t)meta.w);
            const rg_ray r = S.ray[i];
            const float4 b = S.beta[i];
            Path P;
            P.o = mk(r.o[0], r.o[1], r.o[2]);
            P.d = mk(r.d[0], r.d[1], r.d[2]);
            P.beta = f4v(b);
            P.bsdf_pdf = b.w;
            P.tp = f4v(S.tp[i]);
            P.L = f4v(S.L[i]);
            P.passes = meta.x;
            P.medium = meta.y;
            P.alive = true;
            Shadows Sh;
            Sh.n = 0;
            const rg_hit h = sp.hit[i];
            step(sc, R, P, Sh, h);
            const bool survive = P.alive && (uint32_t)(sid.w + 1) < J.max_bounces;
            S.ray[i] = {{P.o.x, P.o.y, P.o.z}, 0.0f, {P.d.x, P.d.y, P.d.z}, MAX_RAY_DISTANCE};
            S.beta[i] = v4(P.beta, P.bsdf_pdf);
            S.tp[i] = v4(P.tp, 0.0f);
            S.L[i] = v4(P.L, survive ? 1.0f : 0.0f);
            S.meta[i] = make_int4(P.passes, P.medium, R.dim, (int)R.k);
            store_shadows(Sh, i, sp.sray, sp.scontrib);
            ns = (uint32_t)Sh.n;
        });
        sid.w += 1;
        S.sid[i] = sid;
        sp.nshadow[i] = ns;
    }
    for (uint32_t k = 0; k < 2; k++) {
        const uint32_t pos = block_slot(k < ns, &sp.ctr->nsh[qi], lds);
        if (k < ns) sp.ssrc[pos] = 2 * i + k;
    }
}

// Any-hit test of every queued NEE ray, persistent waves; each ray traverses
// its own job's tree.
struct PoolShadowSrc {
    const uint32_t *ssrc;
    const int4 *sid;
    const rg_ray *sray;
    uint8_t *socc;
    const DevJobSlot *jobs;
    __device__ __forceinline__ bool load(uint64_t i, DevBvh &bv, PsRay &r) const {
        const uint32_t src = ssrc[i];
        bv = jobs[sid[src >> 1].x].j.bv;
        const rg_ray x = sray[src];
        r = {x.o[0], x.o[1], x.o[2], x.d[0], x.d[1], x.d[2], 0.0f, x.tmax, 1u << (int)x.tmin};
        return true;
    }
    __device__ __forceinline__ void store_closest(uint64_t, uint32_t, float, float, float) const {}
    __device__ __forceinline__ void store_any(uint64_t i, bool o) const { socc[ssrc[i]] = o ? 1 : 0; }
};

template <int kStack, int kW>
__global__ __launch_bounds__(kPtBlock) void k_sanyhit(StreamPool sp, const DevJobSlot *__restrict__ jobs, int qi) {
    RG_LDS_COLUMN(lc, kPtBlock);
    const uint32_t n = sp.ctr->nsh[qi];
    if (blockIdx.x == 0 && threadIdx.x == 0 && n) atomicAdd(&sp.ctr->stats[1], (unsigned long long)n);
    const DevBvh none{};
    trace_ps_loop<kStack, kW, true, kPtBlock, PoolShadowSrc, true>(
        none, PoolShadowSrc{sp.ssrc, sp.st.sid, sp.sray, sp.socc, jobs}, n, lc);
}

// Adds the unoccluded NEE contributions in event order.  A path that ended
// writes its clamped radiance to the job's sample buffer and frees its
// position (one hole-list atomic per block).  Job completion needs no
// counter: a path lives at most max_depth + 2 iterations, so a job is done
// that many iterations after the issue cursor passed its last sample.
__global__ __launch_bounds__(kPtBlock) void k_sresolve(StreamPool sp, const DevJobSlot *__restrict__ jobs, int qi) {
    __shared__ uint32_t lds[kPtBlock / 64 + 1];
    const uint32_t i = blockIdx.x * kPtBlock + threadIdx.x;
    const PathSoA S = sp.st;
    bool fin = false;
    int4 sid = make_int4(-1, 0, 0, 0);
    if (i < sp.cap) sid = S.sid[i];
    if (sid.x >= 0) {
        float4 L = S.L[i];
        const uint32_t ns = sp.nshadow[i];
        for (uint32_t k = 0; k < ns; k++) {
            if (!sp.socc[2 * i + k]) {
                const float4 c = sp.scontrib[2 * i + k];
                L.x = L.x + c.x;
                L.y = L.y + c.y;
                L.z = L.z + c.z;
            }
        }
        if (L.w != 0.0f) {
            S.L[i] = L;
        } else {
            const DevJob &J = jobs[sid.x].j;
            fin = true;
            float cx = fmaxf(L.x, 0.0f), cy = fmaxf(L.y, 0.0f), cz = fmaxf(L.z, 0.0f);
            const float clamp_max = J.clamp;
            if (clamp_max > 0.0f) {
                cx = fminf(cx, clamp_max);
                cy = fminf(cy, clamp_max);
                cz = fminf(cz, clamp_max);
            }
            float *o = J.samples + 3 * ((size_t)(uint32_t)sid.y * J.npix + (uint32_t)sid.z);
            o[0] = cx;
            o[1] = cy;
            o[2] = cz;
            S.sid[i] = make_int4(-1, 0, 0, 0);
        }
    }
    const uint32_t pos = block_slot(fin, &sp.ctr->nholes[qi ^ 1], lds);
    if (fin) sp.holes[pos] = i;
}


// Per-pixel sums of a finished job in pass order, in groups of `group`
// passes (the single-job path's samples in flight), then the HDR resolve
// (mean * exposure, clamped at 0).
__global__ __launch_bounds__(kPtBlock) void k_sfinish(const DevJobSlot *__restrict__ jobs, uint32_t slot) {
    const DevJob &J = jobs[slot].j;
    const uint32_t npix = J.npix;
    const uint32_t pix = blockIdx.x * kPtBlock + threadIdx.x;
    if (pix >= npix) return;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, q0 = 0.0f, q1 = 0.0f, q2 = 0.0f;
    for (uint32_t p0 = 0; p0 < J.spp; p0 += J.group) {
        const uint32_t p1 = min(J.spp, p0 + J.group);
        float sx = 0.0f, sy = 0.0f, sz = 0.0f, qx = 0.0f, qy = 0.0f, qz = 0.0f;
        for (uint32_t p = p0; p < p1; p++) {
            const float *c = J.samples + 3 * ((size_t)p * npix + pix);
            const float cx = c[0], cy = c[1], cz = c[2];
            sx = sx + cx;
            sy = sy + cy;
            sz = sz + cz;
            qx = qx + cx * cx;
            qy = qy + cy * cy;
            qz = qz + cz * cz;
        }
        s0 += sx;
        s1 += sy;
        s2 += sz;
        q0 += qx;
        q1 += qy;
        q2 += qz;
    }
    J.sum[3 * pix] = s0;
    J.sum[3 * pix + 1] = s1;
    J.sum[3 * pix + 2] = s2;
    J.sumsq[3 * pix] = q0;
    J.sumsq[3 * pix + 1] = q1;
    J.sumsq[3 * pix + 2] = q2;
    const float inv = 1.0f / (float)J.spp;
    J.hdr[4 * pix] = fmaxf(s0 * inv * J.exposure, 0.0f);
    J.hdr[4 * pix + 1] = fmaxf(s1 * inv * J.exposure, 0.0f);
    J.hdr[4 * pix + 2] = fmaxf(s2 * inv * J.exposure, 0.0f);
    J.hdr[4 * pix + 3] = 1.0f;
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

int stream_launch_reset(const StreamPool &sp, void *stream) {
    const uint32_t blocks = (sp.cap + kPtBlock - 1) / kPtBlock;
    hipLaunchKernelGGL(k_sreset, dim3(blocks ? blocks : 1), dim3(kPtBlock), 0, (hipStream_t)stream, sp);
    return chk_s(hipGetLastError(), "stream reset launch");
}

template <int kStack, int kW>
static void iteration_impl(const StreamPool &sp, const DevJobSlot *jobs, int qi, const StreamTune &tune, int n_cu,
                           hipStream_t s) {
    const uint32_t flat = (sp.cap + kPtBlock - 1) / kPtBlock;
    const uint32_t regen_blocks = std::min<uint32_t>(flat, (uint32_t)n_cu * 4);
    // persistent traversal grids: what the device holds at once, or fewer
    // when blocks_per_cu caps it
    uint32_t ext_blocks = std::min<uint32_t>(flat, resident_blocks<k_sextend<kStack, kW>>(kPtBlock));
    uint32_t any_blocks = std::min<uint32_t>(2 * flat, resident_blocks<k_sanyhit<kStack, kW>>(kPtBlock));
    if (tune.blocks_per_cu) {
        ext_blocks = std::min<uint32_t>(ext_blocks, (uint32_t)n_cu * tune.blocks_per_cu);
        any_blocks = std::min<uint32_t>(any_blocks, (uint32_t)n_cu * tune.blocks_per_cu);
    }
    hipLaunchKernelGGL(k_sregen, dim3(regen_blocks), dim3(kPtBlock), 0, s, sp, jobs, qi);
    hipLaunchKernelGGL((k_sextend<kStack, kW>), dim3(ext_blocks), dim3(kPtBlock), 0, s, sp, jobs);
    hipLaunchKernelGGL(k_sshade, dim3(flat), dim3(kPtBlock), 0, s, sp, jobs, qi);
    hipLaunchKernelGGL((k_sanyhit<kStack, kW>), dim3(any_blocks), dim3(kPtBlock), 0, s, sp, jobs, qi);
    hipLaunchKernelGGL(k_sresolve, dim3(flat), dim3(kPtBlock), 0, s, sp, jobs, qi);
}

int stream_launch_iteration(const StreamPool &sp, const void *d_jobs, int qi, const StreamTune &tune, int n_cu,
                            void *stream) {
    const DevJobSlot *jobs = (const DevJobSlot *)d_jobs;
    hipStream_t s = (hipStream_t)stream;
    const bool w8 = tune.width == 8;
    switch (tune.stack) {
        case 32: w8 ? iteration_impl<32, 8>(sp, jobs, qi, tune, n_cu, s) : iteration_impl<32, 4>(sp, jobs, qi, tune, n_cu, s); break;
        case 64: w8 ? iteration_impl<64, 8>(sp, jobs, qi, tune, n_cu, s) : iteration_impl<64, 4>(sp, jobs, qi, tune, n_cu, s); break;
        case 128: w8 ? iteration_impl<128, 8>(sp, jobs, qi, tune, n_cu, s) : iteration_impl<128, 4>(sp, jobs, qi, tune, n_cu, s); break;
        case 256: w8 ? iteration_impl<256, 8>(sp, jobs, qi, tune, n_cu, s) : iteration_impl<256, 4>(sp, jobs, qi, tune, n_cu, s); break;
        default: return set_error(RG_E_INVALID, "stack must be 32, 64, 128 or 256");
    }
    return chk_s(hipGetLastError(), "stream iteration launch");
}

int stream_launch_finish(const void *d_jobs, uint32_t slot, uint32_t npix, void *stream) {
    hipLaunchKernelGGL(k_sfinish, dim3((npix + kPtBlock - 1) / kPtBlock), dim3(kPtBlock), 0, (hipStream_t)stream,
                       (const DevJobSlot *)d_jobs, slot);
    return chk_s(hipGetLastError(), "stream finish launch");
}

}  // namespace rg

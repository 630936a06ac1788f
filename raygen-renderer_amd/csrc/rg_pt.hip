// rg_pt.hip — wavefront BSDF path tracer (shading system 5) for gfx950.
//
// One path sample per thread slot.  Each bounce runs three kernels over a
// compacted queue of live paths:
//   extend  closest-hit traversal of every live path's ray
//   shade   one path event: medium free flight, emission + MIS, back-face
//           rules, depth cap, Russian roulette, lobe choice, NEE shadow-ray
//           set-up and the next ray (RayBSDFShaderProvider::shade,
//           rayrenderer.cpp:2776-2896; RayRenderer::tracePath, :1596-1864;
//           bsdf.cpp:54-610); it first adds the previous event's unoccluded
//           NEE contributions (event order), and appends the live paths to
//           the next queue with ballot/popcount compaction (one atomic per
//           block)
//   anyhit  any-hit tests of the queued NEE rays (area light, envmap), one
//           ray per lane: shade appends each event's 0-2 rays to a queue
// followed, after the last bounce, by a per-pixel accumulate kernel that
// adds each finished path's last NEE contributions, clamps and sums the
// samples in a fixed order.
//
// The Mix shader's lobes are chosen stochastically (probability proportional
// to the lobe weights) instead of being evaluated recursively; the estimator
// has the same expectation.  Random draws are consumed in the same order as
// the CPU single-lobe restatement, so results agree with it per path up to
// the last-ulp differences of transcendental functions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "rg_pt_dev.h"

namespace rg {

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
__global__ void k_set_u32(uint32_t *p, uint32_t v) { *p = v; }
// A generation's first kernel: the first queue's length, zero work counters.
__global__ void k_begin_counts(uint32_t *count, uint32_t n0) {
    count[0] = n0;
    count[kExtendWork] = 0;
    count[kShadowWork] = 0;
}

// Camera rays (rayrenderer.cpp:860-934) for the paths of samples
// [sample_base, sample_base + S) of every pixel in an active tile.
__global__ __launch_bounds__(kPtBlock) void k_camera(PTScene sc, PTState st, uint32_t S, uint32_t sample_base,
                                                      const uint8_t *__restrict__ tile_mask) {
    const uint32_t i = blockIdx.x * kPtBlock + threadIdx.x;
    const uint32_t npix = sc.width * sc.height;
    bool active = i < npix * S;
    uint32_t pix = 0;
    if (active) {
        pix = i % npix;
        const uint32_t s = i / npix;
        const uint32_t x = pix % sc.width, y = pix / sc.width;
        if (tile_mask && !tile_mask[(y / 32) * ((sc.width + 31) / 32) + x / 32]) active = false;
        if (active) {
            Rng R = rng_for(sc, pix, (int)(sample_base + s), 0, 0);
            v3 o, d;
            camera_ray(sc, R, x, y, o, d);
            st.ray[i] = {{o.x, o.y, o.z}, 0.0f, {d.x, d.y, d.z}, MAX_RAY_DISTANCE};
            st.beta[i] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
            st.tp[i] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
            st.L[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            st.meta[i] = make_int4(-1, pack_medium(sc.global_medium, -1, false), R.dim, (int)R.k);
            st.nshadow[i] = 0;
        }
    }
    if (!tile_mask) {
        // every sample is live: the queue is the identity and its length
        // (npix * S) was set by the host
        if (active) st.q[0][i] = i;
        if (i == 0) atomicAdd(st.stats + 2, (unsigned long long)npix * S);
        return;
    }
    // live samples of the active tiles: one queue atomic and one counter
    // atomic per block (same-address atomics from every wave of a
    // 1080p pass cost more than the rays)
    __shared__ uint32_t lds[kPtBlock / 64 + 1];
    const uint32_t pos = block_slot(active, st.count + 0, lds);
    if (active) st.q[0][pos] = i;
    const int live = __syncthreads_count(active ? 1 : 0);
    if (threadIdx.x == 0 && live) atomicAdd(st.stats + 2, (unsigned long long)live);
}

// Closest hit of every live path's ray (queue qi), persistent waves.
struct ExtendSrc {
    const rg_ray *ray;
    const uint32_t *q;
    rg_hit *hit;
    uint32_t *work;
    static constexpr uint32_t kStaticQuarters = 1, kGrab = 128;
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

// Block 0 also clears the shadow-ray queue and the next path queue, which
// nothing reads while this kernel runs and the shade kernel after it fills,
// and the shadow-ray traversal's work counter.
// Both traversal kernels run at 7 waves per SIMD (72 VGPRs), as the
// closest-hit kernel: measured faster than the compiler's 8-wave allocation.
template <int kStack, int kW>
__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(7))) void k_extend(DevBvh bv, PTState st, int qi) {
    RG_LDS_COLUMN(lc, kPtBlock);
    const uint32_t count = st.count[qi];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st.count[2] = 0;
        st.count[qi ^ 1] = 0;
        st.count[kShadowWork] = 0;
        if (count) atomicAdd(st.stats + 0, (unsigned long long)count);
    }
    trace_ps_loop<kStack, kW, false, kPtBlock>(bv, ExtendSrc{st.ray, st.q[qi], st.hit, st.count + kExtendWork},
                                               count, lc);
}

template <bool kExt>
__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_shade(PTScene sc, PTState st,
                                                                                           int qi) {
    __shared__ uint32_t lds[kPtBlock / 64 + 1];
    const uint32_t i = blockIdx.x * kPtBlock + threadIdx.x;
    const uint32_t count = st.count[qi];
    if (blockIdx.x * kPtBlock >= count) return;  // whole block past the end
    uint32_t p = 0, ns = 0;
    bool alive = false;
    if (i < count) {
        p = st.q[qi][i];
        const uint32_t npix = sc.width * sc.height;
        const uint32_t pix = p % npix;
        const int sample = (int)(st.sample_base + p / npix);
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
        ns = (uint32_t)S.n;
        st.nshadow[p] = ns;
        alive = P.alive;
    }
    // the event's NEE rays go to the shadow-ray queue, one ray per lane there
    // the point-light and volume slots only when the scene has such lights
    const uint32_t slots = 2 + (sc.nplights ? 1u : 0u) + (sc.nvlights ? 1u : 0u);
    for (uint32_t k = 0; k < slots; k++) {
        const uint32_t pos = block_slot(k < ns, st.count + 2, lds);
        if (k < ns) st.ssrc[pos] = k * st.slot_stride + p;
    }
    // live paths go to the next queue; a finished path keeps its radiance and
    // its last event's NEE rays, which the accumulation adds up
    const uint32_t pos = block_slot(alive, st.count + (qi ^ 1), lds);
    if (alive) st.q[qi ^ 1][pos] = p;
}

// Any-hit test of every queued shadow ray (TriangleBVH::intersectAny,
// bvh.cpp:287-339, with the ray's occluder predicate), persistent waves.
struct ShadowSrc {
    const rg_ray *sray;
    const uint32_t *ssrc;
    uint8_t *socc;
    uint32_t *work;
    static constexpr uint32_t kStaticQuarters = 1, kGrab = 128;
    __device__ __forceinline__ bool load(uint64_t i, const DevBvh &, PsRay &r) const {
        const rg_ray x = sray[ssrc[i]];
        r = {x.o[0], x.o[1], x.o[2], x.d[0], x.d[1], x.d[2], 0.0f, x.tmax, 1u << (int)x.tmin};
        return true;
    }
    __device__ __forceinline__ void store_closest(uint64_t, uint32_t, float, float, float) const {}
    __device__ __forceinline__ void store_any(uint64_t i, bool o) const { socc[ssrc[i]] = o ? 1 : 0; }
};

template <int kStack, int kW>
__global__ __launch_bounds__(kPtBlock) __attribute__((amdgpu_waves_per_eu(7))) void k_anyhit(DevBvh bv, PTState st) {
    RG_LDS_COLUMN(lc, kPtBlock);
    const uint32_t n = st.count[2];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st.count[kExtendWork] = 0;  // for the next extension
        if (n) atomicAdd(st.stats + 1, (unsigned long long)n);
    }
    trace_ps_loop<kStack, kW, true, kPtBlock>(bv, ShadowSrc{st.sray, st.ssrc, st.socc, st.count + kShadowWork}, n,
                                              lc);
}

// Per-pixel sum / sum of squares over the S samples of this generation,
// in sample order.  A sample's value is its path's radiance plus the
// unoccluded NEE contributions of its last event, clamped (>= 0, and at
// clamp_max when that is > 0); back = backColor for an eye ray that crossed
// the global medium without a hit and gathered nothing.
struct BackColor {
    float c[3];
};
// Sums the generation's samples [s0, s0 + S) of every pixel (one pass of
// an adaptive window when s0 > 0).
__global__ __launch_bounds__(kPtBlock) void k_accumulate(PTState st, uint32_t width, uint32_t npix, uint32_t s0,
                                                          uint32_t S, const uint8_t *__restrict__ tile_mask,
                                                          float clamp_max, BackColor back, float *__restrict__ sum,
                                                          float *__restrict__ sumsq) {
    const uint32_t pix = blockIdx.x * kPtBlock + threadIdx.x;
    if (pix >= npix) return;
    if (tile_mask) {
        const uint32_t x = pix % width, y = pix / width;
        if (!tile_mask[(y / 32) * ((width + 31) / 32) + x / 32]) return;
    }
    float sx = 0.0f, sy = 0.0f, sz = 0.0f, qx = 0.0f, qy = 0.0f, qz = 0.0f;
    for (uint32_t s = s0; s < s0 + S; s++) {
        const float4 c = sample_value(st, (uint32_t)((size_t)s * npix + pix), clamp_max, back.c);
        sx = sx + c.x;
        sy = sy + c.y;
        sz = sz + c.z;
        qx = qx + c.x * c.x;
        qy = qy + c.y * c.y;
        qz = qz + c.z * c.z;
    }
    sum[3 * pix] += sx;
    sum[3 * pix + 1] += sy;
    sum[3 * pix + 2] += sz;
    sumsq[3 * pix] += qx;
    sumsq[3 * pix + 1] += qy;
    sumsq[3 * pix + 2] += qz;
}

// HDR output: mean * exposure, clamped at 0 (rayrenderer.cpp:990-1009)
__global__ __launch_bounds__(kPtBlock) void k_resolve(const float *__restrict__ sum, const uint32_t *__restrict__ nspp,
                                                       uint32_t width, uint32_t npix, float exposure,
                                                       float *__restrict__ hdr) {
    const uint32_t pix = blockIdx.x * kPtBlock + threadIdx.x;
    if (pix >= npix) return;
    const uint32_t x = pix % width, y = pix / width;
    const uint32_t n = nspp[(y / 32) * ((width + 31) / 32) + x / 32];
    const float inv = (n > 0) ? 1.0f / (float)n : 0.0f;
    hdr[4 * pix] = fmaxf(sum[3 * pix] * inv * exposure, 0.0f);
    hdr[4 * pix + 1] = fmaxf(sum[3 * pix + 1] * inv * exposure, 0.0f);
    hdr[4 * pix + 2] = fmaxf(sum[3 * pix + 2] * inv * exposure, 0.0f);
    hdr[4 * pix + 3] = 1.0f;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * kPtBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

// Mean per-pixel relative standard error of the mean over a 32x32 tile
// (RayRenderer::computeTileNoise, rayrenderer.cpp:1038-1087), and the tile's
// selection for the next refinement pass (renderAdaptive :1237-1251).
__global__ __launch_bounds__(256) void k_tile_noise(const float *__restrict__ sum, const float *__restrict__ sumsq,
                                                     const uint32_t *__restrict__ nspp, uint32_t W, uint32_t H,
                                                     uint32_t target, float threshold, uint8_t *__restrict__ mask,
                                                     uint32_t *__restrict__ nactive) {
    const uint32_t tiles_x = (W + 31) / 32;
    const uint32_t tile = blockIdx.x;
    const uint32_t x0 = (tile % tiles_x) * 32, y0 = (tile / tiles_x) * 32;
    const uint32_t tw = min(32u, W - x0), th = min(32u, H - y0);
    const uint32_t n = nspp[tile];
    __shared__ double sacc[256];
    __shared__ int scnt[256];
    double acc = 0.0;
    int cnt = 0;
    if (n > 1) {
        const float invN = 1.0f / (float)n;
        for (uint32_t k = threadIdx.x; k < tw * th; k += 256) {
            const size_t pix = (size_t)(y0 + k / tw) * W + (x0 + k % tw);
            const float mr = sum[3 * pix] * invN, mg = sum[3 * pix + 1] * invN, mb = sum[3 * pix + 2] * invN;
            const float mx = fmaxf(fmaxf(mr, mg), mb);
            if (mx < 1e-3f) continue;
            const float vr = fmaxf(0.0f, sumsq[3 * pix] * invN - mr * mr);
            const float vg = fmaxf(0.0f, sumsq[3 * pix + 1] * invN - mg * mg);
            const float vb = fmaxf(0.0f, sumsq[3 * pix + 2] * invN - mb * mb);
            const float rr = sqrtf(vr * invN) / fmaxf(mr, 1e-3f);
            const float rg_ = sqrtf(vg * invN) / fmaxf(mg, 1e-3f);
            const float rb = sqrtf(vb * invN) / fmaxf(mb, 1e-3f);
            acc += (double)fmaxf(fmaxf(rr, rg_), rb);
            cnt++;
        }
    }
    sacc[threadIdx.x] = acc;
    scnt[threadIdx.x] = cnt;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sacc[threadIdx.x] += sacc[threadIdx.x + w];
            scnt[threadIdx.x] += scnt[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float noise;
        if (n <= 1) noise = 1e9f;
        else noise = scnt[0] == 0 ? 0.0f : (float)(sacc[0] / (double)scnt[0]);
        const bool act = n < target && noise > threshold;
        mask[tile] = act ? 1 : 0;
        if (act) atomicAdd(nactive, 1u);
    }
}

__global__ void k_tiles_add(uint32_t *nspp, const uint8_t *mask, uint32_t ntiles, uint32_t add) {
    const uint32_t i = blockIdx.x * kPtBlock + threadIdx.x;
    if (i < ntiles && mask[i]) nspp[i] += add;
}

// ---------------------------------------------------------------------------
// host-side launch sequence for one generation of S samples per pixel
// ---------------------------------------------------------------------------
static int chk(hipError_t e, const char *what) {
    if (e == hipSuccess) return RG_OK;
    return set_error(RG_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

DevBvh dev_bvh(const Scene &s);

template <int kStack, int kW>
static int generation_impl(const Scene &scn, const PTScene &sc, PTState st, uint32_t S, uint32_t sample_base,
                           const uint8_t *d_tile_mask, float *d_sum, float *d_sumsq, float clamp_max,
                           bool accumulate, hipStream_t stream) {
    const uint32_t npix = sc.width * sc.height;
    const uint32_t cap = npix * S;
    const uint32_t blocks = (cap + kPtBlock - 1) / kPtBlock;
    const DevBvh bv = dev_bvh(scn);
    // persistent traversal grids: what the device holds at once, at most one
    // thread per possible ray
    const uint32_t ext_blocks = std::min<uint32_t>(blocks, resident_blocks<k_extend<kStack, kW>>(kPtBlock));
    const uint32_t any_blocks = std::min<uint32_t>(kNeeSlots * blocks, resident_blocks<k_anyhit<kStack, kW>>(kPtBlock));
    st.sample_base = sample_base;
    hipLaunchKernelGGL(k_begin_counts, dim3(1), dim3(1), 0, stream, st.count, d_tile_mask ? 0u : cap);
    hipLaunchKernelGGL(k_camera, dim3(blocks), dim3(kPtBlock), 0, stream, sc, st, S, sample_base, d_tile_mask);
    int qi = 0;
    const int bounces = sc.events;
    for (int b = 0; b < bounces; b++) {
        hipLaunchKernelGGL((k_extend<kStack, kW>), dim3(ext_blocks), dim3(kPtBlock), 0, stream, bv, st, qi);
        if (sc.ext) hipLaunchKernelGGL(k_shade<true>, dim3(blocks), dim3(kPtBlock), 0, stream, sc, st, qi);
        else hipLaunchKernelGGL(k_shade<false>, dim3(blocks), dim3(kPtBlock), 0, stream, sc, st, qi);
        hipLaunchKernelGGL((k_anyhit<kStack, kW>), dim3(any_blocks), dim3(kPtBlock), 0, stream, bv, st);
        qi ^= 1;
    }
    if (accumulate)
        hipLaunchKernelGGL(k_accumulate, dim3((npix + kPtBlock - 1) / kPtBlock), dim3(kPtBlock), 0, stream, st,
                           sc.width, npix, 0u, S, d_tile_mask, clamp_max,
                           BackColor{{sc.back[0], sc.back[1], sc.back[2]}}, d_sum, d_sumsq);
    return chk(hipGetLastError(), "path-tracer launch");
}

int launch_generation(const Scene &scn, const PTScene &sc, const PTState &st, uint32_t S, uint32_t sample_base,
                      const uint8_t *d_tile_mask, float *d_sum, float *d_sumsq, float clamp_max, void *stream,
                      bool accumulate) {
    if (sc.shader != 5)
        return launch_preview(scn, sc, st, S, sample_base, d_tile_mask, d_sum, d_sumsq, clamp_max, stream, accumulate);
    return with_traversal(scn.stack_need, scn.width, [&](auto tag) {
        using T = decltype(tag);
        return generation_impl<T::kStack, T::kW>(scn, sc, st, S, sample_base, d_tile_mask, d_sum, d_sumsq,
                                                  clamp_max, accumulate, (hipStream_t)stream);
    });
}

int launch_accumulate(const PTScene &sc, const PTState &st, uint32_t s0, uint32_t S, const uint8_t *d_tile_mask,
                      float *d_sum, float *d_sumsq, float clamp_max, void *stream) {
    const uint32_t npix = sc.width * sc.height;
    hipLaunchKernelGGL(k_accumulate, dim3((npix + kPtBlock - 1) / kPtBlock), dim3(kPtBlock), 0, (hipStream_t)stream,
                       st, sc.width, npix, s0, S, d_tile_mask, clamp_max,
                       BackColor{{sc.back[0], sc.back[1], sc.back[2]}}, d_sum, d_sumsq);
    return chk(hipGetLastError(), "accumulate launch");
}

int launch_resolve(const float *d_sum, const uint32_t *d_nspp, uint32_t width, uint32_t height, float exposure,
                   float *d_hdr, void *stream) {
    const uint32_t npix = width * height;
    hipLaunchKernelGGL(k_resolve, dim3((npix + kPtBlock - 1) / kPtBlock), dim3(kPtBlock), 0, (hipStream_t)stream,
                       d_sum, d_nspp, width, npix, exposure, d_hdr);
    return chk(hipGetLastError(), "resolve launch");
}

// One thread per triangle: the records upload_pt_scene used to build on the
// host, with the same binary32 operations in the same order (strict FP,
// correctly rounded square root and division), so the values are the same.
__global__ __launch_bounds__(kPtBlock) void k_tri_shade(const float *__restrict__ v9, const float *__restrict__ n9,
                                                         const uint32_t *__restrict__ obj,
                                                         const float *__restrict__ scale, uint32_t n,
                                                         DevTriShade *__restrict__ out) {
    const uint32_t t = blockIdx.x * kPtBlock + threadIdx.x;
    if (t >= n) return;
    const float *v = v9 + 9 * (size_t)t;
    const v3 a = mk(v[0], v[1], v[2]), b = mk(v[3], v[4], v[5]), c = mk(v[6], v[7], v[8]);
    const v3 pd = cross(sub(b, a), sub(c, b));
    DevTriShade r;
    for (int k = 0; k < 9; k++) r.n[k] = n9[9 * (size_t)t + k];
    const float inv = 1.0f / len(pd);
    r.gn[0] = pd.x * inv;
    r.gn[1] = pd.y * inv;
    r.gn[2] = pd.z * inv;
    r.obj = obj[t];
    const v3 q = cross(sub(a, b), sub(a, c));
    r.ti_a = 1.0f / len(q);
    const v3 w = cross(sub(b, a), sub(c, a));
    r.area = 0.5f * len(w);
    r.sampled_area = scale[r.obj] > 0.0f ? scale[r.obj] * r.area : 0.0f;
    out[t] = r;
}

int launch_tri_shade(const float *d_v9, const float *d_n9, const uint32_t *d_obj, const float *d_scale, uint32_t n,
                     DevTriShade *d_out, void *stream) {
    if (n == 0) return RG_OK;
    hipLaunchKernelGGL(k_tri_shade, dim3((n + kPtBlock - 1) / kPtBlock), dim3(kPtBlock), 0, (hipStream_t)stream,
                       d_v9, d_n9, d_obj, d_scale, n, d_out);
    return chk(hipGetLastError(), "shading records launch");
}

int launch_fill_u32(void *p, uint32_t value, uint64_t n, void *stream) {
    if (n == 0) return RG_OK;
    hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)((n + kPtBlock - 1) / kPtBlock)), dim3(kPtBlock), 0,
                       (hipStream_t)stream, (uint32_t *)p, value, n);
    return chk(hipGetLastError(), "fill launch");
}

int launch_tile_noise(const float *d_sum, const float *d_sumsq, const uint32_t *d_nspp, uint32_t width,
                      uint32_t height, uint32_t target, float threshold, uint8_t *d_mask, uint32_t *d_nactive,
                      void *stream) {
    const uint32_t ntiles = ((width + 31) / 32) * ((height + 31) / 32);
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, (hipStream_t)stream, d_nactive, 0u);
    hipLaunchKernelGGL(k_tile_noise, dim3(ntiles), dim3(256), 0, (hipStream_t)stream, d_sum, d_sumsq, d_nspp,
                       width, height, target, threshold, d_mask, d_nactive);
    return chk(hipGetLastError(), "tile noise launch");
}

int launch_tiles_add(uint32_t *d_nspp, const uint8_t *d_mask, uint32_t ntiles, uint32_t add, void *stream) {
    hipLaunchKernelGGL(k_tiles_add, dim3((ntiles + kPtBlock - 1) / kPtBlock), dim3(kPtBlock), 0,
                       (hipStream_t)stream, d_nspp, d_mask, ntiles, add);
    return chk(hipGetLastError(), "tiles add launch");
}

}  // namespace rg

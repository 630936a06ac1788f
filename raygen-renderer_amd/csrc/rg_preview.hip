// rg_preview.hip — shading systems 0-3 (-d 0..3) for gfx950: the eye-ray
// providers of RayRenderer (rayrenderer.cpp:59-77), as one kernel.
//
//   0 RaySimpleShaderProvider            grey 0.1 + max(dot(l, n), 0)
//   1 RayAmbientOcclusionShaderProvider  calcAO within RAY_MAX_DISTANCE
//   2 LambertShaderProvider              lambertTraceLights x 0.3 x albedo
//   3 LambertWithAOShaderProvider        the same, 20 % + 80 % x calcAO(2.0)
//
// None of them bounces: a sample is the eye ray, its provider (calcAO casts
// settings.samples occlusion rays, the Lambert providers one area-light and
// every point light's shadow ray), and for a transparent Lambert surface the
// continuation straight through it (tracePath of the ThicknessRay).  With an
// active global medium every segment runs tracePath's free flight or delta
// tracking, emission integral and in-scattering (NEE toward area lights, the
// environment and emissive volumes, a phase-function bounce, Russian
// roulette), as the reference's traceEyeRay routes it (:1254-1273).
//
// One thread per (pixel, sample) traces all of this inline, with the LDS
// stack columns of the single-ray traversals: the occlusion loop of calcAO is
// uniform across a wave (every lane casts settings.samples rays), so the
// work stays converged without a wavefront round trip per AO ray.  The
// sample's value lands in the path state's radiance slot with no pending
// shadow rays, so the accumulation, adaptive tile decisions and resolve are
// the BSDF path's own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "rg_pt_dev.h"

namespace rg {

namespace {

// Upper bound on the transparent surfaces one sample crosses.  The
// reference's recursion has none; a straight ray meets each triangle at most
// once, so only scenes with more than this many stacked transparent layers
// along one line differ.
constexpr int kMaxCrossings = 256;

template <int kStack, int kW>
struct Tracer {
    DevBvh bv;
    LdsColumn<kPtBlock> col;
    uint32_t closest, any;

    __device__ __forceinline__ rg_hit nearest(v3 o, v3 d) {
        closest++;
        rg_hit h;
        h.prim = 0xFFFFFFFFu;
        h.t = MAX_RAY_DISTANCE;
        h.u = h.v = 0.0f;
        trace_closest_dev<kStack, kPtBlock, kW>(bv, o.x, o.y, o.z, d.x, d.y, d.z, 0.0f, h.t, h.prim, h.u, h.v, col);
        return h;
    }
    __device__ __forceinline__ bool occluded(v3 o, v3 d, float maxt, int pred) {
        any++;
        return trace_any_dev<kStack, kPtBlock, kW>(bv, o.x, o.y, o.z, d.x, d.y, d.z, maxt, 1u << pred, col);
    }
};

// ThicknessRay (raycommon.h:69-71): the origin pushed SURFACE_THICKNESS
// along the normalised direction.
__device__ __forceinline__ v3 thickness_origin(v3 o, v3 d) { return add(o, scl(nrm(d), SURFACE_THICKNESS)); }

// calcAO (rayrenderer.cpp:2381-2407): settings.samples cosine-weighted
// directions (an ldsNext2D pair each); an unoccluded one (no triangle with
// transparency < 0.01 within dist) adds its cosine; the mean.
template <class T>
__device__ __forceinline__ float calc_ao(const PTScene &sc, Rng &R, T &tr, v3 hit, v3 n, float dist) {
    float sum = 0.0f;
    for (uint32_t i = 0; i < sc.ao_samples; i++) {
        const float r1 = lds1(R), r2 = lds1(R);
        const v3 dir = cosine_dir(n, r1, r2);
        if (!tr.occluded(thickness_origin(hit, dir), dir, dist, RG_PRED_OPAQUE)) sum += fmaxf(dot(dir, n), 0.0f);
    }
    return sum / (float)sc.ao_samples;
}

// lambertTraceLights (rayrenderer.cpp:2319-2335): one area light's
// traceAreaLight, every point light's tracePointLight, plus worldColor.
template <class T>
__device__ __forceinline__ v3 lambert_lights(const PTScene &sc, Rng &R, T &tr, v3 hit, v3 n) {
    v3 a = mk(0, 0, 0), p = mk(0, 0, 0);
    if (sc.nlights > 0) {
        const AreaSample as = sample_area(sc, R, hit, n);
        if (as.ok && !tr.occluded(as.o, as.lr, 0.99999f, RG_PRED_AREA)) a = add(a, area_value(as));
    }
    for (uint32_t i = 0; i < sc.nplights; i++) {
        const PointSample ps = point_light_at(sc.plights[i], hit, n);
        if (ps.ok && !tr.occluded(ps.o, ps.lr, 0.99999f, RG_PRED_ENV)) p = add(p, ps.val);
    }
    return add(add(a, p), mk(sc.world[0], sc.world[1], sc.world[2]));
}

struct Local {
    v3 c;     // the provider's value without the continuation
    float t;  // weight of the continuation through the surface (0: none)
};

// The provider's value at a surface (rayrenderer.h:446-475,
// lambert.cpp:14-67): for a transparent Lambert surface, the part before
// "+ tracePath(...) * transparency" and that weight.
template <class T>
__device__ __forceinline__ Local provider(const PTScene &sc, Rng &R, T &tr, uint32_t prim, v3 pos, v3 n,
                                          float u, float v) {
    Local r;
    r.t = 0.0f;
    if (sc.shader == 0) {
        const float c = 0.1f + fmaxf(dot(nrm(mk(1.0f, 1.0f, 1.0f)), n), 0.0f);
        r.c = mk(c, c, c);
        return r;
    }
    if (sc.shader == 1) {
        const float c = fminf(fmaxf(calc_ao(sc, R, tr, pos, n, RAY_MAX_DISTANCE), 0.0f), 1.0f);
        r.c = mk(c, c, c);
        return r;
    }
    const DevObject &m = obj_of(sc, prim);
    const v3 col = ld3(m.color);
    if (m.emission > 0.0f) {
        r.c = scl(col, m.emission);
        return r;
    }
    const v3 light = scl(lambert_lights(sc, R, tr, pos, n), 0.3f);
    v3 c;
    if (sc.shader == 2) c = light;
    else c = add(scl(light, 0.2f), scl(scl(light, 0.8f), calc_ao(sc, R, tr, pos, n, 2.0f)));
    if (sc.color_sampling) c = mul(mul(c, col), albedo_tex(sc, prim, u, v));
    if (m.transparency > 0.0f) {
        r.t = m.transparency;
        if (sc.shader == 2) c = scl(mul(col, add(light, mk(0.5f, 0.5f, 0.5f))), 1.0f - m.transparency);
    }
    r.c = c;
    return r;
}

// traceEyeRay (rayrenderer.cpp:1254-1297) with the provider, surfaces only.
template <class T>
__device__ __forceinline__ v3 eye_surfaces(const PTScene &sc, Rng &R, T &tr, v3 o, v3 d) {
    v3 L = mk(0, 0, 0), w = mk(1, 1, 1);
    for (int k = 0; k <= kMaxCrossings; k++) {
        const rg_hit h = tr.nearest(o, d);
        const bool hit = h.prim != 0xFFFFFFFFu;
        if (hit && (k > 0 || obj_of(sc, h.prim).visible)) {
            const v3 pos = add(o, scl(d, h.t));
            const Local lc = provider(sc, R, tr, h.prim, pos, interp_normal(sc, h.prim, h.u, h.v), h.u, h.v);
            L = add(L, mul(w, lc.c));
            if (!(lc.t > 0.0f)) break;
            w = scl(w, lc.t);
            o = thickness_origin(pos, d);
            continue;
        }
        // the eye ray reads the environment, else backColor; a
        // continuation reads the environment alone (tracePath, :1619-1630)
        v3 env = sample_environment(sc, d);
        if (k == 0 && is_zero(env)) env = mk(sc.back[0], sc.back[1], sc.back[2]);
        L = add(L, mul(w, env));
        break;
    }
    return L;
}

// traceEyeRay through an active global medium: tracePath (:1596-1864) from
// the eye with no shaderParam, iterated.  Lambert surfaces keep the medium
// and pass their caller's shaderParam on, so the throughput used by Russian
// roulette stays (1, 1, 1) and only in-scattering adds passes.
template <class T>
__device__ __forceinline__ v3 eye_medium(const PTScene &sc, Rng &R, T &tr, const DevMedium &med, v3 o, v3 d) {
    v3 L = mk(0, 0, 0), w = mk(1, 1, 1);
    int passes = 0;
    bool eye_hit = false;
    for (int k = 0; k <= kMaxCrossings + sc.max_depth; k++) {
        const rg_hit h = tr.nearest(o, d);
        const bool hit = h.prim != 0xFFFFFFFFu;
        if (k == 0) eye_hit = hit;
        const float maxT = hit ? h.t : RAY_MAX_DISTANCE;
        bool scattered;
        float tf = 0.0f, den = 1.0f;
        if (med.het) {
            scattered = delta_track(med, R, o, d, maxT, tf, den);
            if (!scattered) tf = maxT;
        } else {
            tf = free_flight(med, prng(R));
            scattered = tf < maxT;
        }
        L = add(L, mul(w, emission_integral(med, R, o, d, scattered ? tf : maxT)));
        if (scattered) {
            v3 sw = mk(0, 0, 0);
            if (med.het) {
                const float inv = med.t_max_hero > 0.0f ? 1.0f / med.t_max_hero : 0.0f;
                sw = mk(med.sigma_s[0] * den * inv, med.sigma_s[1] * den * inv, med.sigma_s[2] * den * inv);
            } else {
                const float ph = med.hero * expf(-med.hero * tf);
                const v3 Tr = transmittance(med, tf);
                if (ph > 0.0f)
                    sw = mk(med.sigma_s[0] * Tr.x / ph, med.sigma_s[1] * Tr.y / ph, med.sigma_s[2] * Tr.z / ph);
            }
            if (is_zero(sw)) break;
            const v3 spos = add(o, scl(d, tf));
            passes++;
            if (passes > sc.max_depth) break;
            const v3 proxy = neg(d);
            v3 direct = mk(0, 0, 0);
            {
                const AreaSample as = sample_area(sc, R, spos, proxy);
                if (as.ok && as.pdf > 0.0f && !tr.occluded(as.o, as.lr, 0.99999f, RG_PRED_AREA))
                    direct = add(direct, scl(as.Le, phase_pdf(med, neg(d), as.dir) * (1.0f / as.pdf)));
            }
            {
                const EnvSample es = sample_env_nee(sc, R, spos, proxy);
                if (es.ok && es.pdf > 0.0f && !tr.occluded(es.o, es.dir, 1e30f, RG_PRED_ENV))
                    direct = add(direct, scl(es.Le, phase_pdf(med, neg(d), es.dir) * (1.0f / es.pdf)));
            }
            {
                const VolumeSample vs = sample_volume(sc, R, spos, proxy);
                if (vs.ok && vs.pdf > 0.0f && !tr.occluded(vs.o, vs.lr, 0.99999f, RG_PRED_VOLUME))
                    direct = add(direct, scl(vs.Le, phase_pdf(med, neg(d), vs.dir) * (1.0f / vs.pdf)));
            }
            const float u1 = prng(R), u2 = prng(R);
            const v3 nd = sample_phase(med, d, u1, u2);
            const v3 bsw = mul(w, sw);
            L = add(L, mul(bsw, direct));
            float rrw = 1.0f;
            if (passes >= MIN_RR_DEPTH) {
                const float q = fminf(RR_MAX_PROB, fmaxf(RR_MIN_PROB, maxc(sw)));
                if (prng(R) >= q) break;
                rrw = 1.0f / q;
            }
            w = scl(bsw, rrw);
            o = add(spos, scl(nrm(nd), SURFACE_THICKNESS));
            d = nd;
            continue;
        }
        if (!med.het) {
            const float surv = expf(-med.hero * maxT);
            if (surv > 0.0f) {
                const v3 Tr = transmittance(med, maxT);
                w = mul(w, mk(Tr.x / surv, Tr.y / surv, Tr.z / surv));
            }
        }
        if (!hit) {
            L = add(L, mul(w, sample_environment(sc, d)));
            break;
        }
        // the smooth-shading silhouette fix of the medium branch (:1824-1845)
        v3 n = interp_normal(sc, h.prim, h.u, h.v);
        const v3 g = ld3(sc.shade[h.prim].gn);
        const v3 gN = dot(g, d) <= 0.0f ? g : neg(g);
        if (dot(d, n) > 0.0f && dot(d, gN) < 0.0f) n = gN;
        const v3 pos = add(o, scl(d, h.t));
        const Local lc = provider(sc, R, tr, h.prim, pos, n, h.u, h.v);
        L = add(L, mul(w, lc.c));
        if (!(lc.t > 0.0f)) break;
        w = scl(w, lc.t);
        o = thickness_origin(pos, d);
    }
    if (!eye_hit && is_zero(L)) return mk(sc.back[0], sc.back[1], sc.back[2]);
    return L;
}

// Every sample of the generation: (pixel, sample) = (i % npix, i / npix),
// the wavefront's slot order.  The sample's value goes to st.L[i] with no
// pending shadow rays; stats count its closest-hit and occlusion rays.
template <int kStack, int kW, bool kMed>
__global__ __launch_bounds__(kPtBlock) void k_preview(DevBvh bv, PTScene sc, PTState st, uint32_t S,
                                                       uint32_t sample_base, const uint8_t *__restrict__ tile_mask) {
    RG_LDS_COLUMN(lc, kPtBlock);
    const uint32_t i = blockIdx.x * kPtBlock + threadIdx.x;
    const uint32_t npix = sc.width * sc.height;
    bool active = i < npix * S;
    Tracer<kStack, kW> tr{bv, lc, 0u, 0u};
    if (active) {
        const uint32_t pix = i % npix, s = i / npix;
        const uint32_t x = pix % sc.width, y = pix / sc.width;
        if (tile_mask && !tile_mask[(y / 32) * ((sc.width + 31) / 32) + x / 32]) active = false;
        if (active) {
            Rng R = rng_for(sc, pix, (int)(sample_base + s), 0, 0);
            v3 o, d;
            camera_ray(sc, R, x, y, o, d);
            v3 c;
            if (kMed && sc.media[sc.global_medium].active) c = eye_medium(sc, R, tr, sc.media[sc.global_medium], o, d);
            else c = eye_surfaces(sc, R, tr, o, d);
            st.L[i] = make_float4(c.x, c.y, c.z, 0.0f);
            st.meta[i] = make_int4(-1, pack_medium(-1, -1, false), R.dim, (int)R.k);
            st.nshadow[i] = 0;
        }
    }
    // ray and sample counts: a wave sum, one atomic per wave and counter
    uint32_t a = tr.closest, b = tr.any, n = active ? 1u : 0u;
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
        n += __shfl_xor(n, off);
    }
    if (__lane_id() == 0) {
        if (a) atomicAdd(st.stats + 0, (unsigned long long)a);
        if (b) atomicAdd(st.stats + 1, (unsigned long long)b);
        if (n) atomicAdd(st.stats + 2, (unsigned long long)n);
    }
}

int chk(hipError_t e, const char *what) {
    if (e == hipSuccess) return RG_OK;
    return set_error(RG_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

DevBvh dev_bvh(const Scene &s);

int launch_preview(const Scene &scn, const PTScene &sc, const PTState &st, uint32_t S, uint32_t sample_base,
                   const uint8_t *d_tile_mask, float *d_sum, float *d_sumsq, float clamp_max, void *stream,
                   bool accumulate) {
    const uint32_t npix = sc.width * sc.height;
    const uint64_t n = (uint64_t)npix * S;
    if (n > 0xFFFFFFFFull - kPtBlock) return set_error(RG_E_BUDGET, "too many samples in flight");
    const uint32_t blocks = (uint32_t)((n + kPtBlock - 1) / kPtBlock);
    const DevBvh bv = dev_bvh(scn);
    const bool med = sc.global_medium >= 0;
    const int rc = with_traversal(scn.stack_need, scn.width, [&](auto tag) {
        using T = decltype(tag);
        if (med)
            hipLaunchKernelGGL((k_preview<T::kStack, T::kW, true>), dim3(blocks), dim3(kPtBlock), 0,
                               (hipStream_t)stream, bv, sc, st, S, sample_base, d_tile_mask);
        else
            hipLaunchKernelGGL((k_preview<T::kStack, T::kW, false>), dim3(blocks), dim3(kPtBlock), 0,
                               (hipStream_t)stream, bv, sc, st, S, sample_base, d_tile_mask);
        return chk(hipGetLastError(), "preview launch");
    });
    if (rc) return rc;
    if (!accumulate) return RG_OK;
    return launch_accumulate(sc, st, 0, S, d_tile_mask, d_sum, d_sumsq, clamp_max, stream);
}

}  // namespace rg

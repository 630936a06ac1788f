// rg_pt_dev.h — device functions of the BSDF path tracer (shading system 5)
// shared by the single-job kernels (rg_pt.hip) and the batched streaming
// kernels (rg_stream.hip): vector helpers, random numbers, scene access,
// environment, media, GGX, NEE sampling, one path event (step) and the
// camera ray.  Binary32 arithmetic in explicit order, no contraction.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rg_device.h"
#include "rg_internal.h"
#include "rg_pt.h"

namespace rg {

#define PI_F 3.14159265358979323846f
#define SURFACE_THICKNESS 0.0001f
#define MAX_RAY_DISTANCE 1.0e6f
#define RAY_MAX_DISTANCE 100.0f
#define MIN_RR_DEPTH 3
#define RR_MIN_PROB 0.05f
#define RR_MAX_PROB 0.95f

// ---------------------------------------------------------------------------
// vector helpers: explicit component order, no contraction
// ---------------------------------------------------------------------------
struct v3 {
    float x, y, z;
};
__device__ __forceinline__ v3 mk(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ v3 mul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ v3 scl(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ v3 cross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float len(v3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ v3 nrm(v3 a) {
    const float inv = 1.0f / len(a);
    return scl(a, inv);
}
__device__ __forceinline__ v3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }
__device__ __forceinline__ float maxc(v3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
__device__ __forceinline__ bool is_zero(v3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
__device__ __forceinline__ v3 reflect(v3 d, v3 n) { return sub(d, scl(n, dot(d, n) * 2.0f)); }
__device__ __forceinline__ v3 f4v(float4 a) { return mk(a.x, a.y, a.z); }
__device__ __forceinline__ float4 v4(v3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

// ---------------------------------------------------------------------------
// random numbers: Halton + per-pixel Cranley-Patterson (raycommon.cpp:21-92)
// for the reference's LDS call sites, Philox4x32-10 for the rest
// ---------------------------------------------------------------------------
struct Rng {
    int x, y, sample, dim;
    uint32_t pix, key0, key1, k;
    uint32_t blk[4];
    uint32_t blk_id;
};

__device__ __forceinline__ float prng(Rng &r) {
    const uint32_t b = r.k >> 2;
    if (b != r.blk_id) {
        uint32_t c[4] = {r.pix, (uint32_t)r.sample, b, 0u};
        philox(c, r.key0, r.key1);
        r.blk[0] = c[0];
        r.blk[1] = c[1];
        r.blk[2] = c[2];
        r.blk[3] = c[3];
        r.blk_id = b;
    }
    // constant indices only, so the block stays in registers
    const uint32_t j = r.k & 3;
    const uint32_t w = j == 0 ? r.blk[0] : (j == 1 ? r.blk[1] : (j == 2 ? r.blk[2] : r.blk[3]));
    r.k++;
    return (float)(w >> 8) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ uint32_t hash_pixel(int x, int y, uint32_t salt) {
    uint32_t v = (uint32_t)x * 73856093u ^ (uint32_t)y * 19349663u ^ salt * 83492791u;
    v ^= v >> 16;
    v *= 0x7feb352du;
    v ^= v >> 15;
    v *= 0x846ca68bu;
    v ^= v >> 16;
    return v;
}

__device__ __forceinline__ float radical_inverse(uint32_t i, uint32_t base) {
    const float inv_base = 1.0f / (float)base;
    float inv_base_n = 1.0f;
    uint64_t rev = 0;
    while (i > 0) {
        const uint32_t next = i / base;
        const uint32_t digit = i - next * base;
        rev = rev * base + digit;
        inv_base_n *= inv_base;
        i = next;
    }
    const float v = (float)rev * inv_base_n;
    return v < 1.0f ? v : 0.9999999f;
}

// The first 16 primes, the Halton bases of dimensions 0-15, selected without
// a private-memory table.
__device__ __forceinline__ uint32_t halton_base(int dim) {
    uint32_t b = 53u;
    b = dim == 14 ? 47u : b;
    b = dim == 13 ? 43u : b;
    b = dim == 12 ? 41u : b;
    b = dim == 11 ? 37u : b;
    b = dim == 10 ? 31u : b;
    b = dim == 9 ? 29u : b;
    b = dim == 8 ? 23u : b;
    b = dim == 7 ? 19u : b;
    b = dim == 6 ? 17u : b;
    b = dim == 5 ? 13u : b;
    b = dim == 4 ? 11u : b;
    b = dim == 3 ? 7u : b;
    b = dim == 2 ? 5u : b;
    b = dim == 1 ? 3u : b;
    b = dim == 0 ? 2u : b;
    return b;
}

__device__ __forceinline__ float lds1(Rng &r) {
    if (r.dim >= 16) return prng(r);
    const uint32_t scramble = hash_pixel(r.x, r.y, 0);
    const float v = radical_inverse((uint32_t)r.sample + 1u, halton_base(r.dim));
    const uint32_t h = hash_pixel(r.dim, 0, scramble);
    const float offset = (float)(h >> 8) * (1.0f / (float)(1u << 24));
    float w = v + offset;
    if (w >= 1.0f) w -= 1.0f;
    r.dim++;
    return w;
}

__device__ __forceinline__ uint32_t pick(float u, uint32_t n) {
    const uint32_t i = (uint32_t)(u * (float)n);
    return i < n ? i : n - 1;
}

__device__ __forceinline__ Rng rng_for(const PTScene &sc, uint32_t pix, int sample, int dim, uint32_t k) {
    Rng r;
    r.x = (int)(pix % sc.width);
    r.y = (int)(pix / sc.width);
    r.sample = sample;
    r.dim = dim;
    r.pix = pix;
    r.key0 = sc.seed;
    r.key1 = sc.job;
    r.k = k;
    r.blk_id = 0xFFFFFFFFu;
    return r;
}

// ---------------------------------------------------------------------------
// scene access
// ---------------------------------------------------------------------------
__device__ __forceinline__ const DevObject &obj_of(const PTScene &sc, uint32_t t) {
    return sc.objs[sc.shade[t].obj];
}
__device__ __forceinline__ v3 tri_v(const PTScene &sc, uint32_t t, int k) { return ld3(sc.v9 + 9 * (size_t)t + 3 * k); }
__device__ __forceinline__ v3 tri_n(const PTScene &sc, uint32_t t, int k) { return ld3(sc.shade[t].n + 3 * k); }

// geomNormal (bsdf.cpp:27-30)
__device__ __forceinline__ v3 geom_normal(const PTScene &sc, uint32_t t, v3 shading_n) {
    const v3 g = ld3(sc.shade[t].gn);
    return dot(g, shading_n) >= 0.0f ? g : neg(g);
}

// SurfaceRay (raycommon.h:82-85): offset origin
__device__ __forceinline__ v3 surface_origin(v3 hit, v3 dir, v3 n) {
    const float side = (dot(dir, n) >= 0.0f) ? 1.0f : -1.0f;
    return add(hit, scl(n, side * SURFACE_THICKNESS));
}

// calcVertexInterpolation(info) (rayrenderer.cpp:2354-2362)
__device__ __forceinline__ v3 interp_normal(const PTScene &sc, uint32_t t, float u, float v) {
    const float w = 1.0f - u - v;
    const v3 n = add(add(scl(tri_n(sc, t, 0), w), scl(tri_n(sc, t, 1), u)), scl(tri_n(sc, t, 2), v));
    return nrm(n);
}

// calcVertexInterpolation(rt, hit) (rayrenderer.cpp:2337-2352)
__device__ __forceinline__ v3 interp_normal_at(const PTScene &sc, uint32_t t, v3 p) {
    const v3 f1 = sub(tri_v(sc, t, 0), p), f2 = sub(tri_v(sc, t, 1), p), f3 = sub(tri_v(sc, t, 2), p);
    const float ta = sc.shade[t].ti_a;
    const float a1 = len(cross(f2, f3)) * ta;
    const float a2 = len(cross(f3, f1)) * ta;
    const float a3 = len(cross(f1, f2)) * ta;
    return add(add(scl(tri_n(sc, t, 0), a1), scl(tri_n(sc, t, 1), a2)), scl(tri_n(sc, t, 2), a3));
}

// Texture::sample (texture.cpp:209-235) of the object's albedo texture at the
// hit's interpolated UV (calcVertexInterpolation, rayrenderer.cpp:2354-2362)
// times texTiling: UV folded into [0, 1), nearest texel, linear RGB (sRGB
// textures are converted at upload).  (1, 1, 1) for an untextured object.
__device__ __forceinline__ v3 albedo_tex(const PTScene &sc, uint32_t prim, float u, float v) {
    if (!sc.has_tex) return mk(1.0f, 1.0f, 1.0f);
    const int4 ot = sc.otex[sc.shade[prim].obj];
    if (ot.x < 0) return mk(1.0f, 1.0f, 1.0f);
    const float *q = sc.uv6 + 6 * (size_t)prim;
    const float w = 1.0f - u - v;
    const float su = (q[0] * w + q[2] * u + q[4] * v) * __int_as_float(ot.y);
    const float sv = (q[1] * w + q[3] * u + q[5] * v) * __int_as_float(ot.z);
    if (!(su == su) || !(sv == sv)) return mk(0.0f, 0.0f, 0.0f);
    float fu = su - floorf(su), fv = sv - floorf(sv);
    if (fu < 0.0f || fu >= 1.0f) fu = 0.0f;
    if (fv < 0.0f || fv >= 1.0f) fv = 0.0f;
    const uint4 d = sc.tex_desc[ot.x];
    const int W = (int)d.x, H = (int)d.y;
    int x = (int)(fu * (float)W), y = (int)(fv * (float)H);
    x = x >= W ? W - 1 : (x < 0 ? 0 : x);
    y = y >= H ? H - 1 : (y < 0 ? 0 : y);
    return f4v(sc.texels[(size_t)d.z + (size_t)y * W + x]);
}

// ---------------------------------------------------------------------------
// environment (rayrenderer.cpp:1299-1573)
// ---------------------------------------------------------------------------
__device__ __forceinline__ v3 env_texel(const PTScene &sc, int x, int y) {
    return ld3(sc.env_rgb + 3 * ((size_t)y * sc.env_w + (size_t)x));
}
__device__ __forceinline__ float lum(v3 c) { return fmaxf(0.0f, 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z); }

__device__ __forceinline__ v3 sample_environment(const PTScene &sc, v3 dir) {
    if (!sc.has_env) return mk(0, 0, 0);
    const v3 d = nrm(dir);
    const float cy = sc.env_cos, sy = sc.env_sin;
    const float rx = d.x * cy - d.z * sy;
    const float rz = d.x * sy + d.z * cy;
    const float ry = d.y;
    const float u = 0.5f + atan2f(rz, rx) * (float)(0.5 / M_PI);
    const float v = 0.5f - asinf(fminf(fmaxf(ry, -1.0f), 1.0f)) * (float)(1.0 / M_PI);
    const int W = (int)sc.env_w, H = (int)sc.env_h;
    if (!(u == u) || !(v == v)) return mk(0, 0, 0);
    float fu = u - floorf(u), fv = v - floorf(v);
    if (fu < 0.0f || fu >= 1.0f) fu = 0.0f;
    if (fv < 0.0f || fv >= 1.0f) fv = 0.0f;
    int x = (int)(fu * (float)W), y = (int)(fv * (float)H);
    if (x >= W) x = W - 1;
    if (y >= H) y = H - 1;
    return scl(env_texel(sc, x, y), sc.env_intensity);
}

__device__ __forceinline__ int search_cdf(const float *cdf, int n, float u) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid + 1] <= u) lo = mid + 1;
        else hi = mid;
    }
    return lo < n ? lo : n - 1;
}

__device__ __forceinline__ void sample_env_dir(const PTScene &sc, float u0, float u1, v3 &out, float &pdf) {
    out = mk(0, 1, 0);
    pdf = 0.0f;
    const int W = (int)sc.env_w, H = (int)sc.env_h;
    if (!sc.has_env || sc.env_total <= 0.0f) return;
    const int y = search_cdf(sc.env_marg, H, u0);
    const float *row = sc.env_cond + (size_t)y * (W + 1);
    const int x = search_cdf(row, W, u1);
    const float yp = sc.env_marg[y], ys = sc.env_marg[y + 1] - yp;
    const float xp = row[x], xs = row[x + 1] - xp;
    const float yf = (ys > 0.0f) ? (u0 - yp) / ys : 0.5f;
    const float xf = (xs > 0.0f) ? (u1 - xp) / xs : 0.5f;
    const float u = ((float)x + xf) / (float)W;
    const float v = ((float)y + yf) / (float)H;
    const float phi = (u - 0.5f) * 2.0f * PI_F;
    const float theta = (0.5f - v) * PI_F;
    const float ct = cosf(theta), st = sinf(theta);
    const float dx = ct * cosf(phi), dy = st, dz = ct * sinf(phi);
    const float cy = sc.env_cos, sy = sc.env_sin;
    const float wx = dx * cy + dz * sy;
    const float wz = -dx * sy + dz * cy;
    out = nrm(mk(wx, dy, wz));
    const float l = lum(env_texel(sc, x, y));
    const float sin_t = sinf(PI_F * v);
    if (sin_t <= 0.0f) {
        pdf = 0.0f;
        return;
    }
    const float pdf_uv = (l * sin_t) / sc.env_total;
    pdf = pdf_uv * (float)(W * H) / (2.0f * (float)(M_PI * M_PI) * sin_t);
}

__device__ __forceinline__ float env_dir_pdf(const PTScene &sc, v3 dir) {
    if (!sc.has_env || sc.env_total <= 0.0f) return 0.0f;
    const v3 d = nrm(dir);
    const float cy = sc.env_cos, sy = sc.env_sin;
    const float rx = d.x * cy - d.z * sy;
    const float rz = d.x * sy + d.z * cy;
    const int W = (int)sc.env_w, H = (int)sc.env_h;
    const float u = 0.5f + atan2f(rz, rx) * (float)(0.5 / M_PI);
    const float v = 0.5f - asinf(fminf(fmaxf(d.y, -1.0f), 1.0f)) * (float)(1.0 / M_PI);
    int x = (int)(u * (float)W);
    if (x < 0) x = 0;
    if (x >= W) x = W - 1;
    int y = (int)(v * (float)H);
    if (y < 0) y = 0;
    if (y >= H) y = H - 1;
    const float l = lum(env_texel(sc, x, y));
    const float sin_t = sinf(PI_F * v);
    if (sin_t <= 0.0f) return 0.0f;
    const float pdf_uv = (l * sin_t) / sc.env_total;
    return pdf_uv * (float)(W * H) / (2.0f * (float)(M_PI * M_PI) * sin_t);
}

__device__ __forceinline__ float env_miss_weight(const PTScene &sc, v3 dir, float bsdf_pdf) {
    if (bsdf_pdf > 0.0f) {
        const float ep = env_dir_pdf(sc, dir);
        const float b2 = bsdf_pdf * bsdf_pdf, e2 = ep * ep;
        return (b2 + e2 > 0.0f) ? b2 / (b2 + e2) : 1.0f;
    }
    return 1.0f;
}

// ---------------------------------------------------------------------------
// homogeneous media (medium.cpp:18-123, 475-525)
// ---------------------------------------------------------------------------
__device__ __forceinline__ v3 transmittance(const DevMedium &m, float d) {
    if (d <= 0.0f) return mk(1, 1, 1);
    return mk(expf(-m.sigma_t[0] * d), expf(-m.sigma_t[1] * d), expf(-m.sigma_t[2] * d));
}
__device__ __forceinline__ float ew_channel(float st, float d) {
    if (st < 1e-6f) return d;
    return (1.0f - expf(-st * d)) / st;
}
__device__ __forceinline__ v3 emission_along(const DevMedium &m, float d) {
    if (d <= 0.0f) return mk(0, 0, 0);
    return mk(m.sigma_e[0] * ew_channel(m.sigma_t[0], d), m.sigma_e[1] * ew_channel(m.sigma_t[1], d),
              m.sigma_e[2] * ew_channel(m.sigma_t[2], d));
}
__device__ __forceinline__ float free_flight(const DevMedium &m, float u) {
    if (m.hero <= 0.0f) return 1e30f;
    const float xi = fminf(0.9999999f, fmaxf(0.0f, u));
    return -logf(1.0f - xi) / m.hero;
}
__device__ __forceinline__ float phase_pdf(const DevMedium &m, v3 wo, v3 wi) {
    const float c = dot(wo, wi);
    const float g = m.g;
    const float den = 1.0f + g * g - 2.0f * g * c;
    return (1.0f - g * g) * (float)(1.0 / (4.0 * M_PI)) / (den * sqrtf(fmaxf(1e-8f, den)));
}
__device__ __forceinline__ v3 sample_phase(const DevMedium &m, v3 wo, float u1, float u2) {
    float ct;
    if (fabsf(m.g) < 1e-3f) ct = 1.0f - 2.0f * u1;
    else {
        const float g = m.g;
        const float sq = (1.0f - g * g) / (1.0f - g + 2.0f * g * u1);
        ct = (1.0f + g * g - sq * sq) / (2.0f * g);
    }
    const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    const float phi = 2.0f * PI_F * u2;
    v3 t1;
    if (fabsf(wo.x) > fabsf(wo.y)) {
        const float il = 1.0f / sqrtf(wo.x * wo.x + wo.z * wo.z);
        t1 = mk(-wo.z * il, 0.0f, wo.x * il);
    } else {
        const float il = 1.0f / sqrtf(wo.y * wo.y + wo.z * wo.z);
        t1 = mk(0.0f, wo.z * il, -wo.y * il);
    }
    const v3 t2 = cross(wo, t1);
    return nrm(add(add(scl(t1, st * cosf(phi)), scl(t2, st * sinf(phi))), scl(wo, ct)));
}

// ---------------------------------------------------------------------------
// heterogeneous and emissive media, heat haze (medium.cpp:125-472)
// ---------------------------------------------------------------------------
// value-noise basis: integer lattice hash to [0, 1), smoothstep-weighted
// trilinear interpolation, fBm normalised by the summed amplitudes
__device__ __forceinline__ float lattice(int x, int y, int z) {
    uint32_t h = (uint32_t)x * 374761393u + (uint32_t)y * 668265263u + (uint32_t)z * 2147483647u;
    h = (h ^ (h >> 13)) * 1274126177u;
    h ^= h >> 16;
    return (float)(h & 0x00FFFFFFu) / (float)0x01000000u;
}
__device__ __forceinline__ float fade(float x) { return x * x * (3.0f - 2.0f * x); }
__device__ __forceinline__ float value_noise(float x, float y, float z) {
    const int xi = (int)floorf(x), yi = (int)floorf(y), zi = (int)floorf(z);
    const float u = fade(x - (float)xi), v = fade(y - (float)yi), w = fade(z - (float)zi);
    const float x00 = lattice(xi, yi, zi) * (1.0f - u) + lattice(xi + 1, yi, zi) * u;
    const float x10 = lattice(xi, yi + 1, zi) * (1.0f - u) + lattice(xi + 1, yi + 1, zi) * u;
    const float x01 = lattice(xi, yi, zi + 1) * (1.0f - u) + lattice(xi + 1, yi, zi + 1) * u;
    const float x11 = lattice(xi, yi + 1, zi + 1) * (1.0f - u) + lattice(xi + 1, yi + 1, zi + 1) * u;
    const float y0 = x00 * (1.0f - v) + x10 * v;
    const float y1 = x01 * (1.0f - v) + x11 * v;
    return y0 * (1.0f - w) + y1 * w;
}
__device__ __forceinline__ float fbm01(float x, float y, float z, float freq, int octaves, float gain, float lac) {
    float fx = x * freq, fy = y * freq, fz = z * freq;
    float amp = 1.0f, total = 0.0f, sum = 0.0f;
    const int n = octaves > 0 ? octaves : 1;
    for (int i = 0; i < n; i++) {
        total += amp * value_noise(fx, fy, fz);
        sum += amp;
        amp *= gain;
        fx *= lac;
        fy *= lac;
        fz *= lac;
    }
    return sum > 0.0f ? total / sum : 0.0f;
}

// densityAt (medium.cpp:271-290): 1 without a density field
__device__ __forceinline__ float density_at(const DevMedium &m, v3 p) {
    if (!m.het) return 1.0f;
    const rg_medium &s = m.src;
    const float n = fbm01(p.x - s.noise_offset[0], p.y - s.noise_offset[1], p.z - s.noise_offset[2],
                          s.noise_frequency, s.noise_octaves, s.noise_gain, s.noise_lacunarity);
    return fmaxf(0.0f, s.noise_amplitude * n + s.noise_bias);
}

// emissionAt (medium.cpp:292-404): constant sigma_e (x density), a path
// tube's radial falloff and inner-to-outer blend, or the cone's axial
// gaussian and (1 - v^2)^2 radial profile, all times the density field
__device__ __forceinline__ v3 emission_at(const DevMedium &m, v3 p) {
    const rg_medium &s = m.src;
    if (s.emission_mode == 0)
        return m.het ? scl(mk(m.sigma_e[0], m.sigma_e[1], m.sigma_e[2]), density_at(m, p))
                     : mk(m.sigma_e[0], m.sigma_e[1], m.sigma_e[2]);
    if (s.emission_mode == 2) {
        const uint32_t n = s.n_path;
        if (n < 2) return mk(0, 0, 0);
        float best = __builtin_inff(), bt = 0.0f, br = 0.0f;
        for (uint32_t i = 0; i + 1 < n; i++) {
            const rg_path_point &a = s.path[i], &b = s.path[i + 1];
            const v3 ab = mk(b.p[0] - a.p[0], b.p[1] - a.p[1], b.p[2] - a.p[2]);
            const float ab2 = ab.x * ab.x + ab.y * ab.y + ab.z * ab.z;
            if (ab2 < 1e-12f) continue;
            const v3 ap = mk(p.x - a.p[0], p.y - a.p[1], p.z - a.p[2]);
            float u = (ap.x * ab.x + ap.y * ab.y + ap.z * ab.z) / ab2;
            u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
            const v3 q = mk(a.p[0] + ab.x * u, a.p[1] + ab.y * u, a.p[2] + ab.z * u);
            const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
            const float d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < best) {
                best = d2;
                bt = a.t + (b.t - a.t) * u;
                br = a.radius + (b.radius - a.radius) * u;
            }
        }
        if (br <= 0.0f) return mk(0, 0, 0);
        if (best >= br * br) return mk(0, 0, 0);
        const float r = sqrtf(best) / br;
        const float radial = powf(fmaxf(0.0f, 1.0f - r), s.path_falloff_power);
        if (radial <= 0.0f) return mk(0, 0, 0);
        const float tb = bt < 0.0f ? 0.0f : (bt > 1.0f ? 1.0f : bt);
        const v3 col = add(scl(ld3(s.path_inner), 1.0f - tb), scl(ld3(s.path_outer), tb));
        float k = radial * s.path_intensity * fmaxf(0.0f, s.density);
        if (m.het) k *= density_at(m, p);
        return scl(col, k);
    }
    const v3 rel = mk(p.x - s.cone_origin[0], p.y - s.cone_origin[1], p.z - s.cone_origin[2]);
    const v3 ax = mk(m.cone_axis[0], m.cone_axis[1], m.cone_axis[2]);
    const float at = rel.x * ax.x + rel.y * ax.y + rel.z * ax.z;
    if (s.cone_length <= 0.0f) return mk(0, 0, 0);
    const float u = at / s.cone_length;
    if (u < 0.0f || u > 1.0f) return mk(0, 0, 0);
    const v3 pp = mk(rel.x - ax.x * at, rel.y - ax.y * at, rel.z - ax.z * at);
    const float r2 = pp.x * pp.x + pp.y * pp.y + pp.z * pp.z;
    if (s.cone_radius <= 0.0f) return mk(0, 0, 0);
    const float v2 = r2 / (s.cone_radius * s.cone_radius);
    if (v2 > 1.0f) return mk(0, 0, 0);
    const float dx = (u - s.cone_peak_axial) * s.cone_peak_sharpness;
    const float t = expf(-dx * dx) * ((1.0f - v2) * (1.0f - v2));
    if (t <= 0.0f) return mk(0, 0, 0);
    const v3 col = add(scl(ld3(s.cone_inner), t), scl(ld3(s.cone_outer), 1.0f - t));
    float k = t * s.cone_intensity * fmaxf(0.0f, s.density);
    if (m.het) k *= density_at(m, p);
    return scl(col, k);
}

// sampleDeltaTracking (medium.cpp:406-440): exponential steps at the
// bound's rate, each accepted with probability density / peak; at most 64
// (inlined: a call taking the generator by reference would move it to
// scratch memory for the whole kernel)
__device__ __forceinline__ bool delta_track(const DevMedium &m, Rng &R, v3 o, v3 d, float maxT, float &out_t,
                                         float &out_den) {
    if (m.t_max_hero <= 0.0f || maxT <= 0.0f) return false;
    const float peak = m.hero > 0.0f ? m.t_max_hero / m.hero : 1.0f;
    if (peak <= 0.0f) return false;
    float t = 0.0f;
    for (int i = 0; i < 64; i++) {
        const float xi = fminf(0.9999999f, fmaxf(0.0f, prng(R)));
        t += -logf(1.0f - xi) / m.t_max_hero;
        if (t >= maxT) return false;
        const float den = density_at(m, add(o, scl(d, t)));
        if (prng(R) < den / peak) {
            out_t = t;
            out_den = den;
            return true;
        }
    }
    return false;
}

// emissionIntegralAlongRay (medium.cpp:442-472): the closed form for
// constant emission, else one jittered sample per stratum of Tr x sigma_e
__device__ __forceinline__ v3 emission_integral(const DevMedium &m, Rng &R, v3 o, v3 d, float maxT) {
    const rg_medium &s = m.src;
    if (s.emission_mode == 0) return emission_along(m, maxT);
    if (maxT <= 0.0f) return mk(0, 0, 0);
    const int cnt = s.emission_mode == 2 ? s.path_samples : s.cone_samples;
    const int N = cnt > 0 ? cnt : 1;
    const float dt = maxT / (float)N;
    v3 acc = mk(0, 0, 0);
    for (int i = 0; i < N; i++) {
        const float t = ((float)i + prng(R)) * dt;
        const v3 se = emission_at(m, add(o, scl(d, t)));
        if (is_zero(se)) continue;
        const v3 Tr = transmittance(m, t);
        acc.x += se.x * Tr.x * dt;
        acc.y += se.y * Tr.y * dt;
        acc.z += se.z * Tr.z * dt;
    }
    return acc;
}

// iorAt / iorGradientAt / bendRay (medium.cpp:188-269): n = 1 + amplitude x
// fade x 2 (fbm - 1/2), central differences at half a noise cell, and an
// eikonal march of iorMarchSteps midpoint steps
__device__ __forceinline__ float ior_at(const DevMedium &m, v3 p) {
    const rg_medium &s = m.src;
    const float n = fbm01(p.x - s.ior_offset[0], p.y - s.ior_offset[1], p.z - s.ior_offset[2], s.ior_frequency,
                          s.ior_octaves, s.ior_gain, s.ior_lacunarity);
    float am = 1.0f;
    if (s.ior_falloff != 0.0f && s.cone_length > 0.0f) {
        const float rx = p.x - s.cone_origin[0], ry = p.y - s.cone_origin[1], rz = p.z - s.cone_origin[2];
        const float at = rx * m.cone_axis[0] + ry * m.cone_axis[1] + rz * m.cone_axis[2];
        float u = at / s.cone_length;
        u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
        const float mag = s.ior_falloff < 0.0f ? -s.ior_falloff : s.ior_falloff;
        const float t = s.ior_falloff >= 0.0f ? u : 1.0f - u;
        am = 1.0f - mag * t;
        if (am < 0.0f) am = 0.0f;
    }
    return 1.0f + s.ior_amplitude * am * (n - 0.5f) * 2.0f;
}
struct Ray3 {
    v3 o, d;
};
__device__ __forceinline__ Ray3 bend_ray(const DevMedium &m, v3 o, v3 d, float length) {
    if (length <= 0.0f || !m.haze) return {add(o, scl(d, length)), d};
    const int N = m.src.ior_march_steps;
    const float ds = length / (float)N;
    const float f = m.src.ior_frequency;
    const float h = f > 0.0f ? 0.5f / f : 0.05f;
    const float i2h = 0.5f / h;
    v3 pos = o, dir = d;
    for (int i = 0; i < N; i++) {
        const v3 mid = mk(pos.x + dir.x * (ds * 0.5f), pos.y + dir.y * (ds * 0.5f), pos.z + dir.z * (ds * 0.5f));
        const float n = ior_at(m, mid);
        const float xp = ior_at(m, mk(mid.x + h, mid.y, mid.z)), xm = ior_at(m, mk(mid.x - h, mid.y, mid.z));
        const float yp = ior_at(m, mk(mid.x, mid.y + h, mid.z)), ym = ior_at(m, mk(mid.x, mid.y - h, mid.z));
        const float zp = ior_at(m, mk(mid.x, mid.y, mid.z + h)), zm = ior_at(m, mk(mid.x, mid.y, mid.z - h));
        const v3 g = mk((xp - xm) * i2h, (yp - ym) * i2h, (zp - zm) * i2h);
        const float gd = g.x * dir.x + g.y * dir.y + g.z * dir.z;
        const v3 pp = mk(g.x - dir.x * gd, g.y - dir.y * gd, g.z - dir.z * gd);
        const float in = n > 1e-6f ? 1.0f / n : 1.0f;
        const v3 nd = mk(dir.x + pp.x * (ds * in), dir.y + pp.y * (ds * in), dir.z + pp.z * (ds * in));
        const float l2 = nd.x * nd.x + nd.y * nd.y + nd.z * nd.z;
        if (l2 > 1e-12f) {
            const float il = 1.0f / sqrtf(l2);
            dir = mk(nd.x * il, nd.y * il, nd.z * il);
        }
        pos = mk(pos.x + dir.x * ds, pos.y + dir.y * ds, pos.z + dir.z * ds);
    }
    return {pos, dir};
}

// ---------------------------------------------------------------------------
// sampling helpers
// ---------------------------------------------------------------------------
// cosineWeightedDirection (raycommon.cpp:94-115)
__device__ __forceinline__ v3 cosine_dir(v3 n, float r1, float r2) {
    const float theta = acosf(sqrtf(1.0f - r1));
    const float phi = 2.0f * PI_F * r2;
    const float x = sinf(theta) * cosf(phi);
    const float y = sinf(theta) * sinf(phi);
    const float z = cosf(theta);
    v3 tg;
    if (fabsf(n.x) > fabsf(n.y)) tg = nrm(cross(mk(0, 1, 0), n));
    else tg = nrm(cross(mk(1, 0, 0), n));
    const v3 bt = cross(n, tg);
    return nrm(add(add(scl(tg, x), scl(bt, y)), scl(n, z)));
}

// ldsPointInTriangle (raycommon.cpp:117-122)
__device__ __forceinline__ v3 point_in_tri(const PTScene &sc, uint32_t t, float u, float v) {
    if (u + v > 1.0f) {
        u = 1.0f - u;
        v = 1.0f - v;
    }
    const v3 a = tri_v(sc, t, 0);
    return add(add(a, scl(sub(tri_v(sc, t, 1), a), u)), scl(sub(tri_v(sc, t, 2), a), v));
}

// randomRayInHemisphere (ugm, not vendored): uniform hemisphere about n
__device__ __forceinline__ v3 hemisphere_dir(v3 n, float u1, float u2) {
    const float z = u1;
    const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    const float phi = 2.0f * PI_F * u2;
    v3 tg;
    if (fabsf(n.x) > fabsf(n.y)) tg = nrm(cross(mk(0, 1, 0), n));
    else tg = nrm(cross(mk(1, 0, 0), n));
    const v3 bt = cross(n, tg);
    return add(add(scl(tg, r * cosf(phi)), scl(bt, r * sinf(phi))), scl(n, z));
}

__device__ __forceinline__ float fresnel_schlick(float c, float ior) {
    float r0 = (1.0f - ior) / (1.0f + ior);
    r0 = r0 * r0;
    const float m = 1.0f - c;
    return r0 + (1.0f - r0) * (m * m) * (m * m) * m;
}

// refract (bsdf.cpp:39-52)
__device__ __forceinline__ v3 refract_dir(v3 d, v3 n, float r) {
    const v3 nl = dot(d, n) < 0.0f ? n : neg(n);
    const bool into = dot(nl, n) > 0.0f;
    if (into) r = 1.0f / r;
    const float c = dot(d, nl);
    const float t = 1.0f - r * r * (1.0f - c * c);
    if (t < 0.0f) return reflect(d, n);
    return nrm(sub(scl(d, r), scl(n, (into ? 1.0f : -1.0f) * (c * r + sqrtf(t)))));
}

struct Ggx {
    v3 t, b, n;
    float ax, ay;
    v3 V;
    float g1v;
    v3 F0;
};

__device__ __forceinline__ float smith_g1(v3 w, float ax, float ay) {
    if (w.z <= 0.0f) return 0.0f;
    const float ax2 = (ax * w.x) * (ax * w.x);
    const float ay2 = (ay * w.y) * (ay * w.y);
    const float lambda = 0.5f * (sqrtf(1.0f + (ax2 + ay2) / (w.z * w.z)) - 1.0f);
    return 1.0f / (1.0f + lambda);
}
// Heitz 2018 VNDF sampling (bsdf.cpp:139-161)
__device__ __forceinline__ v3 sample_vndf(v3 Ve, float ax, float ay, float u1, float u2) {
    const v3 Vh = nrm(mk(ax * Ve.x, ay * Ve.y, Ve.z));
    const float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    const v3 T1 = (lensq > 0.0f) ? scl(mk(-Vh.y, Vh.x, 0.0f), 1.0f / sqrtf(lensq)) : mk(1, 0, 0);
    const v3 T2 = cross(Vh, T1);
    const float r = sqrtf(u1);
    const float phi = 2.0f * PI_F * u2;
    const float t1 = r * cosf(phi);
    float t2 = r * sinf(phi);
    const float s = 0.5f * (1.0f + Vh.z);
    t2 = (1.0f - s) * sqrtf(fmaxf(0.0f, 1.0f - t1 * t1)) + s * t2;
    const v3 Nh = add(add(scl(T1, t1), scl(T2, t2)), scl(Vh, sqrtf(fmaxf(0.0f, 1.0f - t1 * t1 - t2 * t2))));
    return nrm(mk(ax * Nh.x, ay * Nh.y, fmaxf(0.0f, Nh.z)));
}
__device__ __forceinline__ v3 schlick_f(v3 F0, float vdoth) {
    const float a1 = fmaxf(0.0f, 1.0f - vdoth);
    const float a5 = a1 * a1 * a1 * a1 * a1;
    return add(F0, scl(sub(mk(1, 1, 1), F0), a5));
}
__device__ __forceinline__ float ggx_d(v3 H, float ax, float ay) {
    const float hx = H.x / ax, hy = H.y / ay;
    const float den = hx * hx + hy * hy + H.z * H.z;
    return 1.0f / (PI_F * ax * ay * den * den);
}
// evalGlossy (bsdf.cpp:258-280)
__device__ __forceinline__ bool eval_glossy(const Ggx &G, v3 dw, v3 &fcos, float &pdf) {
    const v3 Ll = mk(dot(dw, G.t), dot(dw, G.b), dot(dw, G.n));
    if (Ll.z <= 0.0f) return false;
    v3 Hl = add(G.V, Ll);
    const float h2 = Hl.x * Hl.x + Hl.y * Hl.y + Hl.z * Hl.z;
    if (h2 <= 0.0f) return false;
    Hl = scl(Hl, 1.0f / sqrtf(h2));
    if (Hl.z <= 0.0f) return false;
    const float vdh = fmaxf(0.0f, dot(G.V, Hl));
    if (vdh <= 0.0f) return false;
    const float D = ggx_d(Hl, G.ax, G.ay);
    const float g1l = smith_g1(Ll, G.ax, G.ay);
    const v3 F = schlick_f(G.F0, vdh);
    fcos = scl(F, D * G.g1v * g1l / (4.0f * G.V.z));
    pdf = G.g1v * D / (4.0f * G.V.z);
    return true;
}
__device__ __forceinline__ bool ggx_setup(const DevObject &m, v3 alb, v3 n, v3 in_dir, Ggx &G) {
    G.F0 = (m.metallic >= 1.0f) ? alb : add(scl(alb, m.metallic), scl(mk(0.04f, 0.04f, 0.04f), 1.0f - m.metallic));
    const v3 up = (fabsf(n.y) > 0.99f) ? mk(1, 0, 0) : mk(0, 1, 0);
    G.t = nrm(cross(up, n));
    G.b = cross(n, G.t);
    if (fabsf(m.aniso_rotation) > 0.0f) {
        const float a = m.aniso_rotation * (float)(M_PI / 180.0);
        const float ca = cosf(a), sa = sinf(a);
        const v3 t2 = add(scl(G.t, ca), scl(G.b, sa));
        const v3 b2 = sub(scl(G.b, ca), scl(G.t, sa));
        G.t = t2;
        G.b = b2;
    }
    G.n = n;
    const float an = fmaxf(-0.99f, fminf(0.99f, m.anisotropy));
    const float r2 = m.roughness * m.roughness;
    G.ax = fmaxf(1e-3f, r2 * (1.0f - an));
    G.ay = fmaxf(1e-3f, r2 * (1.0f + an));
    const v3 V = neg(in_dir);
    G.V = mk(dot(V, G.t), dot(V, G.b), dot(V, n));
    if (G.V.z <= 0.0f) return false;
    G.g1v = smith_g1(G.V, G.ax, G.ay);
    return true;
}

// ---------------------------------------------------------------------------
// shadow-ray emission: NEE samples are drawn in the shade kernel and tested
// in the shadow kernel; a contribution is added only when unoccluded
// ---------------------------------------------------------------------------
// Up to kNeeSlots NEE rays per event, written to the path's slots
// k * slot_stride + p as they are drawn (ray: tmin = predicate; contribution).
struct Shadows {
    int n;
    uint32_t p, stride;
    rg_ray *sray;
    float4 *scontrib;
};

__device__ __forceinline__ Shadows shadows_of(const PTState &st, uint32_t p) {
    Shadows S;
    S.n = 0;
    S.p = p;
    S.stride = st.slot_stride;
    S.sray = st.sray;
    S.scontrib = st.scontrib;
    return S;
}

__device__ __forceinline__ void push_shadow(Shadows &S, v3 o, v3 d, float tmax, int pred, v3 contrib) {
    const uint32_t k = (uint32_t)S.n * S.stride + S.p;
    S.sray[k] = {{o.x, o.y, o.z}, (float)pred, {d.x, d.y, d.z}, tmax};
    S.scontrib[k] = v4(contrib, 0.0f);
    S.n++;
}

struct AreaSample {
    bool ok;
    v3 dir, Le, o, lr;
    float pdf, co, cl, r2, sa;
};

// sampleAreaLightForNEE (rayrenderer.cpp:1976-2036) up to the shadow test
__device__ __forceinline__ AreaSample sample_area(const PTScene &sc, Rng &R, v3 hit, v3 n) {
    AreaSample s;
    s.ok = false;
    if (sc.nlights == 0) return s;
    const uint32_t li = pick(prng(R), sc.nlights);
    const DevLight L = sc.lights[li];
    const uint32_t tri = sc.light_tris[L.first + pick(prng(R), L.count)];
    const float u = lds1(R), v = lds1(R);
    const v3 p = point_in_tri(sc, tri, u, v);
    const v3 lr = sub(p, hit);
    const v3 dir = nrm(lr);
    const float co = dot(dir, n);
    if (co <= 0.0f) return s;
    const v3 ln = interp_normal_at(sc, tri, p);
    const float cl = dot(neg(dir), ln);
    if (cl <= 0.0f) return s;
    const DevObject &lo = sc.objs[L.obj];
    const float sa = (float)L.count * sc.shade[tri].area;
    const float r2 = lr.x * lr.x + lr.y * lr.y + lr.z * lr.z;
    s.ok = true;
    s.dir = dir;
    s.pdf = r2 / (cl * sa);
    s.Le = scl(ld3(lo.color), lo.emission);
    s.co = co;
    s.cl = cl;
    s.r2 = r2;
    s.sa = sa;
    s.lr = lr;
    s.o = surface_origin(hit, lr, n);
    return s;
}

// traceAreaLight's value (rayrenderer.cpp:1916-1974): Lambert NEE incl. 1/pi and MIS
__device__ __forceinline__ v3 area_value(const AreaSample &s) {
    const float pl = s.r2 / (s.cl * s.sa);
    const float pb = s.co / PI_F;
    const float pl2 = pl * pl, pb2 = pb * pb;
    const float w = pl2 / (pl2 + pb2);
    return scl(scl(s.Le, s.co * s.cl * s.sa / (PI_F * s.r2)), w);
}

// tracePointLight (rayrenderer.cpp:2205-2269) up to the shadow test: the
// shadow ray toward the light (unnormalised, tmax 0.99999, the environment
// rays' occluders) and the light's value when unoccluded, color x emission x
// dotToLight / |lightray|^2 clamped to [0, 1] per channel; ok = false when
// that value is zero whatever the occlusion.
struct PointSample {
    bool ok;
    v3 o, lr, val;
};

__device__ __forceinline__ float smoothstep01(float e0, float e1, float x) {
    float t = (x - e0) / (e1 - e0);
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

// One point light seen from hit (tracePointLight up to the shadow test).
__device__ __forceinline__ PointSample point_light_at(const DevPointLight &L, v3 hit, v3 n) {
    PointSample s;
    s.ok = false;
    const v3 lr = sub(mk(L.pos[0], L.pos[1], L.pos[2]), hit);
    const v3 dir = nrm(lr);
    const float dto = dot(dir, n);
    if (!(dto > 0.0f)) return s;
    float dtl = dot(dir, mk(L.normal[0], L.normal[1], L.normal[2]));
    if (L.spot) dtl = dtl * smoothstep01(fmaxf(L.spot_cos - 0.1f, 0.0f), fminf(L.spot_cos + 0.1f, 1.0f), dtl);
    else dtl = fabsf(dto);
    if (!(dtl > 0.0f)) return s;
    const float da = powf(len(lr), -2.0f);
    const float lum = L.emission * dtl * da;
    const v3 c = mk(fminf(fmaxf(L.color[0] * lum, 0.0f), 1.0f), fminf(fmaxf(L.color[1] * lum, 0.0f), 1.0f),
                    fminf(fmaxf(L.color[2] * lum, 0.0f), 1.0f));
    if (is_zero(c)) return s;
    s.ok = true;
    s.o = surface_origin(hit, lr, n);
    s.lr = lr;
    s.val = c;
    return s;
}

// traceLight's point-light part (rayrenderer.cpp:2281-2290): the only light,
// or one picked uniformly (a rand() draw)
__device__ __forceinline__ PointSample sample_point(const PTScene &sc, Rng &R, v3 hit, v3 n) {
    if (sc.nplights == 0) {
        PointSample s;
        s.ok = false;
        return s;
    }
    const uint32_t li = sc.nplights == 1 ? 0u : pick(prng(R), sc.nplights);
    return point_light_at(sc.plights[li], hit, n);
}

// sampleVolumeLightForNEE (rayrenderer.cpp:2066-2203) up to the shadow
// test: one emissive volume picked uniformly (a rand() draw), equiangular
// sampling (Kulla & Fajardo 2012) of its cone axis, or of the owning
// object's location for constant emission; Le = sigma_e at the point times
// the medium's homogeneous transmittance over the distance; pdf = the
// solid-angle pdf times the number of volumes.  The shadow ray (unnormalised,
// tmax 0.99999) is tested against the volume lights' occluders.
struct VolumeSample {
    bool ok;
    v3 o, lr, dir, Le;
    float pdf;
};

__device__ __forceinline__ VolumeSample sample_volume(const PTScene &sc, Rng &R, v3 hit, v3 n) {
    VolumeSample s;
    s.ok = false;
    if (sc.nvlights == 0) return s;
    const DevVolumeLight ev = sc.vlights[pick(prng(R), sc.nvlights)];
    const DevMedium &m = sc.media[ev.medium];
    v3 a, b;
    if (m.src.emission_mode == 1) {
        a = ld3(m.src.cone_origin);
        b = add(a, scl(mk(m.cone_axis[0], m.cone_axis[1], m.cone_axis[2]), m.src.cone_length));
    } else {
        a = b = mk(ev.pos[0], ev.pos[1], ev.pos[2]);
    }
    const v3 sd = sub(b, a);
    const float sl = len(sd);
    const v3 axis = sl > 1e-6f ? scl(sd, 1.0f / sl) : mk(0.0f, 0.0f, -1.0f);
    const v3 toA = sub(a, hit);
    const float along = dot(toA, axis);
    const v3 perp = sub(toA, scl(axis, along));
    const float D = len(perp);
    if (!(D > 1e-5f) && sl <= 1e-6f) return s;
    const float thA = atanf(along / fmaxf(D, 1e-5f));
    const float thB = atanf((along + sl) / fmaxf(D, 1e-5f));
    if (!(fabsf(thB - thA) > 1e-6f) && sl > 1e-6f) return s;
    const float u = prng(R);
    const float th = thA + u * (thB - thA);
    const float ts = D * tanf(th);
    const v3 q = sl > 1e-6f ? add(a, scl(axis, ts - along)) : a;
    const v3 tq = sub(q, hit);
    const float r2 = tq.x * tq.x + tq.y * tq.y + tq.z * tq.z;
    const float r = sqrtf(r2);
    if (!(r > 1e-5f)) return s;
    v3 se = emission_at(m, q);
    if (is_zero(se)) return s;
    if (m.hero > 0.0f) se = mul(se, transmittance(m, r));
    const float pdf_t = D / (fmaxf(1e-6f, fabsf(thB - thA)) * (D * D + ts * ts));
    s.ok = true;
    s.o = surface_origin(hit, tq, n);
    s.lr = tq;
    s.dir = scl(tq, 1.0f / r);
    s.Le = se;
    s.pdf = pdf_t * r2 * (float)sc.nvlights;
    return s;
}

struct EnvSample {
    bool ok;
    v3 dir, Le, o;
    float pdf, co;
};

// sampleEnvmapForNEE (rayrenderer.cpp:2038-2065) up to the shadow test
__device__ __forceinline__ EnvSample sample_env_nee(const PTScene &sc, Rng &R, v3 hit, v3 n) {
    EnvSample s;
    s.ok = false;
    if (!sc.has_env || sc.env_total <= 0.0f) return s;
    const float u0 = prng(R), u1 = prng(R);
    v3 dir;
    float pdf;
    sample_env_dir(sc, u0, u1, dir, pdf);
    if (pdf <= 0.0f) return s;
    const float co = dot(dir, n);
    if (co <= 0.0f) return s;
    s.ok = true;
    s.dir = dir;
    s.pdf = pdf;
    s.Le = sample_environment(sc, dir);
    s.co = co;
    s.o = surface_origin(hit, dir, n);
    return s;
}

// ---------------------------------------------------------------------------
// path state access
// ---------------------------------------------------------------------------
struct Path {
    v3 o, d, beta, tp, L;
    float bsdf_pdf;
    int passes, medium;
    int chroma;  // chromatic dispersion band (BSDFParam::chromaChannel), -1 = none chosen
    bool alive;
    bool bg;     // eye ray through an active global medium that hit nothing
};

// The path state's medium word: the medium index (signed, low 16 bits), the
// dispersion band + 1 (bits 16-17; 0 = none chosen) and the background flag
// (bit 18: the eye ray crossed an active global medium without a hit, so a
// sample with no radiance at all reads backColor, rayrenderer.cpp:1263-1273).
__device__ __forceinline__ int pack_medium(int medium, int chroma, bool bg) {
    return (int)(((uint32_t)medium & 0xFFFFu) | ((uint32_t)(chroma + 1) << 16) | (bg ? (1u << 18) : 0u));
}
__device__ __forceinline__ int medium_of(int w) { return (int)(int16_t)(w & 0xFFFF); }
__device__ __forceinline__ int chroma_of(int w) { return (int)(((uint32_t)w >> 16) & 3u) - 1; }
__device__ __forceinline__ bool bg_of(int w) { return ((uint32_t)w >> 18) & 1u; }

__device__ __forceinline__ int crossed_medium(const PTScene &sc, const DevObject &m, bool entering, int cur) {
    if (entering) return m.interior_medium >= 0 ? m.interior_medium : cur;
    return sc.global_medium;
}

// TransparencyShader as one path step (bsdf.cpp:513-546)
__device__ __forceinline__ void step_transparency(const PTScene &sc, Path &p, uint32_t prim, v3 hit, v3 n, float beta_scale,
                                  bool fresh_param) {
    const DevObject &m = obj_of(sc, prim);
    p.beta = scl(p.beta, beta_scale);
    p.tp = scl(p.tp, m.transparency);
    p.medium = crossed_medium(sc, m, dot(p.d, n) < 0.0f, p.medium);
    const v3 gN = geom_normal(sc, prim, n);
    p.o = surface_origin(hit, p.d, gN);
    if (fresh_param) {  // a fresh BSDFParam: no MIS pdf, no dispersion band
        p.bsdf_pdf = 0.0f;
        p.chroma = -1;
    }
}

// MixShader (bsdf.cpp:579-610) with one lobe drawn in proportion to its
// weight; tex = the albedo texture sample at the hit (1 when untextured)
// kExt: the scene uses a heterogeneous density, cone or path emission, heat
// haze or emissive-volume NEE; the kernels of other scenes are compiled
// without that code (it costs the common path registers).
template <bool kExt>
__device__ __forceinline__ void step_mix(const PTScene &sc, Rng &R, Path &p, Shadows &S, uint32_t prim, v3 hit, v3 n,
                                         v3 tex) {
    const DevObject &m = obj_of(sc, prim);
    const float wd = 1.0f - m.glossy - m.refraction;
    const float w0 = wd > 0.00001f ? wd : 0.0f;
    const float w1 = m.glossy > 0.00001f ? m.glossy : 0.0f;
    const float w2 = m.refraction > 0.00001f ? m.refraction : 0.0f;
    const float W = w0 + w1 + w2;
    if (!(W > 0.0f)) {
        p.alive = false;
        return;
    }
    const float u = prng(R) * W;
    int lobe = (u < w0) ? 0 : (u < w0 + w1 ? 1 : 2);
    if (lobe == 2 && w2 == 0.0f) lobe = w1 > 0.0f ? 1 : 0;
    if (lobe == 1 && w1 == 0.0f) lobe = 0;
    const float wl = lobe == 0 ? w0 : (lobe == 1 ? w1 : w2);
    p.beta = scl(p.beta, W);
    p.tp = scl(p.tp, wl);
    if (lobe == 0) {
        // DiffuseShader (bsdf.cpp:54-108): albedo = colour x texture, 1
        // without colour sampling
        const v3 alb = sc.color_sampling ? mul(ld3(m.color), tex) : mk(1.0f, 1.0f, 1.0f);
        const float r1 = lds1(R), r2 = lds1(R);
        const v3 dir = cosine_dir(n, r1, r2);
        const float pdf = fmaxf(0.0f, dot(dir, n)) * (float)(1.0 / M_PI);
        const v3 ba = mul(p.beta, alb);
        // traceLight (area, then point lights), then traceEnvmapLight
        const AreaSample as = sample_area(sc, R, hit, n);
        if (as.ok) push_shadow(S, as.o, as.lr, 0.99999f, RG_PRED_AREA, mul(ba, area_value(as)));
        const PointSample ps = sample_point(sc, R, hit, n);
        if (ps.ok) push_shadow(S, ps.o, ps.lr, 0.99999f, RG_PRED_ENV, mul(ba, ps.val));
        const EnvSample es = sample_env_nee(sc, R, hit, n);
        if (es.ok) {
            float w = 1.0f;
            if (pdf > 0.0f) {
                const float e2 = es.pdf * es.pdf, b2 = pdf * pdf;
                w = e2 / (e2 + b2);
            }
            push_shadow(S, es.o, es.dir, 1e30f, RG_PRED_ENV, mul(ba, scl(es.Le, es.co * w / (PI_F * es.pdf))));
        }
        if constexpr (kExt) {
            // emissive-volume NEE (bsdf.cpp:92-103), no MIS
            const VolumeSample vs = sample_volume(sc, R, hit, n);
            if (vs.ok && vs.pdf > 0.0f) {
                const float cl = fmaxf(0.0f, dot(vs.dir, n));
                push_shadow(S, vs.o, vs.lr, 0.99999f, RG_PRED_VOLUME,
                            mul(ba, scl(vs.Le, cl * (float)(1.0 / M_PI) / vs.pdf)));
            }
        }
        p.beta = ba;
        p.tp = mul(p.tp, alb);
        p.bsdf_pdf = pdf;
        p.o = surface_origin(hit, dir, geom_normal(sc, prim, n));
        p.d = dir;
        return;
    }
    if (lobe == 1) {
        // GlossyShader (bsdf.cpp:173-374): albedo = colour x texture drives
        // F0 (the colour alone without colour sampling)
        const v3 alb = sc.color_sampling ? mul(ld3(m.color), tex) : ld3(m.color);
        if (m.roughness < 1e-3f) {
            const v3 F0 = (m.metallic >= 1.0f) ? alb
                                               : add(scl(alb, m.metallic), scl(mk(0.04f, 0.04f, 0.04f), 1.0f - m.metallic));
            const v3 gN = geom_normal(sc, prim, n);
            v3 r = reflect(p.d, n);
            if (dot(r, gN) <= 0.0f) r = reflect(p.d, gN);
            const float ci = fmaxf(0.0f, -dot(p.d, n));
            const v3 F = schlick_f(F0, ci);
            p.beta = mul(p.beta, F);
            p.tp = mul(p.tp, F);
            p.bsdf_pdf = 0.0f;
            p.o = surface_origin(hit, r, gN);
            p.d = r;
            return;
        }
        Ggx G;
        if (!ggx_setup(m, alb, n, p.d, G)) {
            p.alive = false;
            return;
        }
        {
            const AreaSample as = sample_area(sc, R, hit, n);
            v3 fc;
            float pb;
            if (as.ok && eval_glossy(G, as.dir, fc, pb)) {
                const float pl2 = as.pdf * as.pdf, pb2 = pb * pb;
                push_shadow(S, as.o, as.lr, 0.99999f, RG_PRED_AREA,
                            mul(p.beta, scl(mul(as.Le, fc), (pl2 / (pl2 + pb2)) / as.pdf)));
            }
        }
        {
            const EnvSample es = sample_env_nee(sc, R, hit, n);
            v3 fc;
            float pb;
            if (es.ok && eval_glossy(G, es.dir, fc, pb)) {
                const float pe2 = es.pdf * es.pdf, pb2 = pb * pb;
                push_shadow(S, es.o, es.dir, 1e30f, RG_PRED_ENV,
                            mul(p.beta, scl(mul(es.Le, fc), (pe2 / (pe2 + pb2)) / es.pdf)));
            }
        }
        if constexpr (kExt) {
            // emissive-volume NEE (bsdf.cpp:318-329), no MIS
            const VolumeSample vs = sample_volume(sc, R, hit, n);
            v3 fc;
            float pb;
            if (vs.ok && vs.pdf > 0.0f && eval_glossy(G, vs.dir, fc, pb))
                push_shadow(S, vs.o, vs.lr, 0.99999f, RG_PRED_VOLUME, mul(p.beta, scl(mul(vs.Le, fc), 1.0f / vs.pdf)));
        }
        const float u1 = prng(R), u2 = prng(R);
        const v3 H = sample_vndf(G.V, G.ax, G.ay, u1, u2);
        const v3 Ll = sub(scl(H, 2.0f * dot(G.V, H)), G.V);
        if (Ll.z <= 0.0f) {
            p.alive = false;
            return;
        }
        const v3 L = add(add(scl(G.t, Ll.x), scl(G.b, Ll.y)), scl(n, Ll.z));
        const v3 gN = geom_normal(sc, prim, n);
        if (dot(L, gN) <= 0.0f) {
            p.alive = false;
            return;
        }
        const float vdh = fmaxf(0.0f, dot(G.V, H));
        const v3 F = schlick_f(G.F0, vdh);
        const v3 wgt = scl(F, smith_g1(Ll, G.ax, G.ay));
        p.beta = mul(p.beta, wgt);
        p.tp = mul(p.tp, wgt);
        p.bsdf_pdf = G.g1v * ggx_d(H, G.ax, G.ay) / (4.0f * G.V.z);
        p.o = surface_origin(hit, L, gN);
        p.d = L;
        return;
    }
    // RefractionShader (bsdf.cpp:376-483).  Chromatic dispersion
    // (bsdf.cpp:394-420): the path's first dispersive interface draws its
    // band, kept for every later one; the band's index of refraction is
    // refraction_ratio * (1 - s), * 1, * (1 + s), and the estimator weight
    // is masked to 3 in that band.
    const float ct = fabsf(dot(p.d, n));
    float ior = m.refraction_ratio;
    v3 mask = mk(1.0f, 1.0f, 1.0f);
    if (m.chroma_dispersion > 0.0f) {
        bool first = false;
        if (p.chroma < 0) {
            p.chroma = (int)fminf(2.0f, floorf(prng(R) * 3.0f));
            first = true;
        }
        const float s = m.chroma_dispersion;
        if (p.chroma == 0) ior = m.refraction_ratio * (1.0f - s);
        else if (p.chroma == 2) ior = m.refraction_ratio * (1.0f + s);
        if (first) mask = mk(p.chroma == 0 ? 3.0f : 0.0f, p.chroma == 1 ? 3.0f : 0.0f, p.chroma == 2 ? 3.0f : 0.0f);
    }
    const float fr = fresnel_schlick(ct, ior);
    const bool refl = prng(R) < fr;
    v3 dir = refl ? reflect(p.d, n) : refract_dir(p.d, n, ior);
    if (m.roughness > 0.0f) {
        const float u1 = prng(R), u2 = prng(R);
        dir = nrm(add(dir, scl(hemisphere_dir(n, u1, u2), m.roughness)));
    }
    const v3 gN = geom_normal(sc, prim, n);
    if (refl && dot(dir, gN) <= 0.0f) dir = reflect(p.d, gN);
    const v3 col = ld3(m.color);
    p.beta = mul(mul(p.beta, col), mask);
    p.tp = mul(p.tp, col);
    p.bsdf_pdf = 0.0f;
    const float di = dot(p.d, n), dout = dot(dir, n);
    if (di * dout > 0.0f) p.medium = crossed_medium(sc, m, di < 0.0f, p.medium);
    p.o = surface_origin(hit, dir, gN);
    p.d = dir;
}

// RayBSDFShaderProvider::shade (rayrenderer.cpp:2776-2896) as one path event
template <bool kExt>
__device__ __forceinline__ void step_surface(const PTScene &sc, Rng &R, Path &p, Shadows &S, uint32_t prim, v3 hit, v3 n,
                                             v3 tex) {
    const DevObject &m = obj_of(sc, prim);
    const bool eye = p.passes < 0;
    if (m.emission > 0.0f) {
        const v3 E = scl(ld3(m.color), m.emission);
        if (!eye && p.bsdf_pdf > 0.0f) {
            const float cl = -dot(p.d, n);
            if (cl > 0.0f) {
                const v3 dd = sub(hit, p.o);
                const float r2 = dd.x * dd.x + dd.y * dd.y + dd.z * dd.z;
                const float pl = r2 / (cl * sc.shade[prim].sampled_area);
                const float pb = p.bsdf_pdf;
                const float pb2 = pb * pb, pl2 = pl * pl;
                p.L = add(p.L, mul(p.beta, scl(E, pb2 / (pb2 + pl2))));
            }
        } else {
            p.L = add(p.L, mul(p.beta, E));
        }
        p.alive = false;
        return;
    }
    if (dot(p.d, n) > 0.0f) {
        if (m.transparency > 0.001f) {
            if (!eye && p.passes + 1 <= sc.max_depth) {
                p.passes = p.passes + 1;
                step_transparency(sc, p, prim, hit, n, m.transparency, true);
                return;
            }
            p.alive = false;
            return;
        } else if (m.refraction < 0.001f && m.glossy > 0.001f) {
            p.alive = false;
            return;
        }
    }
    if (!eye) {
        if (p.passes + 1 >= sc.max_depth) {
            // the light term times colour x texture, zero without colour
            // sampling (rayrenderer.cpp:2839-2856)
            if (1.0f - m.glossy - m.refraction > 0.00001f && sc.color_sampling) {
                const v3 bc = mul(p.beta, mul(ld3(m.color), tex));
                const AreaSample as = sample_area(sc, R, hit, n);
                if (as.ok) push_shadow(S, as.o, as.lr, 0.99999f, RG_PRED_AREA, mul(bc, area_value(as)));
                const PointSample ps = sample_point(sc, R, hit, n);
                if (ps.ok) push_shadow(S, ps.o, ps.lr, 0.99999f, RG_PRED_ENV, mul(bc, ps.val));
            }
            p.alive = false;
            return;
        }
        p.passes = p.passes + 1;
        if (p.passes >= MIN_RR_DEPTH) {
            float q = maxc(p.tp);
            q = fminf(RR_MAX_PROB, fmaxf(RR_MIN_PROB, q));
            if (prng(R) >= q) {
                p.alive = false;
                return;
            }
            p.beta = scl(p.beta, 1.0f / q);
        }
        if (m.transparency > 0.001f) {
            step_transparency(sc, p, prim, hit, n, m.transparency, false);
            return;
        }
        step_mix<kExt>(sc, R, p, S, prim, hit, n, tex);
        return;
    }
    p.passes = 0;
    if (m.transparency > 0.01f) {
        if (prng(R) < m.transparency) {
            step_transparency(sc, p, prim, hit, n, 1.0f, true);
            // the reference evaluates the Mix branch first on the same
            // BSDFParam (rayrenderer.cpp:2889-2894); its refraction lobe
            // leaves the dispersion band it drew (bsdf.cpp:402-407), which
            // the transparency path then carries without the 3x mask
            if (m.refraction > 0.00001f && m.chroma_dispersion > 0.0f)
                p.chroma = (int)fminf(2.0f, floorf(prng(R) * 3.0f));
            return;
        }
    }
    step_mix<kExt>(sc, R, p, S, prim, hit, n, tex);
}

// RayRenderer::tracePath (rayrenderer.cpp:1596-1864) as one path event
template <bool kExt>
__device__ __forceinline__ void step(const PTScene &sc, Rng &R, Path &p, Shadows &S, const rg_hit &h) {
    const bool eye = p.passes < 0;
    const bool hit = h.prim != 0xFFFFFFFFu;
    const v3 pos = add(p.o, scl(p.d, h.t));
    const DevMedium *med = (p.medium >= 0 && sc.media[p.medium].active) ? &sc.media[p.medium] : nullptr;
    if (!med) {
        if (hit && (!eye || obj_of(sc, h.prim).visible)) {
            step_surface<kExt>(sc, R, p, S, h.prim, pos, interp_normal(sc, h.prim, h.u, h.v),
                         albedo_tex(sc, h.prim, h.u, h.v));
            return;
        }
        v3 env = sample_environment(sc, p.d);
        if (eye) {
            if (is_zero(env)) env = mk(sc.back[0], sc.back[1], sc.back[2]);
        } else {
            env = scl(env, env_miss_weight(sc, p.d, p.bsdf_pdf));
        }
        p.L = add(p.L, mul(p.beta, env));
        p.alive = false;
        return;
    }
    if (kExt && med->haze) {
        // heat haze (rayrenderer.cpp:1645-1663): the ray bends through the
        // volume and continues in the outer medium as the path's next event,
        // with the same passes, throughput and BSDF pdf (a heat-haze global
        // medium is refused, so the outer one is never heat haze)
        const Ray3 b = bend_ray(*med, p.o, p.d, fminf(hit ? h.t : RAY_MAX_DISTANCE, 1000.0f));
        p.o = b.o;
        p.d = b.d;
        p.medium = sc.global_medium;
        return;
    }
    if (eye && !hit) p.bg = true;
    const float maxT = hit ? h.t : RAY_MAX_DISTANCE;
    bool scattered;
    float tf = 0.0f, den = 1.0f;
    if (kExt && med->het) {
        // delta tracking through the fBm density (rayrenderer.cpp:1677-1686)
        scattered = delta_track(*med, R, p.o, p.d, maxT, tf, den);
        if (!scattered) tf = maxT;
    } else {
        tf = free_flight(*med, prng(R));
        scattered = tf < maxT;
    }
    const float elen = scattered ? tf : maxT;
    p.L = add(p.L, mul(p.beta, kExt ? emission_integral(*med, R, p.o, p.d, elen) : emission_along(*med, elen)));
    if (scattered) {
        v3 sw = mk(0, 0, 0);
        if (kExt && med->het) {
            const float inv = med->t_max_hero > 0.0f ? 1.0f / med->t_max_hero : 0.0f;
            sw = mk(med->sigma_s[0] * den * inv, med->sigma_s[1] * den * inv, med->sigma_s[2] * den * inv);
        } else {
            const float ph = med->hero * expf(-med->hero * tf);
            const v3 Tr = transmittance(*med, tf);
            if (ph > 0.0f)
                sw = mk(med->sigma_s[0] * Tr.x / ph, med->sigma_s[1] * Tr.y / ph, med->sigma_s[2] * Tr.z / ph);
        }
        if (is_zero(sw)) {
            p.alive = false;
            return;
        }
        const v3 spos = add(p.o, scl(p.d, tf));
        const int passes = eye ? 1 : p.passes + 1;
        if (passes > sc.max_depth) {
            p.alive = false;
            return;
        }
        const v3 bsw = mul(p.beta, sw);
        const v3 proxy = neg(p.d);
        {
            const AreaSample as = sample_area(sc, R, spos, proxy);
            if (as.ok && as.pdf > 0.0f)
                push_shadow(S, as.o, as.lr, 0.99999f, RG_PRED_AREA,
                            mul(bsw, scl(as.Le, phase_pdf(*med, neg(p.d), as.dir) * (1.0f / as.pdf))));
        }
        {
            const EnvSample es = sample_env_nee(sc, R, spos, proxy);
            if (es.ok && es.pdf > 0.0f)
                push_shadow(S, es.o, es.dir, 1e30f, RG_PRED_ENV,
                            mul(bsw, scl(es.Le, phase_pdf(*med, neg(p.d), es.dir) * (1.0f / es.pdf))));
        }
        if constexpr (kExt) {
            const VolumeSample vs = sample_volume(sc, R, spos, proxy);
            if (vs.ok && vs.pdf > 0.0f)
                push_shadow(S, vs.o, vs.lr, 0.99999f, RG_PRED_VOLUME,
                            mul(bsw, scl(vs.Le, phase_pdf(*med, neg(p.d), vs.dir) * (1.0f / vs.pdf))));
        }
        const float u1 = prng(R), u2 = prng(R);
        const v3 nd = sample_phase(*med, p.d, u1, u2);
        float rrw = 1.0f;
        if (passes >= MIN_RR_DEPTH) {
            float q = maxc(mul(p.tp, sw));
            q = fminf(RR_MAX_PROB, fmaxf(RR_MIN_PROB, q));
            if (prng(R) >= q) {
                p.alive = false;
                return;
            }
            rrw = 1.0f / q;
        }
        p.beta = scl(bsw, rrw);
        p.passes = passes;
        p.bsdf_pdf = 0.0f;
        p.o = add(spos, scl(nrm(nd), SURFACE_THICKNESS));
        p.d = nd;
        return;
    }
    v3 sw = mk(1, 1, 1);
    if (!(kExt && med->het)) {
        const float surv = expf(-med->hero * maxT);
        if (surv > 0.0f) {
            const v3 Tr = transmittance(*med, maxT);
            sw = mk(Tr.x / surv, Tr.y / surv, Tr.z / surv);
        }
    }
    p.beta = mul(p.beta, sw);
    if (hit) {
        v3 n = interp_normal(sc, h.prim, h.u, h.v);
        const v3 g = ld3(sc.shade[h.prim].gn);
        const v3 gN = dot(g, p.d) <= 0.0f ? g : neg(g);
        if (dot(p.d, n) > 0.0f && dot(p.d, gN) < 0.0f) n = gN;
        step_surface<kExt>(sc, R, p, S, h.prim, pos, n, albedo_tex(sc, h.prim, h.u, h.v));
        return;
    }
    v3 env = sample_environment(sc, p.d);
    if (!eye) env = scl(env, env_miss_weight(sc, p.d, p.bsdf_pdf));
    p.L = add(p.L, mul(p.beta, env));
    p.alive = false;
}


// Primary ray through pixel (x, y) (rayrenderer.cpp:860-934): antialias
// jitter, then thin-lens depth of field with a disc or polygonal aperture.
__device__ __forceinline__ void camera_ray(const PTScene &sc, Rng &R, uint32_t x, uint32_t y, v3 &o, v3 &d) {
    const float dx = ((float)x + 0.5f - sc.half_w) * sc.view_scale;
    const float dy = -((float)y + 0.5f - sc.half_h) * sc.view_scale;
    const float jx = lds1(R), jy = lds1(R);
    const float jxo = sc.antialias ? (jx - 0.5f) : 0.0f;
    const float jyo = sc.antialias ? (jy - 0.5f) : 0.0f;
    if (sc.dof >= 0.001f && sc.aperture > 0.0f) {
        const float px = dx + jxo * sc.view_scale;
        const float py = dy - jyo * sc.view_scale;
        const v3 focal = mk(px * sc.dof, py * sc.dof, -sc.dof);
        const float du = lds1(R), dv = lds1(R);
        float ox, oy;
        if (sc.blades >= 3) {
            const int N = sc.blades;
            const float u0 = du * (float)N;
            const int wedge = (int)fminf((float)(N - 1), floorf(u0));
            float s2 = u0 - (float)wedge, t2 = dv;
            if (s2 + t2 > 1.0f) {
                s2 = 1.0f - s2;
                t2 = 1.0f - t2;
            }
            const float stp = 2.0f * PI_F / (float)N;
            const float a0 = sc.ap_rot + stp * (float)wedge;
            const float a1 = a0 + stp;
            ox = (cosf(a0) * s2 + cosf(a1) * t2) * sc.aperture;
            oy = (sinf(a0) * s2 + sinf(a1) * t2) * sc.aperture;
        } else {
            const float a = 2.0f * du - 1.0f, b = 2.0f * dv - 1.0f;
            float r, phi;
            if (a == 0.0f && b == 0.0f) {
                r = 0.0f;
                phi = 0.0f;
            } else if (a * a > b * b) {
                r = a;
                phi = (float)(M_PI / 4.0) * (b / a);
            } else {
                r = b;
                phi = (float)(M_PI / 2.0) - (float)(M_PI / 4.0) * (a / b);
            }
            ox = r * cosf(phi) * sc.aperture;
            oy = r * sinf(phi) * sc.aperture;
        }
        o = mk(ox, oy, 0.0f);
        d = nrm(sub(focal, o));
    } else {
        o = mk(0, 0, 0);
        d = nrm(mk(dx + jxo * sc.view_scale, dy - jyo * sc.view_scale, -1.0f));
    }
}

constexpr int kPtBlock = 256;

// Position in a queue for every active lane with pred set, in lane order
// (0xFFFFFFFF for the others): one atomic per wave.
__device__ __forceinline__ uint32_t wave_slot(bool pred, uint32_t *counter) {
    const uint64_t m = __ballot(pred ? 1 : 0);
    if (m == 0) return 0xFFFFFFFFu;
    const uint32_t lane = __lane_id();
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, (int)leader);
    return pred ? base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull)) : 0xFFFFFFFFu;
}

// A path's radiance with the unoccluded NEE contributions of its last event
// added in event order (the shadow rays' results arrive after the event that
// cast them: the next event of a live path adds them first, the accumulation
// adds them for a finished one).
__device__ __forceinline__ v3 with_last_nee(const PTState &st, uint32_t p) {
    float4 L = st.L[p];
    const uint32_t ns = st.nshadow[p];
    for (uint32_t k = 0; k < ns; k++) {
        if (!st.socc[k * st.slot_stride + p]) {
            const float4 c = st.scontrib[k * st.slot_stride + p];
            L.x = L.x + c.x;
            L.y = L.y + c.y;
            L.z = L.z + c.z;
        }
    }
    return mk(L.x, L.y, L.z);
}

// A finished sample's value: radiance clamped at 0 and, when clamp_max > 0,
// at clamp_max (the reference's firefly clamp, rayrenderer.cpp:985-988).  An
// eye ray that crossed an active global medium without a hit and gathered no
// radiance reads backColor (traceEyeRay, rayrenderer.cpp:1263-1273).
__device__ __forceinline__ float4 sample_value(const PTState &st, uint32_t p, float clamp_max, const float *back) {
    v3 L = with_last_nee(st, p);
    if (bg_of(st.meta[p].y) && is_zero(L)) L = mk(back[0], back[1], back[2]);
    float4 c = make_float4(fmaxf(L.x, 0.0f), fmaxf(L.y, 0.0f), fmaxf(L.z, 0.0f), 1.0f);
    if (clamp_max > 0.0f) {
        c.x = fminf(c.x, clamp_max);
        c.y = fminf(c.y, clamp_max);
        c.z = fminf(c.z, clamp_max);
    }
    return c;
}

// Position in a queue for every thread of the block with pred set, in
// thread order (0xFFFFFFFF for the others): one atomic per block.  Every
// thread of the block must call it; `lds` holds kPtBlock / 64 + 1 words.
__device__ __forceinline__ uint32_t block_slot(bool pred, uint32_t *counter, uint32_t *lds) {
    constexpr uint32_t kWaves = kPtBlock / 64;
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    const uint64_t m = __ballot(pred ? 1 : 0);
    if (lane == 0) lds[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (uint32_t w = 0; w < kWaves; w++) {
            const uint32_t c = lds[w];
            lds[w] = tot;
            tot += c;
        }
        lds[kWaves] = tot ? atomicAdd(counter, tot) : 0u;
    }
    __syncthreads();
    const uint32_t pos = lds[kWaves] + lds[wave] + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    __syncthreads();  // the words are reused by the next call
    return pred ? pos : 0xFFFFFFFFu;
}

// Appends `value` for every lane with pred set: one atomic per wave.
__device__ __forceinline__ void wave_append(bool pred, uint32_t value, uint32_t *queue, uint32_t *counter) {
    const uint32_t pos = wave_slot(pred, counter);
    if (pred) queue[pos] = value;
}

}  // namespace rg

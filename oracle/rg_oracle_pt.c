/*
 * rg_oracle_pt.c — CPU restatement of RayGen's BSDF path tracer, shading
 * system 5 (TEST INFRASTRUCTURE: the checker, never the product).
 *
 * Two estimators over the same scene and the same random-number sources:
 *
 *  mode 0 "recursive": restates the reference's recursion call for call —
 *    RayRenderer::traceEyeRay / tracePath / RayBSDFShaderProvider::shade
 *    (rayrenderer.cpp:1254-1297, 1596-1864, 2776-2896), MixShader evaluating
 *    every lobe (bsdf.cpp:579-610), Diffuse/Glossy/Refraction/Transparency
 *    shaders (bsdf.cpp:54-546), area-light and envmap NEE with power-heuristic
 *    MIS (rayrenderer.cpp:1916-2065, 2271-2293, 1484-1530), homogeneous media
 *    (rayrenderer.cpp:1645-1864; medium.cpp:18-123, 475-525).
 *
 *  mode 1 "single-lobe": the same estimator with one stochastically chosen
 *    lobe per vertex (probability proportional to the Mix weights) — the
 *    iterative form the GPU wavefront kernels run.  Its random draws are
 *    consumed in exactly the order the GPU consumes them, so the two agree
 *    per pixel up to the last-ulp differences of transcendental functions.
 *
 * Random numbers: ldsNext1D (Halton + per-pixel Cranley-Patterson,
 * raycommon.cpp:21-92) for the reference's LDS call sites (camera jitter,
 * cosine directions, light points); every randomValue()/rand() call site of
 * the reference draws from Philox4x32-10 keyed by (seed, job) with counter
 * (pixel, sample, draw/4) — the reference's own PRNG is time-seeded and not
 * reproducible, so image parity with it is statistical.
 *
 * ugm routines that are not vendored with the reference and are therefore
 * parity unpinned: vec3::normalize (taken as v * (1/|v|)),
 * randomRayInHemisphere (taken as a uniform hemisphere direction),
 * DEGREE_TO_RADIAN (binary32), and for point lights RADIAN_TO_DEGREE (x *
 * 180 / pi), smoothstep (GLSL) and color3 clamp (per channel to [0, 1]).
 * Cubemaps, heterogeneous / cone / path media and heat haze are outside
 * this path.
 */
#include "rg_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#define PI_F 3.14159265358979323846f
#define SURFACE_THICKNESS 0.0001f
#define MAX_RAY_DISTANCE 1.0e6f
#define RAY_MAX_DISTANCE 100.0f
#define MIN_RR_DEPTH 3
#define RR_MIN_PROB 0.05f
#define RR_MAX_PROB 0.95f

/* ------------------------------------------------------------------------ */
typedef struct { float x, y, z; } v3;

static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 scl(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float len(v3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
static inline v3 nrm(v3 a) { float inv = 1.0f / len(a); return scl(a, inv); }
static inline v3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }
static inline float maxc(v3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
static inline int is_zero(v3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
static inline v3 reflect(v3 d, v3 n) { return sub(d, scl(n, dot(d, n) * 2.0f)); }

/* ------------------------------------------------------------------------ */
/* prepared scene                                                           */
/* ------------------------------------------------------------------------ */
typedef struct {
    v3 sigma_t, sigma_s_eff, sigma_e_eff;
    float hero, g;
    int active;
    const or_medium *src;   /* authored fields (noise, cone, path, haze) */
    int het;                /* fBm density field */
    float t_max_hero, peak; /* delta-tracking bound sigma_t_hero * peak */
    v3 cone_axis;           /* normalised (render space) */
    float density;          /* authored density (scales cone / path emission) */
    int haze;               /* isHeatHaze */
} medium_t;

typedef struct {
    const or_scene_desc *d;
    or_node *nodes;
    int64_t nnodes;
    uint32_t *perm;
    v3 *gn;           /* ti.normalizedpd */
    float *ti_a;      /* ti.a */
    float *area;
    float *sampled_area; /* triCount(object) * area, for emitters */
    /* lights: objects with emission > 0 and at least one triangle */
    uint32_t nlights;
    uint32_t *light_obj;
    uint32_t *light_first; /* into light_tris */
    uint32_t *light_count;
    uint32_t *light_tris;
    medium_t *media;
    /* envmap */
    int has_env;
    float *env_marg;  /* H+1 */
    float *env_cond;  /* H*(W+1) */
    float env_total;
    /* camera */
    float view_scale, half_w, half_h, aperture, dof, ap_rot;
    int blades;
    int max_depth;
    /* enableColorSampling, backColor (rgb) */
    int color_sampling;
    v3 back;
    /* shaderProvider, worldColor, calcAO's sample count */
    int shader;
    v3 world;
    int ao_samples;
    /* linear RGB texels of each texture (sRGB decoded once) */
    float **tex_lin;
} prep_t;

/* ------------------------------------------------------------------------ */
/* random numbers                                                            */
/* ------------------------------------------------------------------------ */
typedef struct {
    int x, y, sample, dim;
    uint32_t pix;
    uint32_t key0, key1;
    uint32_t k;          /* next Philox draw */
    uint32_t blk[4];
    uint32_t blk_id;
} rng_t;

static float prng(rng_t *r) {
    const uint32_t b = r->k >> 2;
    if (b != r->blk_id) {
        const uint32_t ctr[4] = {r->pix, (uint32_t)r->sample, b, 0u};
        const uint32_t key[2] = {r->key0, r->key1};
        or_philox4x32_10(ctr, key, r->blk);
        r->blk_id = b;
    }
    const uint32_t w = r->blk[r->k & 3];
    r->k++;
    return (float)(w >> 8) * (1.0f / 16777216.0f);
}

static float lds1(rng_t *r) {
    if (r->dim >= 16) return prng(r);
    const float v = or_lds_sample(r->x, r->y, r->sample, r->dim);
    r->dim++;
    return v;
}

static inline uint32_t pick(float u, uint32_t n) {
    uint32_t i = (uint32_t)(u * (float)n);
    return i < n ? i : n - 1;
}

/* ------------------------------------------------------------------------ */
/* geometry helpers                                                          */
/* ------------------------------------------------------------------------ */
static inline v3 tri_v(const prep_t *P, uint32_t t, int k) { return ld3(P->d->v9 + 9 * (size_t)t + 3 * k); }
static inline v3 tri_n(const prep_t *P, uint32_t t, int k) { return ld3(P->d->n9 + 9 * (size_t)t + 3 * k); }
static inline const or_object *obj_of(const prep_t *P, uint32_t t) { return &P->d->objs[P->d->obj_of_tri[t]]; }

/* geomNormal (bsdf.cpp:27-30) */
static inline v3 geom_normal(const prep_t *P, uint32_t t, v3 shading_n) {
    v3 g = P->gn[t];
    return dot(g, shading_n) >= 0.0f ? g : neg(g);
}

/* SurfaceRay (raycommon.h:82-85): returns the offset origin */
static inline v3 surface_origin(v3 hit, v3 dir, v3 n) {
    const float side = (dot(dir, n) >= 0.0f) ? 1.0f : -1.0f;
    return add(hit, scl(n, side * SURFACE_THICKNESS));
}

typedef struct {
    int hit;
    uint32_t prim;
    float t, u, v;
    v3 pos;
} isect_t;

/* findNearestTriangle: closest hit with info.t = 1e6 (raycommon.h:174) */
static isect_t nearest(const prep_t *P, v3 o, v3 d) {
    or_ray r = {{o.x, o.y, o.z}, 0.0f, {d.x, d.y, d.z}, MAX_RAY_DISTANCE};
    or_hit h;
    or_closest1(P->nodes, P->nnodes, P->d->v9, P->perm, &r, &h);
    isect_t s;
    s.hit = h.prim != 0xFFFFFFFFu;
    s.prim = h.prim;
    s.t = h.t;
    s.u = h.u;
    s.v = h.v;
    s.pos = add(o, scl(d, h.t));
    return s;
}

static int occluded(const prep_t *P, v3 o, v3 d, float maxt, int pred) {
    or_ray r = {{o.x, o.y, o.z}, 0.0f, {d.x, d.y, d.z}, maxt};
    uint8_t mask_pred = (uint8_t)pred;
    return or_any1_pred(P->nodes, P->nnodes, P->d->v9, P->perm, &r, P->d->obj_of_tri,
                        P->d->objs, mask_pred);
}

/* calcVertexInterpolation(info) (rayrenderer.cpp:2354-2362): normal only */
static v3 interp_normal(const prep_t *P, const isect_t *h) {
    const float w = 1.0f - h->u - h->v;
    v3 n = add(add(scl(tri_n(P, h->prim, 0), w), scl(tri_n(P, h->prim, 1), h->u)),
               scl(tri_n(P, h->prim, 2), h->v));
    return nrm(n);
}

/* Texture::sample (texture.cpp:209-235) of the hit object's albedo texture
 * at the barycentric UV (calcVertexInterpolation, rayrenderer.cpp:2354-2362)
 * times texTiling; (1, 1, 1) when the object has no texture. */
static v3 albedo_tex(const prep_t *P, const isect_t *h) {
    const or_scene_desc *d = P->d;
    if (!d->ntex || !d->tex_of_obj) return mk(1.0f, 1.0f, 1.0f);
    const uint32_t o = d->obj_of_tri[h->prim];
    const int32_t t = d->tex_of_obj[o];
    if (t < 0) return mk(1.0f, 1.0f, 1.0f);
    const float *q = d->uv6 + 6 * (size_t)h->prim;
    const float w = 1.0f - h->u - h->v;
    const float su = (q[0] * w + q[2] * h->u + q[4] * h->v) * d->tiling[2 * o];
    const float sv = (q[1] * w + q[3] * h->u + q[5] * h->v) * d->tiling[2 * o + 1];
    if (!(su == su) || !(sv == sv)) return mk(0.0f, 0.0f, 0.0f);
    float fu = su - floorf(su), fv = sv - floorf(sv);
    if (fu < 0.0f || fu >= 1.0f) fu = 0.0f;
    if (fv < 0.0f || fv >= 1.0f) fv = 0.0f;
    const int W = (int)d->tex[t].w, H = (int)d->tex[t].h;
    int x = (int)(fu * (float)W), y = (int)(fv * (float)H);
    if (x >= W) x = W - 1;
    if (y >= H) y = H - 1;
    if (x < 0) x = 0;
    if (y < 0) y = 0;
    return ld3(P->tex_lin[t] + 3 * ((size_t)y * W + x));
}

/* srgbToLinear (texture.cpp:203-206) */
static float srgb_to_linear(float c) {
    if (c <= 0.04045f) return c * (1.0f / 12.92f);
    return powf((c + 0.055f) * (1.0f / 1.055f), 2.4f);
}

/* calcVertexInterpolation(rt, hit) (rayrenderer.cpp:2337-2352): area weights */
static v3 interp_normal_at(const prep_t *P, uint32_t t, v3 p) {
    const v3 f1 = sub(tri_v(P, t, 0), p), f2 = sub(tri_v(P, t, 1), p), f3 = sub(tri_v(P, t, 2), p);
    const float a1 = len(cross(f2, f3)) * P->ti_a[t];
    const float a2 = len(cross(f3, f1)) * P->ti_a[t];
    const float a3 = len(cross(f1, f2)) * P->ti_a[t];
    return add(add(scl(tri_n(P, t, 0), a1), scl(tri_n(P, t, 1), a2)), scl(tri_n(P, t, 2), a3));
}

/* ------------------------------------------------------------------------ */
/* environment (rayrenderer.cpp:1299-1573; scene.cpp:330-463)                */
/* ------------------------------------------------------------------------ */
static v3 env_texel(const prep_t *P, int x, int y) {
    const float *p = P->d->env.rgb + 3 * ((size_t)y * P->d->env.w + (size_t)x);
    return mk(p[0], p[1], p[2]);
}

static v3 sample_environment(const prep_t *P, v3 dir) {
    if (!P->has_env) return mk(0, 0, 0);
    const v3 d = nrm(dir);
    const float yaw = P->d->env.rotation_deg * (float)(M_PI / 180.0);
    const float cy = cosf(yaw), sy = sinf(yaw);
    const float rx = d.x * cy - d.z * sy;
    const float rz = d.x * sy + d.z * cy;
    const float ry = d.y;
    const float u = 0.5f + atan2f(rz, rx) * (float)(0.5 / M_PI);
    const float v = 0.5f - asinf(fminf(fmaxf(ry, -1.0f), 1.0f)) * (float)(1.0 / M_PI);
    /* Texture::sample, nearest (texture.cpp:209-241) */
    const int W = (int)P->d->env.w, H = (int)P->d->env.h;
    if (!(u == u) || !(v == v)) return mk(0, 0, 0);
    float fu = u - floorf(u), fv = v - floorf(v);
    if (fu < 0.0f || fu >= 1.0f) fu = 0.0f;
    if (fv < 0.0f || fv >= 1.0f) fv = 0.0f;
    int x = (int)(fu * (float)W), y = (int)(fv * (float)H);
    if (x >= W) x = W - 1;
    if (y >= H) y = H - 1;
    return scl(env_texel(P, x, y), P->d->env.intensity);
}

static int search_cdf(const float *cdf, int n, float u) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid + 1] <= u) lo = mid + 1; else hi = mid;
    }
    return lo < n ? lo : n - 1;
}

static float lum(v3 c) { return fmaxf(0.0f, 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z); }

static void sample_env_dir(const prep_t *P, float u0, float u1, v3 *out, float *pdf) {
    *out = mk(0, 1, 0);
    *pdf = 0.0f;
    const int W = (int)P->d->env.w, H = (int)P->d->env.h;
    if (!P->has_env || P->env_total <= 0.0f) return;
    const int y = search_cdf(P->env_marg, H, u0);
    const float *row = P->env_cond + (size_t)y * (W + 1);
    const int x = search_cdf(row, W, u1);
    const float yp = P->env_marg[y], ys = P->env_marg[y + 1] - yp;
    const float xp = row[x], xs = row[x + 1] - xp;
    const float yf = (ys > 0.0f) ? (u0 - yp) / ys : 0.5f;
    const float xf = (xs > 0.0f) ? (u1 - xp) / xs : 0.5f;
    const float u = ((float)x + xf) / (float)W;
    const float v = ((float)y + yf) / (float)H;
    const float phi = (u - 0.5f) * 2.0f * PI_F;
    const float theta = (0.5f - v) * PI_F;
    const float ct = cosf(theta), st = sinf(theta);
    const float dx = ct * cosf(phi), dy = st, dz = ct * sinf(phi);
    const float yaw = P->d->env.rotation_deg * (float)(M_PI / 180.0);
    const float cy = cosf(yaw), sy = sinf(yaw);
    const float wx = dx * cy + dz * sy;
    const float wz = -dx * sy + dz * cy;
    *out = nrm(mk(wx, dy, wz));
    const float l = lum(env_texel(P, x, y));
    const float sin_t = sinf(PI_F * v);
    if (sin_t <= 0.0f) { *pdf = 0.0f; return; }
    const float pdf_uv = (l * sin_t) / P->env_total;
    *pdf = pdf_uv * (float)(W * H) / (2.0f * (float)(M_PI * M_PI) * sin_t);
}

static float env_dir_pdf(const prep_t *P, v3 dir) {
    if (!P->has_env || P->env_total <= 0.0f) return 0.0f;
    const v3 d = nrm(dir);
    const float yaw = P->d->env.rotation_deg * (float)(M_PI / 180.0);
    const float cy = cosf(yaw), sy = sinf(yaw);
    const float rx = d.x * cy - d.z * sy;
    const float rz = d.x * sy + d.z * cy;
    const int W = (int)P->d->env.w, H = (int)P->d->env.h;
    const float u = 0.5f + atan2f(rz, rx) * (float)(0.5 / M_PI);
    const float v = 0.5f - asinf(fminf(fmaxf(d.y, -1.0f), 1.0f)) * (float)(1.0 / M_PI);
    int x = (int)(u * (float)W); if (x < 0) x = 0; if (x >= W) x = W - 1;
    int y = (int)(v * (float)H); if (y < 0) y = 0; if (y >= H) y = H - 1;
    const float l = lum(env_texel(P, x, y));
    const float sin_t = sinf(PI_F * v);
    if (sin_t <= 0.0f) return 0.0f;
    const float pdf_uv = (l * sin_t) / P->env_total;
    return pdf_uv * (float)(W * H) / (2.0f * (float)(M_PI * M_PI) * sin_t);
}

/* ------------------------------------------------------------------------ */
/* media (medium.cpp:18-123, 475-525)                                        */
/* ------------------------------------------------------------------------ */
static v3 transmittance(const medium_t *m, float d) {
    if (d <= 0.0f) return mk(1, 1, 1);
    return mk(expf(-m->sigma_t.x * d), expf(-m->sigma_t.y * d), expf(-m->sigma_t.z * d));
}
static float ew_channel(float st, float d) {
    if (st < 1e-6f) return d;
    return (1.0f - expf(-st * d)) / st;
}
static v3 emission_along(const medium_t *m, float d) {
    if (d <= 0.0f) return mk(0, 0, 0);
    return mk(m->sigma_e_eff.x * ew_channel(m->sigma_t.x, d), m->sigma_e_eff.y * ew_channel(m->sigma_t.y, d),
              m->sigma_e_eff.z * ew_channel(m->sigma_t.z, d));
}
static float free_flight(const medium_t *m, float u) {
    if (m->hero <= 0.0f) return 1e30f;
    const float xi = fminf(0.9999999f, fmaxf(0.0f, u));
    return -logf(1.0f - xi) / m->hero;
}
/* value-noise basis of the density and refractive-index fields
 * (medium.cpp:125-186): integer lattice hash to [0, 1), smoothstep-weighted
 * trilinear interpolation, fBm normalised by the summed amplitudes */
static float lattice(int x, int y, int z) {
    uint32_t h = (uint32_t)x * 374761393u + (uint32_t)y * 668265263u + (uint32_t)z * 2147483647u;
    h = (h ^ (h >> 13)) * 1274126177u;
    h ^= h >> 16;
    return (float)(h & 0x00FFFFFFu) / (float)0x01000000u;
}
static float fade(float x) { return x * x * (3.0f - 2.0f * x); }
static float value_noise(float x, float y, float z) {
    const int xi = (int)floorf(x), yi = (int)floorf(y), zi = (int)floorf(z);
    const float u = fade(x - (float)xi), v = fade(y - (float)yi), w = fade(z - (float)zi);
    float c[2][2][2];
    for (int k = 0; k < 2; k++)
        for (int j = 0; j < 2; j++)
            for (int i = 0; i < 2; i++) c[k][j][i] = lattice(xi + i, yi + j, zi + k);
    const float x00 = c[0][0][0] * (1.0f - u) + c[0][0][1] * u;
    const float x10 = c[0][1][0] * (1.0f - u) + c[0][1][1] * u;
    const float x01 = c[1][0][0] * (1.0f - u) + c[1][0][1] * u;
    const float x11 = c[1][1][0] * (1.0f - u) + c[1][1][1] * u;
    const float y0 = x00 * (1.0f - v) + x10 * v;
    const float y1 = x01 * (1.0f - v) + x11 * v;
    return y0 * (1.0f - w) + y1 * w;
}
static float fbm01(float x, float y, float z, float freq, int octaves, float gain, float lac) {
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

/* densityAt (medium.cpp:271-290): 1 without a density field */
static float density_at(const medium_t *m, v3 p) {
    if (!m->het) return 1.0f;
    const or_medium *s = m->src;
    const float n = fbm01(p.x - s->noise_offset[0], p.y - s->noise_offset[1], p.z - s->noise_offset[2],
                          s->noise_frequency, s->noise_octaves, s->noise_gain, s->noise_lacunarity);
    return fmaxf(0.0f, s->noise_amplitude * n + s->noise_bias);
}

/* emissionAt (medium.cpp:292-404): constant sigma_e (x density), a path
 * tube's radial falloff and inner-to-outer blend, or the cone's axial
 * gaussian and (1 - v^2)^2 radial profile, all times the density field */
static v3 emission_at(const medium_t *m, v3 p) {
    const or_medium *s = m->src;
    if (s->emission_mode == 0) return m->het ? scl(m->sigma_e_eff, density_at(m, p)) : m->sigma_e_eff;
    if (s->emission_mode == 2) {
        const uint32_t n = s->n_path;
        if (n < 2) return mk(0, 0, 0);
        float best = INFINITY, bt = 0.0f, br = 0.0f;
        for (uint32_t i = 0; i + 1 < n; i++) {
            const or_path_point *a = &s->path[i], *b = &s->path[i + 1];
            const v3 ab = mk(b->p[0] - a->p[0], b->p[1] - a->p[1], b->p[2] - a->p[2]);
            const float ab2 = ab.x * ab.x + ab.y * ab.y + ab.z * ab.z;
            if (ab2 < 1e-12f) continue;
            const v3 ap = mk(p.x - a->p[0], p.y - a->p[1], p.z - a->p[2]);
            float u = (ap.x * ab.x + ap.y * ab.y + ap.z * ab.z) / ab2;
            u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
            const v3 q = mk(a->p[0] + ab.x * u, a->p[1] + ab.y * u, a->p[2] + ab.z * u);
            const float dx = p.x - q.x, dy = p.y - q.y, dz = p.z - q.z;
            const float d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < best) {
                best = d2;
                bt = a->t + (b->t - a->t) * u;
                br = a->radius + (b->radius - a->radius) * u;
            }
        }
        if (br <= 0.0f) return mk(0, 0, 0);
        if (best >= br * br) return mk(0, 0, 0);
        const float r = sqrtf(best) / br;
        const float radial = powf(fmaxf(0.0f, 1.0f - r), s->path_falloff_power);
        if (radial <= 0.0f) return mk(0, 0, 0);
        const float tb = bt < 0.0f ? 0.0f : (bt > 1.0f ? 1.0f : bt);
        const v3 col = add(scl(ld3(s->path_inner), 1.0f - tb), scl(ld3(s->path_outer), tb));
        float k = radial * s->path_intensity * fmaxf(0.0f, m->density);
        if (m->het) k *= density_at(m, p);
        return scl(col, k);
    }
    const v3 rel = mk(p.x - s->cone_origin[0], p.y - s->cone_origin[1], p.z - s->cone_origin[2]);
    const v3 ax = m->cone_axis;
    const float at = rel.x * ax.x + rel.y * ax.y + rel.z * ax.z;
    if (s->cone_length <= 0.0f) return mk(0, 0, 0);
    const float u = at / s->cone_length;
    if (u < 0.0f || u > 1.0f) return mk(0, 0, 0);
    const v3 pp = mk(rel.x - ax.x * at, rel.y - ax.y * at, rel.z - ax.z * at);
    const float r2 = pp.x * pp.x + pp.y * pp.y + pp.z * pp.z;
    if (s->cone_radius <= 0.0f) return mk(0, 0, 0);
    const float v2 = r2 / (s->cone_radius * s->cone_radius);
    if (v2 > 1.0f) return mk(0, 0, 0);
    const float dx = (u - s->cone_peak_axial) * s->cone_peak_sharpness;
    const float t = expf(-dx * dx) * ((1.0f - v2) * (1.0f - v2));
    if (t <= 0.0f) return mk(0, 0, 0);
    const v3 col = add(scl(ld3(s->cone_inner), t), scl(ld3(s->cone_outer), 1.0f - t));
    float k = t * s->cone_intensity * fmaxf(0.0f, m->density);
    if (m->het) k *= density_at(m, p);
    return scl(col, k);
}

/* sampleDeltaTracking (medium.cpp:406-440): exponential steps at the
 * bound's rate, each accepted with probability density / peak; at most 64 */
static int delta_track(const medium_t *m, rng_t *R, v3 o, v3 d, float maxT, float *out_t, float *out_den) {
    if (m->t_max_hero <= 0.0f || maxT <= 0.0f) return 0;
    const float peak = m->hero > 0.0f ? m->t_max_hero / m->hero : 1.0f;
    if (peak <= 0.0f) return 0;
    float t = 0.0f;
    for (int i = 0; i < 64; i++) {
        const float xi = fminf(0.9999999f, fmaxf(0.0f, prng(R)));
        t += -logf(1.0f - xi) / m->t_max_hero;
        if (t >= maxT) return 0;
        const float den = density_at(m, add(o, scl(d, t)));
        if (prng(R) < den / peak) {
            *out_t = t;
            *out_den = den;
            return 1;
        }
    }
    return 0;
}

/* emissionIntegralAlongRay (medium.cpp:442-472): the closed form for
 * constant emission, else one jittered sample per stratum of Tr x sigma_e */
static v3 emission_integral(const medium_t *m, rng_t *R, v3 o, v3 d, float maxT) {
    const or_medium *s = m->src;
    if (s->emission_mode == 0) return emission_along(m, maxT);
    if (maxT <= 0.0f) return mk(0, 0, 0);
    const int cnt = s->emission_mode == 2 ? s->path_samples : s->cone_samples;
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

/* iorAt / iorGradientAt / bendRay (medium.cpp:188-269): n = 1 + amplitude x
 * fade x 2 (fbm - 1/2), central differences at half a noise cell, and an
 * eikonal march of iorMarchSteps midpoint steps */
static float ior_at(const medium_t *m, v3 p) {
    const or_medium *s = m->src;
    const float n = fbm01(p.x - s->ior_offset[0], p.y - s->ior_offset[1], p.z - s->ior_offset[2], s->ior_frequency,
                          s->ior_octaves, s->ior_gain, s->ior_lacunarity);
    float am = 1.0f;
    if (s->ior_falloff != 0.0f && s->cone_length > 0.0f) {
        const float rx = p.x - s->cone_origin[0], ry = p.y - s->cone_origin[1], rz = p.z - s->cone_origin[2];
        const float at = rx * m->cone_axis.x + ry * m->cone_axis.y + rz * m->cone_axis.z;
        float u = at / s->cone_length;
        u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
        const float mag = s->ior_falloff < 0.0f ? -s->ior_falloff : s->ior_falloff;
        const float t = s->ior_falloff >= 0.0f ? u : 1.0f - u;
        am = 1.0f - mag * t;
        if (am < 0.0f) am = 0.0f;
    }
    return 1.0f + s->ior_amplitude * am * (n - 0.5f) * 2.0f;
}
static v3 ior_grad(const medium_t *m, v3 p) {
    const float f = m->src->ior_frequency;
    const float h = f > 0.0f ? 0.5f / f : 0.05f;
    const float xp = ior_at(m, mk(p.x + h, p.y, p.z)), xm = ior_at(m, mk(p.x - h, p.y, p.z));
    const float yp = ior_at(m, mk(p.x, p.y + h, p.z)), ym = ior_at(m, mk(p.x, p.y - h, p.z));
    const float zp = ior_at(m, mk(p.x, p.y, p.z + h)), zm = ior_at(m, mk(p.x, p.y, p.z - h));
    const float i2h = 0.5f / h;
    return mk((xp - xm) * i2h, (yp - ym) * i2h, (zp - zm) * i2h);
}
static void bend_ray(const medium_t *m, v3 o, v3 d, float length, v3 *out_o, v3 *out_d) {
    if (length <= 0.0f || !m->haze) {
        *out_o = add(o, scl(d, length));
        *out_d = d;
        return;
    }
    const int N = m->src->ior_march_steps;
    const float ds = length / (float)N;
    v3 pos = o, dir = d;
    for (int i = 0; i < N; i++) {
        const v3 mid = mk(pos.x + dir.x * (ds * 0.5f), pos.y + dir.y * (ds * 0.5f), pos.z + dir.z * (ds * 0.5f));
        const float n = ior_at(m, mid);
        const v3 g = ior_grad(m, mid);
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
    *out_o = pos;
    *out_d = dir;
}

static float phase_pdf(const medium_t *m, v3 wo, v3 wi) {
    const float c = dot(wo, wi);
    const float g = m->g;
    const float den = 1.0f + g * g - 2.0f * g * c;
    return (1.0f - g * g) * (float)(1.0 / (4.0 * M_PI)) / (den * sqrtf(fmaxf(1e-8f, den)));
}
static v3 sample_phase(const medium_t *m, v3 wo, float u1, float u2) {
    float ct;
    if (fabsf(m->g) < 1e-3f) ct = 1.0f - 2.0f * u1;
    else {
        const float g = m->g;
        const float sq = (1.0f - g * g) / (1.0f - g + 2.0f * g * u1);
        ct = (1.0f + g * g - sq * sq) / (2.0f * g);
    }
    const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    const float phi = 2.0f * PI_F * u2;
    v3 t1, t2;
    if (fabsf(wo.x) > fabsf(wo.y)) {
        const float il = 1.0f / sqrtf(wo.x * wo.x + wo.z * wo.z);
        t1 = mk(-wo.z * il, 0.0f, wo.x * il);
    } else {
        const float il = 1.0f / sqrtf(wo.y * wo.y + wo.z * wo.z);
        t1 = mk(0.0f, wo.z * il, -wo.y * il);
    }
    t2 = cross(wo, t1);
    return nrm(add(add(scl(t1, st * cosf(phi)), scl(t2, st * sinf(phi))), scl(wo, ct)));
}
static const medium_t *medium_at(const prep_t *P, int idx) {
    if (idx < 0) return NULL;
    return &P->media[idx];
}

/* ------------------------------------------------------------------------ */
/* sampling helpers                                                          */
/* ------------------------------------------------------------------------ */
/* cosineWeightedDirection (raycommon.cpp:94-115) */
static v3 cosine_dir(v3 n, float r1, float r2) {
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

/* ldsPointInTriangle (raycommon.cpp:117-122) */
static v3 point_in_tri(const prep_t *P, uint32_t t, float u, float v) {
    if (u + v > 1.0f) { u = 1.0f - u; v = 1.0f - v; }
    const v3 a = tri_v(P, t, 0);
    return add(add(a, scl(sub(tri_v(P, t, 1), a), u)), scl(sub(tri_v(P, t, 2), a), v));
}

/* randomRayInHemisphere (ugm, not vendored): uniform hemisphere about n */
static v3 hemisphere_dir(v3 n, float u1, float u2) {
    const float z = u1;
    const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    const float phi = 2.0f * PI_F * u2;
    v3 tg;
    if (fabsf(n.x) > fabsf(n.y)) tg = nrm(cross(mk(0, 1, 0), n));
    else tg = nrm(cross(mk(1, 0, 0), n));
    const v3 bt = cross(n, tg);
    return add(add(scl(tg, r * cosf(phi)), scl(bt, r * sinf(phi))), scl(n, z));
}

static float fresnel_schlick(float c, float ior) {
    float r0 = (1.0f - ior) / (1.0f + ior);
    r0 = r0 * r0;
    const float m = 1.0f - c;
    return r0 + (1.0f - r0) * (m * m) * (m * m) * m;
}

/* refract (bsdf.cpp:39-52) */
static v3 refract_dir(v3 d, v3 n, float r) {
    const v3 nl = dot(d, n) < 0.0f ? n : neg(n);
    const int into = dot(nl, n) > 0.0f;
    if (into) r = 1.0f / r;
    const float c = dot(d, nl);
    const float t = 1.0f - r * r * (1.0f - c * c);
    if (t < 0.0f) return reflect(d, n);
    return nrm(sub(scl(d, r), scl(n, (into ? 1.0f : -1.0f) * (c * r + sqrtf(t)))));
}

/* GGX helpers (bsdf.cpp:129-170) */
typedef struct { v3 t, b, n; float ax, ay; v3 V; float g1v; v3 F0; } ggx_t;

static void tangent_frame(v3 n, v3 *t, v3 *b) {
    const v3 up = (fabsf(n.y) > 0.99f) ? mk(1, 0, 0) : mk(0, 1, 0);
    *t = nrm(cross(up, n));
    *b = cross(n, *t);
}
static float smith_g1(v3 w, float ax, float ay) {
    if (w.z <= 0.0f) return 0.0f;
    const float ax2 = (ax * w.x) * (ax * w.x);
    const float ay2 = (ay * w.y) * (ay * w.y);
    const float lambda = 0.5f * (sqrtf(1.0f + (ax2 + ay2) / (w.z * w.z)) - 1.0f);
    return 1.0f / (1.0f + lambda);
}
static v3 sample_vndf(v3 Ve, float ax, float ay, float u1, float u2) {
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
static v3 schlick_f(v3 F0, float vdoth) {
    const float a1 = fmaxf(0.0f, 1.0f - vdoth);
    const float a5 = a1 * a1 * a1 * a1 * a1;
    return add(F0, scl(sub(mk(1, 1, 1), F0), a5));
}
static float ggx_d(v3 H, float ax, float ay) {
    const float hx = H.x / ax, hy = H.y / ay;
    const float den = hx * hx + hy * hy + H.z * H.z;
    return 1.0f / (PI_F * ax * ay * den * den);
}
/* evalGlossy (bsdf.cpp:258-280) */
static int eval_glossy(const ggx_t *G, v3 dw, v3 *fcos, float *pdf) {
    const v3 Ll = mk(dot(dw, G->t), dot(dw, G->b), dot(dw, G->n));
    if (Ll.z <= 0.0f) return 0;
    v3 Hl = add(G->V, Ll);
    const float h2 = Hl.x * Hl.x + Hl.y * Hl.y + Hl.z * Hl.z;
    if (h2 <= 0.0f) return 0;
    Hl = scl(Hl, 1.0f / sqrtf(h2));
    if (Hl.z <= 0.0f) return 0;
    const float vdh = fmaxf(0.0f, dot(G->V, Hl));
    if (vdh <= 0.0f) return 0;
    const float D = ggx_d(Hl, G->ax, G->ay);
    const float g1l = smith_g1(Ll, G->ax, G->ay);
    const v3 F = schlick_f(G->F0, vdh);
    *fcos = scl(F, D * G->g1v * g1l / (4.0f * G->V.z));
    *pdf = G->g1v * D / (4.0f * G->V.z);
    return 1;
}
/* returns 0 when the view is below the shading plane */
static int ggx_setup(const or_object *m, v3 alb, v3 n, v3 in_dir, ggx_t *G) {
    G->F0 = (m->metallic >= 1.0f) ? alb
            : add(scl(alb, m->metallic), scl(mk(0.04f, 0.04f, 0.04f), 1.0f - m->metallic));
    tangent_frame(n, &G->t, &G->b);
    if (fabsf(m->aniso_rotation) > 0.0f) {
        const float a = m->aniso_rotation * (float)(M_PI / 180.0);
        const float ca = cosf(a), sa = sinf(a);
        const v3 t2 = add(scl(G->t, ca), scl(G->b, sa));
        const v3 b2 = sub(scl(G->b, ca), scl(G->t, sa));
        G->t = t2;
        G->b = b2;
    }
    G->n = n;
    const float an = fmaxf(-0.99f, fminf(0.99f, m->anisotropy));
    const float r2 = m->roughness * m->roughness;
    G->ax = fmaxf(1e-3f, r2 * (1.0f - an));
    G->ay = fmaxf(1e-3f, r2 * (1.0f + an));
    const v3 V = neg(in_dir);
    G->V = mk(dot(V, G->t), dot(V, G->b), dot(V, n));
    if (G->V.z <= 0.0f) return 0;
    G->g1v = smith_g1(G->V, G->ax, G->ay);
    return 1;
}

/* ------------------------------------------------------------------------ */
/* light sampling                                                            */
/* ------------------------------------------------------------------------ */
typedef struct { v3 dir; float pdf; v3 Le; int ok; } nee_t;

/* sampleAreaLightForNEE (rayrenderer.cpp:1976-2036); traceAreaLight is the
 * same draw with the Lambert MIS weight folded in (:1916-1974). */
static nee_t sample_area(const prep_t *P, rng_t *R, v3 hit, v3 n, float *cos_obj, float *cos_light,
                         float *r2_out, float *sampled_area) {
    nee_t s;
    memset(&s, 0, sizeof s);
    if (P->nlights == 0) return s;
    const uint32_t li = pick(prng(R), P->nlights);
    const uint32_t cnt = P->light_count[li];
    const uint32_t tri = P->light_tris[P->light_first[li] + pick(prng(R), cnt)];
    const float u = lds1(R), v = lds1(R);
    const v3 p = point_in_tri(P, tri, u, v);
    const v3 lr = sub(p, hit);
    const v3 dir = nrm(lr);
    const float co = dot(dir, n);
    if (co <= 0.0f) return s;
    const v3 ln = interp_normal_at(P, tri, p);
    const float cl = dot(neg(dir), ln);
    if (cl <= 0.0f) return s;
    if (occluded(P, surface_origin(hit, lr, n), lr, 0.99999f, 1)) return s;
    const or_object *lo = &P->d->objs[P->light_obj[li]];
    const float sa = (float)cnt * P->area[tri];
    const float r2 = lr.x * lr.x + lr.y * lr.y + lr.z * lr.z;
    s.ok = 1;
    s.dir = dir;
    s.pdf = r2 / (cl * sa);
    s.Le = scl(ld3(lo->color), lo->emission);
    *cos_obj = co;
    *cos_light = cl;
    *r2_out = r2;
    *sampled_area = sa;
    return s;
}

/* traceAreaLight's returned value (Lambert NEE incl. 1/pi and MIS) */
static v3 trace_area_light(const prep_t *P, rng_t *R, v3 hit, v3 n) {
    float co, cl, r2, sa;
    nee_t s = sample_area(P, R, hit, n, &co, &cl, &r2, &sa);
    if (!s.ok) return mk(0, 0, 0);
    const float pl = r2 / (cl * sa);
    const float pb = co / PI_F;
    const float pl2 = pl * pl, pb2 = pb * pb;
    const float w = pl2 / (pl2 + pb2);
    return scl(scl(s.Le, co * cl * sa / (PI_F * r2)), w);
}

/* ugm smoothstep (not vendored; the GLSL definition) */
static float smoothstep_f(float e0, float e1, float x) {
    float t = (x - e0) / (e1 - e0);
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

static float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

/* tracePointLight (rayrenderer.cpp:2205-2269): shadow ray toward the light
 * (unnormalised, maxt 0.99999, blocked by non-emissive triangles with
 * transparency < 0.01 or refraction > 0.1: the environment predicate); the
 * unblocked value is color x emission x dotToLight / |lightray|^2 with
 * dotToLight = |dotToObject| for an omni light, the spot falloff
 * dot(l, n_light) x smoothstep(cos - 0.1, cos + 0.1, dot(l, n_light)) with
 * cos = cosf(RADIAN_TO_DEGREE(spotRange / 2)) for a spot, each channel
 * clamped to [0, 1] (ugm clamp).  RADIAN_TO_DEGREE(x) is taken as
 * x * 180 / pi, its name's meaning (ugm is not vendored). */
static v3 trace_point_light(const prep_t *P, const or_point_light *L, v3 hit, v3 n) {
    const v3 lr = sub(ld3(L->pos), hit);
    if (occluded(P, surface_origin(hit, lr, n), lr, 0.99999f, 2)) return mk(0, 0, 0);
    const v3 dir = nrm(lr);
    const float dto = dot(dir, n);
    float dtl = dot(dir, ld3(L->normal));
    if (!(dto > 0.0f)) return mk(0, 0, 0);
    if (L->spot_range > 0.0f) {
        const float sc = cosf(L->spot_range * 0.5f * (float)(180.0 / M_PI));
        dtl = dtl * smoothstep_f(fmaxf(sc - 0.1f, 0.0f), fminf(sc + 0.1f, 1.0f), dtl);
    } else {
        dtl = fabsf(dto);
    }
    if (!(dtl > 0.0f)) return mk(0, 0, 0);
    const float da = powf(len(lr), -2.0f);
    const float lum = L->emission * dtl * da;
    return mk(clamp01(L->color[0] * lum), clamp01(L->color[1] * lum), clamp01(L->color[2] * lum));
}

/* HomogeneousMedium::prepare (medium.cpp:18-57) and isActive (medium.h:214-223) */
static void prep_medium(medium_t *q, const or_medium *m) {
    memset(q, 0, sizeof *q);
    const float den = fmaxf(0.0f, m->density);
    q->sigma_s_eff = scl(ld3(m->sigma_s), den);
    q->sigma_e_eff = scl(ld3(m->sigma_e), den);
    q->sigma_t = add(scl(ld3(m->sigma_a), den), q->sigma_s_eff);
    q->hero = (q->sigma_t.x + q->sigma_t.y + q->sigma_t.z) * (1.0f / 3.0f);
    q->g = m->g;
    q->src = m;
    q->het = m->density_field == 1;
    q->peak = q->het ? fmaxf(0.0f, m->noise_amplitude * 1.0f + m->noise_bias) : 1.0f;
    q->t_max_hero = q->het ? q->hero * q->peak : q->hero;
    const float al2 = m->cone_axis[0] * m->cone_axis[0] + m->cone_axis[1] * m->cone_axis[1] +
                      m->cone_axis[2] * m->cone_axis[2];
    q->cone_axis = al2 > 1e-12f ? scl(ld3(m->cone_axis), 1.0f / sqrtf(al2)) : mk(0.0f, 0.0f, -1.0f);
    q->density = m->density;
    q->haze = m->heat_haze && m->ior_amplitude > 0.0f && m->ior_march_steps > 0;
    q->active = q->hero > 0.0f || (m->emission_mode == 1 && m->cone_intensity > 0.0f) ||
                (m->emission_mode == 2 && m->path_intensity > 0.0f && m->n_path >= 2) ||
                (m->heat_haze && m->ior_amplitude > 0.0f) || !is_zero(q->sigma_e_eff);
}

/* sampleVolumeLightForNEE (rayrenderer.cpp:2066-2203): one emissive volume
 * picked uniformly (a rand() draw); equiangular sampling (Kulla & Fajardo
 * 2012) of its cone axis, or of the owning object's location for constant
 * emission; shadow ray with the volume lights' occluders; sigma_e at the
 * point, times the medium's homogeneous transmittance over the distance;
 * solid-angle pdf times the number of volumes. */
static int sample_volume(const prep_t *P, rng_t *R, v3 hit, v3 n, v3 *out_dir, float *out_pdf, v3 *out_le) {
    const uint32_t N = P->d->vlights ? P->d->nvlights : 0;
    if (N == 0) return 0;
    const or_volume_light *ev = &P->d->vlights[pick(prng(R), N)];
    const medium_t *m = medium_at(P, (int)ev->medium);
    if (!m) return 0;
    v3 a, b;
    if (m->src->emission_mode == 1) {
        a = ld3(m->src->cone_origin);
        b = add(a, scl(m->cone_axis, m->src->cone_length));
    } else {
        a = b = ld3(ev->pos);
    }
    const v3 sd = sub(b, a);
    const float sl = len(sd);
    const v3 axis = sl > 1e-6f ? scl(sd, 1.0f / sl) : mk(0.0f, 0.0f, -1.0f);
    const v3 toA = sub(a, hit);
    const float along = dot(toA, axis);
    const v3 perp = sub(toA, scl(axis, along));
    const float D = len(perp);
    if (!(D > 1e-5f) && sl <= 1e-6f) return 0;
    const float thA = atanf(along / fmaxf(D, 1e-5f));
    const float thB = atanf((along + sl) / fmaxf(D, 1e-5f));
    if (!(fabsf(thB - thA) > 1e-6f) && sl > 1e-6f) return 0;
    const float u = prng(R);
    const float th = thA + u * (thB - thA);
    const float ts = D * tanf(th);
    const v3 q = sl > 1e-6f ? add(a, scl(axis, ts - along)) : a;
    const v3 tq = sub(q, hit);
    const float r2 = tq.x * tq.x + tq.y * tq.y + tq.z * tq.z;
    const float r = sqrtf(r2);
    if (!(r > 1e-5f)) return 0;
    if (occluded(P, surface_origin(hit, tq, n), tq, 0.99999f, 3)) return 0;
    v3 se = emission_at(m, q);
    if (is_zero(se)) return 0;
    if (m->hero > 0.0f) se = mul(se, transmittance(m, r));
    const float pdf_t = D / (fmaxf(1e-6f, fabsf(thB - thA)) * (D * D + ts * ts));
    *out_dir = scl(tq, 1.0f / r);
    *out_pdf = pdf_t * r2 * (float)N;
    *out_le = se;
    return 1;
}

/* traceLight (rayrenderer.cpp:2271-2293): one area light's traceAreaLight
 * plus one point light's tracePointLight (the only one, or one picked
 * uniformly by a rand() draw) */
static v3 trace_light(const prep_t *P, rng_t *R, v3 hit, v3 n) {
    v3 c = trace_area_light(P, R, hit, n);
    const uint32_t np = P->d->plights ? P->d->nplights : 0;
    if (np > 0) {
        const uint32_t li = np == 1 ? 0u : pick(prng(R), np);
        c = add(c, trace_point_light(P, &P->d->plights[li], hit, n));
    }
    return c;
}

/* sampleEnvmapForNEE (rayrenderer.cpp:2038-2065) */
static nee_t sample_env_nee(const prep_t *P, rng_t *R, v3 hit, v3 n, float *cos_obj) {
    nee_t s;
    memset(&s, 0, sizeof s);
    if (!P->has_env || P->env_total <= 0.0f) return s;
    const float u0 = prng(R), u1 = prng(R);
    v3 dir;
    float pdf;
    sample_env_dir(P, u0, u1, &dir, &pdf);
    if (pdf <= 0.0f) return s;
    const float co = dot(dir, n);
    if (co <= 0.0f) return s;
    if (occluded(P, surface_origin(hit, dir, n), dir, 1e30f, 2)) return s;
    s.ok = 1;
    s.dir = dir;
    s.pdf = pdf;
    s.Le = sample_environment(P, dir);
    *cos_obj = co;
    return s;
}

/* traceEnvmapLight (rayrenderer.cpp:1484-1530) */
static v3 trace_env_light(const prep_t *P, rng_t *R, v3 hit, v3 n, float bsdf_pdf) {
    float co;
    nee_t s = sample_env_nee(P, R, hit, n, &co);
    if (!s.ok) return mk(0, 0, 0);
    float w = 1.0f;
    if (bsdf_pdf > 0.0f) {
        const float e2 = s.pdf * s.pdf, b2 = bsdf_pdf * bsdf_pdf;
        w = e2 / (e2 + b2);
    }
    return scl(s.Le, co * w / (PI_F * s.pdf));
}

static float env_miss_weight(const prep_t *P, v3 dir, float bsdf_pdf) {
    if (bsdf_pdf > 0.0f) {
        const float ep = env_dir_pdf(P, dir);
        const float b2 = bsdf_pdf * bsdf_pdf, e2 = ep * ep;
        return (b2 + e2 > 0.0f) ? b2 / (b2 + e2) : 1.0f;
    }
    return 1.0f;
}

/* ======================================================================== */
/* mode 0: recursive restatement                                            */
/* ======================================================================== */
typedef struct {
    int passes;
    v3 tp;          /* BSDFParam::throughput */
    float bsdf_pdf; /* BSDFParam::bsdfSampledPdf */
    int medium;     /* BSDFParam::currentMedium */
    int chroma;     /* BSDFParam::chromaChannel, -1 = none chosen */
} param_t;

static v3 trace_path(const prep_t *P, rng_t *R, v3 o, v3 d, param_t *sp);

static v3 shade(const prep_t *P, rng_t *R, const isect_t *h, v3 in_o, v3 in_d, v3 n, param_t *sp);

/* DiffuseShader (bsdf.cpp:54-108) */
static v3 diffuse_shade(const prep_t *P, rng_t *R, const isect_t *h, v3 in_o, v3 n, param_t *pm) {
    (void)in_o;
    const or_object *m = obj_of(P, h->prim);
    const float r1 = lds1(R), r2 = lds1(R);
    const v3 dir = cosine_dir(n, r1, r2);
    const v3 o = surface_origin(h->pos, dir, geom_normal(P, h->prim, n));
    /* albedo: colour x texture, 1 without colour sampling (bsdf.cpp:68-73) */
    const v3 alb = P->color_sampling ? mul(ld3(m->color), albedo_tex(P, h)) : mk(1.0f, 1.0f, 1.0f);
    const v3 st = pm->tp;
    const float sp = pm->bsdf_pdf;
    pm->tp = mul(pm->tp, alb);
    const float cs = fmaxf(0.0f, dot(dir, n));
    pm->bsdf_pdf = cs * (float)(1.0 / M_PI);
    v3 c = trace_path(P, R, o, dir, pm);
    c = add(c, trace_light(P, R, h->pos, n));
    c = add(c, trace_env_light(P, R, h->pos, n, pm->bsdf_pdf));
    {
        v3 vd, vle;
        float vp = 0.0f;
        if (sample_volume(P, R, h->pos, n, &vd, &vp, &vle) && vp > 0.0f) {
            const float cl = fmaxf(0.0f, dot(vd, n));
            c = add(c, scl(vle, cl * (float)(1.0 / M_PI) / vp));
        }
    }
    pm->tp = st;
    pm->bsdf_pdf = sp;
    return mul(c, alb);
}

/* GlossyShader (bsdf.cpp:173-374) */
static v3 glossy_shade(const prep_t *P, rng_t *R, const isect_t *h, v3 in_d, v3 n, param_t *pm) {
    const or_object *m = obj_of(P, h->prim);
    /* albedo: colour x texture, the colour alone without colour sampling
     * (bsdf.cpp:185-188) */
    const v3 alb = P->color_sampling ? mul(ld3(m->color), albedo_tex(P, h)) : ld3(m->color);
    const v3 F0 = (m->metallic >= 1.0f) ? alb
                  : add(scl(alb, m->metallic), scl(mk(0.04f, 0.04f, 0.04f), 1.0f - m->metallic));
    if (m->roughness < 1e-3f) {
        const v3 gN = geom_normal(P, h->prim, n);
        v3 r = reflect(in_d, n);
        if (dot(r, gN) <= 0.0f) r = reflect(in_d, gN);
        const float ci = fmaxf(0.0f, -dot(in_d, n));
        const v3 F = schlick_f(F0, ci);
        const v3 st = pm->tp;
        const float sp = pm->bsdf_pdf;
        pm->tp = mul(pm->tp, F);
        pm->bsdf_pdf = 0.0f;
        const v3 inc = trace_path(P, R, surface_origin(h->pos, r, gN), r, pm);
        pm->tp = st;
        pm->bsdf_pdf = sp;
        return mul(inc, F);
    }
    ggx_t G;
    if (!ggx_setup(m, alb, n, in_d, &G)) return mk(0, 0, 0);
    v3 direct = mk(0, 0, 0);
    {
        float co, cl, r2, sa;
        nee_t s = sample_area(P, R, h->pos, n, &co, &cl, &r2, &sa);
        v3 fc;
        float pb;
        if (s.ok && eval_glossy(&G, s.dir, &fc, &pb)) {
            const float pl2 = s.pdf * s.pdf, pb2 = pb * pb;
            const float w = pl2 / (pl2 + pb2);
            direct = add(direct, scl(mul(s.Le, fc), w / s.pdf));
        }
    }
    {
        float co;
        nee_t s = sample_env_nee(P, R, h->pos, n, &co);
        v3 fc;
        float pb;
        if (s.ok && eval_glossy(&G, s.dir, &fc, &pb)) {
            const float pe2 = s.pdf * s.pdf, pb2 = pb * pb;
            const float w = pe2 / (pe2 + pb2);
            direct = add(direct, scl(mul(s.Le, fc), w / s.pdf));
        }
    }
    {
        v3 vd, vle, fc;
        float vp = 0.0f, pb;
        if (sample_volume(P, R, h->pos, n, &vd, &vp, &vle) && vp > 0.0f && eval_glossy(&G, vd, &fc, &pb))
            direct = add(direct, scl(mul(vle, fc), 1.0f / vp));
    }
    const float u1 = prng(R), u2 = prng(R);
    const v3 H = sample_vndf(G.V, G.ax, G.ay, u1, u2);
    const v3 Ll = sub(scl(H, 2.0f * dot(G.V, H)), G.V);
    if (Ll.z <= 0.0f) return direct;
    const v3 L = add(add(scl(G.t, Ll.x), scl(G.b, Ll.y)), scl(n, Ll.z));
    const v3 gN = geom_normal(P, h->prim, n);
    if (dot(L, gN) <= 0.0f) return direct;
    const float vdh = fmaxf(0.0f, dot(G.V, H));
    const v3 F = schlick_f(G.F0, vdh);
    const float g1o = smith_g1(Ll, G.ax, G.ay);
    const v3 wgt = scl(F, g1o);
    const float pdf = G.g1v * ggx_d(H, G.ax, G.ay) / (4.0f * G.V.z);
    const v3 st = pm->tp;
    const float sp = pm->bsdf_pdf;
    pm->tp = mul(pm->tp, wgt);
    pm->bsdf_pdf = pdf;
    const v3 inc = trace_path(P, R, surface_origin(h->pos, L, gN), L, pm);
    pm->tp = st;
    pm->bsdf_pdf = sp;
    return add(direct, mul(inc, wgt));
}

/* medium index after crossing the interface (bsdf.cpp:462-474, 530-538) */
static int crossed_medium(const prep_t *P, const or_object *m, int entering, int cur) {
    if (entering) return m->interior_medium >= 0 ? m->interior_medium : cur;
    return P->d->global_medium;
}

/* Chromatic dispersion of RefractionShader (bsdf.cpp:394-420): the first
 * dispersive interface of a path draws its band (0, 1, 2 with probability
 * 1/3 each), stored in *chroma for every later interface; the band's index
 * of refraction is refraction_ratio * (1 - s), * 1, * (1 + s).  Returns the
 * index of refraction and sets *mask to 3 in the band drawn here (1, 1, 1
 * when the band was drawn earlier or the material does not disperse). */
static float dispersion_ior(const or_object *m, rng_t *R, int *chroma, v3 *mask) {
    float ior = m->refraction_ratio;
    *mask = mk(1, 1, 1);
    if (m->chroma_dispersion > 0.0f) {
        int first = 0;
        if (*chroma < 0) {
            *chroma = (int)fminf(2.0f, floorf(prng(R) * 3.0f));
            first = 1;
        }
        const float s = m->chroma_dispersion;
        if (*chroma == 0) ior = m->refraction_ratio * (1.0f - s);
        else if (*chroma == 2) ior = m->refraction_ratio * (1.0f + s);
        if (first) *mask = mk(*chroma == 0 ? 3.0f : 0.0f, *chroma == 1 ? 3.0f : 0.0f, *chroma == 2 ? 3.0f : 0.0f);
    }
    return ior;
}

/* RefractionShader (bsdf.cpp:376-483) */
static v3 refraction_shade(const prep_t *P, rng_t *R, const isect_t *h, v3 in_d, v3 n, param_t *pm) {
    const or_object *m = obj_of(P, h->prim);
    const float ct = fabsf(dot(in_d, n));
    v3 mask;
    const float ior = dispersion_ior(m, R, &pm->chroma, &mask);
    const float fr = fresnel_schlick(ct, ior);
    const int refl = prng(R) < fr;
    v3 dir = refl ? reflect(in_d, n) : refract_dir(in_d, n, ior);
    if (m->roughness > 0.0f) {
        const float u1 = prng(R), u2 = prng(R);
        dir = nrm(add(dir, scl(hemisphere_dir(n, u1, u2), m->roughness)));
    }
    const v3 gN = geom_normal(P, h->prim, n);
    if (refl && dot(dir, gN) <= 0.0f) dir = reflect(in_d, gN);
    const v3 st = pm->tp;
    const float sp = pm->bsdf_pdf;
    const int smed = pm->medium;
    const v3 col = ld3(m->color);
    pm->tp = mul(pm->tp, col);
    pm->bsdf_pdf = 0.0f;
    const float di = dot(in_d, n), dout = dot(dir, n);
    if (di * dout > 0.0f) pm->medium = crossed_medium(P, m, di < 0.0f, pm->medium);
    const v3 c = trace_path(P, R, surface_origin(h->pos, dir, gN), dir, pm);
    pm->tp = st;
    pm->bsdf_pdf = sp;
    pm->medium = smed;
    return mul(mul(c, col), mask);
}

/* TransparencyShader (bsdf.cpp:513-546) */
static v3 transparency_shade(const prep_t *P, rng_t *R, const isect_t *h, v3 in_d, v3 n, param_t *pm) {
    const or_object *m = obj_of(P, h->prim);
    const v3 st = pm->tp;
    const int smed = pm->medium;
    pm->tp = scl(pm->tp, m->transparency);
    pm->medium = crossed_medium(P, m, dot(in_d, n) < 0.0f, pm->medium);
    const v3 gN = geom_normal(P, h->prim, n);
    const v3 c = trace_path(P, R, surface_origin(h->pos, in_d, gN), in_d, pm);
    pm->tp = st;
    pm->medium = smed;
    return scl(c, m->transparency);
}

/* MixShader (bsdf.cpp:579-610) */
static v3 mix_shade(const prep_t *P, rng_t *R, const isect_t *h, v3 in_o, v3 in_d, v3 n, param_t *pm) {
    const or_object *m = obj_of(P, h->prim);
    v3 c = mk(0, 0, 0);
    const float wd = 1.0f - m->glossy - m->refraction;
    const v3 st = pm->tp;
    if (wd > 0.00001f) {
        pm->tp = scl(st, wd);
        c = add(c, scl(diffuse_shade(P, R, h, in_o, n, pm), wd));
    }
    if (m->glossy > 0.00001f) {
        pm->tp = scl(st, m->glossy);
        c = add(c, scl(glossy_shade(P, R, h, in_d, n, pm), m->glossy));
    }
    if (m->refraction > 0.00001f) {
        pm->tp = scl(st, m->refraction);
        c = add(c, scl(refraction_shade(P, R, h, in_d, n, pm), m->refraction));
    }
    pm->tp = st;
    return c;
}

static v3 preview_shade(const prep_t *P, rng_t *R, const isect_t *h, v3 in_d, v3 n, param_t *sp);

/* RayBSDFShaderProvider::shade (rayrenderer.cpp:2776-2896); the provider
 * of shading systems 0-3 instead (rayrenderer.cpp:59-77) */
static v3 shade(const prep_t *P, rng_t *R, const isect_t *h, v3 in_o, v3 in_d, v3 n, param_t *sp) {
    if (P->shader != 5) return preview_shade(P, R, h, in_d, n, sp);
    const or_object *m = obj_of(P, h->prim);
    param_t pm;
    pm.passes = 0;
    pm.tp = mk(1, 1, 1);
    pm.bsdf_pdf = 0.0f;
    pm.medium = sp ? sp->medium : P->d->global_medium;
    pm.chroma = -1; /* a fresh BSDFParam: the back-face and eye branches keep it */
    if (m->emission > 0.0f) {
        const v3 E = scl(ld3(m->color), m->emission);
        if (sp && sp->bsdf_pdf > 0.0f) {
            const float cl = -dot(in_d, n);
            if (cl <= 0.0f) return mk(0, 0, 0);
            const v3 dd = sub(h->pos, in_o);
            const float r2 = dd.x * dd.x + dd.y * dd.y + dd.z * dd.z;
            const float pl = r2 / (cl * P->sampled_area[h->prim]);
            const float pb = sp->bsdf_pdf;
            const float pb2 = pb * pb, pl2 = pl * pl;
            return scl(E, pb2 / (pb2 + pl2));
        }
        return E;
    }
    if (dot(in_d, n) > 0.0f) {
        if (m->transparency > 0.001f) {
            if (sp && sp->passes + 1 <= P->max_depth) {
                pm.passes = sp->passes + 1;
                pm.tp = sp->tp;
                return transparency_shade(P, R, h, in_d, n, &pm);
            }
            return mk(0, 0, 0);
        } else if (m->refraction < 0.001f && m->glossy > 0.001f) {
            return mk(0, 0, 0);
        }
    }
    if (sp) {
        if (sp->passes + 1 >= P->max_depth) {
            /* colour x texture, zero without colour sampling (rayrenderer.cpp:2839-2856) */
            if (1.0f - m->glossy - m->refraction > 0.00001f) {
                const v3 l = trace_light(P, R, h->pos, n);
                return P->color_sampling ? mul(l, mul(ld3(m->color), albedo_tex(P, h))) : mk(0, 0, 0);
            }
            return mk(0, 0, 0);
        }
        pm.passes = sp->passes + 1;
        pm.tp = sp->tp;
        pm.chroma = sp->chroma; /* rayrenderer.cpp:2863 */
        pm.bsdf_pdf = sp->bsdf_pdf;
        float rrw = 1.0f;
        if (pm.passes >= MIN_RR_DEPTH) {
            float q = maxc(pm.tp);
            q = fminf(RR_MAX_PROB, fmaxf(RR_MIN_PROB, q));
            if (prng(R) >= q) return mk(0, 0, 0);
            rrw = 1.0f / q;
        }
        if (m->transparency > 0.001f) return scl(transparency_shade(P, R, h, in_d, n, &pm), rrw);
        return scl(mix_shade(P, R, h, in_o, in_d, n, &pm), rrw);
    }
    if (m->transparency > 0.01f) {
        const v3 a = scl(mix_shade(P, R, h, in_o, in_d, n, &pm), 1.0f - m->transparency);
        return add(a, transparency_shade(P, R, h, in_d, n, &pm));
    }
    return mix_shade(P, R, h, in_o, in_d, n, &pm);
}

/* RayRenderer::tracePath (rayrenderer.cpp:1596-1864): surfaces, heat haze
 * (a bent ray continuing in the outer medium), and homogeneous (free
 * flight) or fBm (delta tracking) media with constant, cone or path
 * emission */
static v3 trace_path(const prep_t *P, rng_t *R, v3 o, v3 d, param_t *sp) {
    isect_t h = nearest(P, o, d);
    const int midx = sp ? sp->medium : P->d->global_medium;
    const medium_t *med = medium_at(P, midx);
    if (!med || !med->active) {
        if (h.hit) return shade(P, R, &h, o, d, interp_normal(P, &h), sp);
        v3 env = sample_environment(P, d);
        if (sp) env = scl(env, env_miss_weight(P, d, sp->bsdf_pdf));
        return env;
    }
    if (med->haze) {
        /* the global medium is never heat haze (refused when preparing) */
        v3 bo, bd;
        bend_ray(med, o, d, fminf(h.hit ? h.t : RAY_MAX_DISTANCE, 1000.0f), &bo, &bd);
        const int saved = sp ? sp->medium : -1;
        if (sp) sp->medium = P->d->global_medium;
        const v3 r = trace_path(P, R, bo, bd, sp);
        if (sp) sp->medium = saved;
        return r;
    }
    const float maxT = h.hit ? h.t : RAY_MAX_DISTANCE;
    int scattered;
    float tf = 0.0f, den = 1.0f;
    if (med->het) {
        scattered = delta_track(med, R, o, d, maxT, &tf, &den);
        if (!scattered) tf = maxT;
    } else {
        tf = free_flight(med, prng(R));
        scattered = tf < maxT;
    }
    const v3 memi = emission_integral(med, R, o, d, scattered ? tf : maxT);
    if (scattered) {
        v3 sw = mk(0, 0, 0);
        if (med->het) {
            const float inv = med->t_max_hero > 0.0f ? 1.0f / med->t_max_hero : 0.0f;
            const v3 ss = scl(med->sigma_s_eff, den);
            sw = mk(ss.x * inv, ss.y * inv, ss.z * inv);
        } else {
            const float ph = med->hero * expf(-med->hero * tf);
            const v3 Tr = transmittance(med, tf);
            if (ph > 0.0f)
                sw = mk(med->sigma_s_eff.x * Tr.x / ph, med->sigma_s_eff.y * Tr.y / ph,
                        med->sigma_s_eff.z * Tr.z / ph);
        }
        if (is_zero(sw)) return memi;
        const v3 pos = add(o, scl(d, tf));
        const int passes = sp ? sp->passes + 1 : 1;
        if (passes > P->max_depth) return memi;
        v3 direct = mk(0, 0, 0);
        const v3 proxy = neg(d);
        {
            float co, cl, r2, sa;
            nee_t s = sample_area(P, R, pos, proxy, &co, &cl, &r2, &sa);
            if (s.ok && s.pdf > 0.0f) direct = add(direct, scl(s.Le, phase_pdf(med, neg(d), s.dir) * (1.0f / s.pdf)));
        }
        {
            float co;
            nee_t s = sample_env_nee(P, R, pos, proxy, &co);
            if (s.ok && s.pdf > 0.0f) direct = add(direct, scl(s.Le, phase_pdf(med, neg(d), s.dir) * (1.0f / s.pdf)));
        }
        {
            v3 vd, vle;
            float vp = 0.0f;
            if (sample_volume(P, R, pos, proxy, &vd, &vp, &vle) && vp > 0.0f)
                direct = add(direct, scl(vle, phase_pdf(med, neg(d), vd) * (1.0f / vp)));
        }
        const float u1 = prng(R), u2 = prng(R);
        const v3 nd = sample_phase(med, d, u1, u2);
        param_t pm;
        pm.passes = passes;
        pm.medium = midx;
        pm.tp = sp ? sp->tp : mk(1, 1, 1);
        pm.chroma = sp ? sp->chroma : -1; /* rayrenderer.cpp:1781-1784 */
        pm.bsdf_pdf = 0.0f;
        float rrw = 1.0f;
        if (passes >= MIN_RR_DEPTH) {
            float q = maxc(mul(pm.tp, sw));
            q = fminf(RR_MAX_PROB, fmaxf(RR_MIN_PROB, q));
            if (prng(R) >= q) return add(memi, mul(sw, direct));
            rrw = 1.0f / q;
        }
        const v3 no = add(pos, scl(nrm(nd), SURFACE_THICKNESS));
        const v3 ind = scl(trace_path(P, R, no, nd, &pm), rrw);
        return add(memi, mul(sw, add(direct, ind)));
    }
    v3 sw = mk(1, 1, 1);
    if (!med->het) {
        const float surv = expf(-med->hero * maxT);
        if (surv > 0.0f) {
            const v3 Tr = transmittance(med, maxT);
            sw = mk(Tr.x / surv, Tr.y / surv, Tr.z / surv);
        }
    }
    if (h.hit) {
        v3 n = interp_normal(P, &h);
        const v3 g = P->gn[h.prim];
        const v3 gN = dot(g, d) <= 0.0f ? g : neg(g);
        if (dot(d, n) > 0.0f && dot(d, gN) < 0.0f) n = gN;
        return add(memi, mul(sw, shade(P, R, &h, o, d, n, sp)));
    }
    v3 env = sample_environment(P, d);
    if (sp) env = scl(env, env_miss_weight(P, d, sp->bsdf_pdf));
    return add(memi, mul(sw, env));
}

/* RayRenderer::traceEyeRay (rayrenderer.cpp:1254-1297) */
static v3 trace_eye(const prep_t *P, rng_t *R, v3 o, v3 d) {
    const medium_t *gm = medium_at(P, P->d->global_medium);
    if (gm && gm->active) {
        isect_t h = nearest(P, o, d);
        const v3 c = trace_path(P, R, o, d, NULL);
        if (h.hit || !is_zero(c)) return mk(fmaxf(c.x, 0), fmaxf(c.y, 0), fmaxf(c.z, 0));
        return P->back;
    }
    isect_t h = nearest(P, o, d);
    if (h.hit && obj_of(P, h.prim)->visible) {
        const v3 c = shade(P, R, &h, o, d, interp_normal(P, &h), NULL);
        return mk(fmaxf(c.x, 0), fmaxf(c.y, 0), fmaxf(c.z, 0));
    }
    const v3 env = sample_environment(P, d);
    if (!is_zero(env)) return env;
    /* settings.backColor, default (1, 0.95, 0.9) * 0.2 (rayrenderer.h:152) */
    return P->back;
}

/* ======================================================================== */
/* shading systems 0-3 (-d 0..3, rayrenderer.cpp:59-77)                     */
/* ======================================================================== */
/* ThicknessRay (raycommon.h:69-71): the origin pushed SURFACE_THICKNESS
 * along the normalised direction */
static inline v3 thickness_origin(v3 o, v3 d) { return add(o, scl(nrm(d), SURFACE_THICKNESS)); }

/* calcAO (rayrenderer.cpp:2381-2407): settings.samples cosine-weighted
 * directions about the normal (cosineWeightedDirection, an ldsNext2D pair
 * each); a direction whose ThicknessRay meets no triangle with transparency
 * < 0.01 within traceDistance adds its cosine; the sum over the samples */
static float calc_ao(const prep_t *P, rng_t *R, v3 hit, v3 n, float dist) {
    float sum = 0.0f;
    for (int i = 0; i < P->ao_samples; i++) {
        const float r1 = lds1(R), r2 = lds1(R);
        const v3 dir = cosine_dir(n, r1, r2);
        if (!occluded(P, thickness_origin(hit, dir), dir, dist, 4)) sum += fmaxf(dot(dir, n), 0.0f);
    }
    return sum / (float)P->ao_samples;
}

/* lambertTraceLights (rayrenderer.cpp:2319-2335): one area light's
 * traceAreaLight (the light picked by a rand() draw inside
 * trace_area_light), the sum of every point light's tracePointLight, plus
 * settings.worldColor */
static v3 lambert_lights(const prep_t *P, rng_t *R, v3 hit, v3 n) {
    v3 a = mk(0, 0, 0), p = mk(0, 0, 0);
    if (P->nlights > 0) a = add(a, trace_area_light(P, R, hit, n));
    const uint32_t np = P->d->plights ? P->d->nplights : 0;
    for (uint32_t i = 0; i < np; i++) p = add(p, trace_point_light(P, &P->d->plights[i], hit, n));
    return add(add(a, p), P->world);
}

/* RaySimpleShaderProvider and RayAmbientOcclusionShaderProvider
 * (rayrenderer.h:446-475; traceDistance RAY_MAX_DISTANCE),
 * LambertShaderProvider and LambertWithAOShaderProvider (lambert.cpp:14-67):
 * a transparent surface continues with tracePath of the ThicknessRay along
 * the incoming direction, with the caller's shaderParam */
static v3 preview_shade(const prep_t *P, rng_t *R, const isect_t *h, v3 in_d, v3 n, param_t *sp) {
    const or_object *m = obj_of(P, h->prim);
    if (P->shader == 0) {
        const float c = 0.1f + fmaxf(dot(nrm(mk(1.0f, 1.0f, 1.0f)), n), 0.0f);
        return mk(c, c, c);
    }
    if (P->shader == 1) {
        const float c = clamp01(calc_ao(P, R, h->pos, n, RAY_MAX_DISTANCE));
        return mk(c, c, c);
    }
    if (m->emission > 0.0f) return scl(ld3(m->color), m->emission);
    const v3 light = scl(lambert_lights(P, R, h->pos, n), 0.3f);
    v3 c;
    if (P->shader == 2) c = light;
    else c = add(scl(light, 0.2f), scl(scl(light, 0.8f), calc_ao(P, R, h->pos, n, 2.0f)));
    if (P->color_sampling) c = mul(mul(c, ld3(m->color)), albedo_tex(P, h));
    if (m->transparency > 0.0f) {
        const v3 next = scl(trace_path(P, R, thickness_origin(h->pos, in_d), in_d, sp), m->transparency);
        if (P->shader == 2)
            c = add(scl(mul(ld3(m->color), add(light, mk(0.5f, 0.5f, 0.5f))), 1.0f - m->transparency), next);
        else
            c = add(c, next);
    }
    return c;
}

/* ======================================================================== */
/* mode 1: single-lobe iterative twin of the GPU wavefront                  */
/* ======================================================================== */
typedef struct {
    v3 o, d, beta, tp, L;
    float bsdf_pdf;
    int passes; /* passes of the spawning param; -1 = eye ray */
    int medium;
    int chroma; /* chromatic dispersion band, -1 = none chosen */
    int alive;
} path_t;

/* add an NEE contribution after its shadow test (already tested) */
static inline void add_L(path_t *p, v3 c) { p->L = add(p->L, c); }

static void twin_surface(const prep_t *P, rng_t *R, path_t *p, const isect_t *h, v3 n);

/* TransparencyShader as one path step.  bsdfSampledPdf and the dispersion
 * band are carried over from the parent on the front-face branch and start
 * at 0 and none on the back-face and eye branches, where the reference
 * builds a fresh BSDFParam. */
static void twin_transparency(const prep_t *P, path_t *p, const isect_t *h, v3 n, float beta_scale,
                              int fresh_param) {
    const or_object *m = obj_of(P, h->prim);
    p->beta = scl(p->beta, beta_scale);
    p->tp = scl(p->tp, m->transparency);
    p->medium = crossed_medium(P, m, dot(p->d, n) < 0.0f, p->medium);
    const v3 gN = geom_normal(P, h->prim, n);
    p->o = surface_origin(h->pos, p->d, gN);
    if (fresh_param) {
        p->bsdf_pdf = 0.0f;
        p->chroma = -1;
    }
}

static void twin_step(const prep_t *P, rng_t *R, path_t *p) {
    const int eye = p->passes < 0;
    isect_t h = nearest(P, p->o, p->d);
    const medium_t *med = medium_at(P, p->medium);
    if (!med || !med->active) {
        if (h.hit && (!eye || obj_of(P, h.prim)->visible)) {
            twin_surface(P, R, p, &h, interp_normal(P, &h));
            return;
        }
        v3 env = sample_environment(P, p->d);
        if (eye) {
            if (is_zero(env)) env = P->back;
        } else {
            env = scl(env, env_miss_weight(P, p->d, p->bsdf_pdf));
        }
        add_L(p, mul(p->beta, env));
        p->alive = 0;
        return;
    }
    if (med->haze) {
        /* the bent ray continues in the outer medium as the path's next
         * event, with the same passes, throughput and BSDF pdf */
        v3 bo, bd;
        bend_ray(med, p->o, p->d, fminf(h.hit ? h.t : RAY_MAX_DISTANCE, 1000.0f), &bo, &bd);
        p->o = bo;
        p->d = bd;
        p->medium = P->d->global_medium;
        return;
    }
    const float maxT = h.hit ? h.t : RAY_MAX_DISTANCE;
    int scattered;
    float tf = 0.0f, den = 1.0f;
    if (med->het) {
        scattered = delta_track(med, R, p->o, p->d, maxT, &tf, &den);
        if (!scattered) tf = maxT;
    } else {
        tf = free_flight(med, prng(R));
        scattered = tf < maxT;
    }
    add_L(p, mul(p->beta, emission_integral(med, R, p->o, p->d, scattered ? tf : maxT)));
    if (scattered) {
        v3 sw = mk(0, 0, 0);
        if (med->het) {
            const float inv = med->t_max_hero > 0.0f ? 1.0f / med->t_max_hero : 0.0f;
            const v3 ss = scl(med->sigma_s_eff, den);
            sw = mk(ss.x * inv, ss.y * inv, ss.z * inv);
        } else {
            const float ph = med->hero * expf(-med->hero * tf);
            const v3 Tr = transmittance(med, tf);
            if (ph > 0.0f)
                sw = mk(med->sigma_s_eff.x * Tr.x / ph, med->sigma_s_eff.y * Tr.y / ph,
                        med->sigma_s_eff.z * Tr.z / ph);
        }
        if (is_zero(sw)) { p->alive = 0; return; }
        const v3 pos = add(p->o, scl(p->d, tf));
        const int passes = eye ? 1 : p->passes + 1;
        if (passes > P->max_depth) { p->alive = 0; return; }
        const v3 bsw = mul(p->beta, sw);
        const v3 proxy = neg(p->d);
        {
            float co, cl, r2, sa;
            nee_t s = sample_area(P, R, pos, proxy, &co, &cl, &r2, &sa);
            if (s.ok && s.pdf > 0.0f) add_L(p, mul(bsw, scl(s.Le, phase_pdf(med, neg(p->d), s.dir) * (1.0f / s.pdf))));
        }
        {
            float co;
            nee_t s = sample_env_nee(P, R, pos, proxy, &co);
            if (s.ok && s.pdf > 0.0f) add_L(p, mul(bsw, scl(s.Le, phase_pdf(med, neg(p->d), s.dir) * (1.0f / s.pdf))));
        }
        {
            v3 vd, vle;
            float vp = 0.0f;
            if (sample_volume(P, R, pos, proxy, &vd, &vp, &vle) && vp > 0.0f)
                add_L(p, mul(bsw, scl(vle, phase_pdf(med, neg(p->d), vd) * (1.0f / vp))));
        }
        const float u1 = prng(R), u2 = prng(R);
        const v3 nd = sample_phase(med, p->d, u1, u2);
        float rrw = 1.0f;
        if (passes >= MIN_RR_DEPTH) {
            float q = maxc(mul(p->tp, sw));
            q = fminf(RR_MAX_PROB, fmaxf(RR_MIN_PROB, q));
            if (prng(R) >= q) { p->alive = 0; return; }
            rrw = 1.0f / q;
        }
        p->beta = scl(bsw, rrw);
        p->passes = passes;
        p->bsdf_pdf = 0.0f;
        p->o = add(pos, scl(nrm(nd), SURFACE_THICKNESS));
        p->d = nd;
        return;
    }
    v3 sw = mk(1, 1, 1);
    if (!med->het) {
        const float surv = expf(-med->hero * maxT);
        if (surv > 0.0f) {
            const v3 Tr = transmittance(med, maxT);
            sw = mk(Tr.x / surv, Tr.y / surv, Tr.z / surv);
        }
    }
    p->beta = mul(p->beta, sw);
    if (h.hit) {
        v3 n = interp_normal(P, &h);
        const v3 g = P->gn[h.prim];
        const v3 gN = dot(g, p->d) <= 0.0f ? g : neg(g);
        if (dot(p->d, n) > 0.0f && dot(p->d, gN) < 0.0f) n = gN;
        twin_surface(P, R, p, &h, n);
        return;
    }
    v3 env = sample_environment(P, p->d);
    if (!eye) env = scl(env, env_miss_weight(P, p->d, p->bsdf_pdf));
    add_L(p, mul(p->beta, env));
    p->alive = 0;
}

static void twin_mix(const prep_t *P, rng_t *R, path_t *p, const isect_t *h, v3 n) {
    const or_object *m = obj_of(P, h->prim);
    const float wd = 1.0f - m->glossy - m->refraction;
    const float w0 = wd > 0.00001f ? wd : 0.0f;
    const float w1 = m->glossy > 0.00001f ? m->glossy : 0.0f;
    const float w2 = m->refraction > 0.00001f ? m->refraction : 0.0f;
    const float W = w0 + w1 + w2;
    if (!(W > 0.0f)) { p->alive = 0; return; }
    const float u = prng(R) * W;
    int lobe = (u < w0) ? 0 : (u < w0 + w1 ? 1 : 2);
    if (lobe == 2 && w2 == 0.0f) lobe = w1 > 0.0f ? 1 : 0;
    if (lobe == 1 && w1 == 0.0f) lobe = 0;
    const float wl = lobe == 0 ? w0 : (lobe == 1 ? w1 : w2);
    p->beta = scl(p->beta, W);
    p->tp = scl(p->tp, wl);
    const v3 hit = h->pos;
    if (lobe == 0) {
        const v3 alb = P->color_sampling ? mul(ld3(m->color), albedo_tex(P, h)) : mk(1.0f, 1.0f, 1.0f);
        const float r1 = lds1(R), r2 = lds1(R);
        const v3 dir = cosine_dir(n, r1, r2);
        const float pdf = fmaxf(0.0f, dot(dir, n)) * (float)(1.0 / M_PI);
        const v3 ba = mul(p->beta, alb);
        add_L(p, mul(ba, trace_light(P, R, hit, n)));
        add_L(p, mul(ba, trace_env_light(P, R, hit, n, pdf)));
        {
            v3 vd, vle;
            float vp = 0.0f;
            if (sample_volume(P, R, hit, n, &vd, &vp, &vle) && vp > 0.0f) {
                const float cl = fmaxf(0.0f, dot(vd, n));
                add_L(p, mul(ba, scl(vle, cl * (float)(1.0 / M_PI) / vp)));
            }
        }
        p->beta = ba;
        p->tp = mul(p->tp, alb);
        p->bsdf_pdf = pdf;
        p->o = surface_origin(hit, dir, geom_normal(P, h->prim, n));
        p->d = dir;
        return;
    }
    if (lobe == 1) {
        const v3 alb = P->color_sampling ? mul(ld3(m->color), albedo_tex(P, h)) : ld3(m->color);
        if (m->roughness < 1e-3f) {
            const v3 F0 = (m->metallic >= 1.0f) ? alb
                          : add(scl(alb, m->metallic), scl(mk(0.04f, 0.04f, 0.04f), 1.0f - m->metallic));
            const v3 gN = geom_normal(P, h->prim, n);
            v3 r = reflect(p->d, n);
            if (dot(r, gN) <= 0.0f) r = reflect(p->d, gN);
            const float ci = fmaxf(0.0f, -dot(p->d, n));
            const v3 F = schlick_f(F0, ci);
            p->beta = mul(p->beta, F);
            p->tp = mul(p->tp, F);
            p->bsdf_pdf = 0.0f;
            p->o = surface_origin(hit, r, gN);
            p->d = r;
            return;
        }
        ggx_t G;
        if (!ggx_setup(m, alb, n, p->d, &G)) { p->alive = 0; return; }
        {
            float co, cl, r2, sa;
            nee_t s = sample_area(P, R, hit, n, &co, &cl, &r2, &sa);
            v3 fc;
            float pb;
            if (s.ok && eval_glossy(&G, s.dir, &fc, &pb)) {
                const float pl2 = s.pdf * s.pdf, pb2 = pb * pb;
                add_L(p, mul(p->beta, scl(mul(s.Le, fc), (pl2 / (pl2 + pb2)) / s.pdf)));
            }
        }
        {
            float co;
            nee_t s = sample_env_nee(P, R, hit, n, &co);
            v3 fc;
            float pb;
            if (s.ok && eval_glossy(&G, s.dir, &fc, &pb)) {
                const float pe2 = s.pdf * s.pdf, pb2 = pb * pb;
                add_L(p, mul(p->beta, scl(mul(s.Le, fc), (pe2 / (pe2 + pb2)) / s.pdf)));
            }
        }
        {
            v3 vd, vle, fc;
            float vp = 0.0f, pb;
            if (sample_volume(P, R, hit, n, &vd, &vp, &vle) && vp > 0.0f && eval_glossy(&G, vd, &fc, &pb))
                add_L(p, mul(p->beta, scl(mul(vle, fc), 1.0f / vp)));
        }
        const float u1 = prng(R), u2 = prng(R);
        const v3 H = sample_vndf(G.V, G.ax, G.ay, u1, u2);
        const v3 Ll = sub(scl(H, 2.0f * dot(G.V, H)), G.V);
        if (Ll.z <= 0.0f) { p->alive = 0; return; }
        const v3 L = add(add(scl(G.t, Ll.x), scl(G.b, Ll.y)), scl(n, Ll.z));
        const v3 gN = geom_normal(P, h->prim, n);
        if (dot(L, gN) <= 0.0f) { p->alive = 0; return; }
        const float vdh = fmaxf(0.0f, dot(G.V, H));
        const v3 F = schlick_f(G.F0, vdh);
        const v3 wgt = scl(F, smith_g1(Ll, G.ax, G.ay));
        p->beta = mul(p->beta, wgt);
        p->tp = mul(p->tp, wgt);
        p->bsdf_pdf = G.g1v * ggx_d(H, G.ax, G.ay) / (4.0f * G.V.z);
        p->o = surface_origin(hit, L, gN);
        p->d = L;
        return;
    }
    {
        const float ct = fabsf(dot(p->d, n));
        v3 mask;
        const float ior = dispersion_ior(m, R, &p->chroma, &mask);
        const float fr = fresnel_schlick(ct, ior);
        const int refl = prng(R) < fr;
        v3 dir = refl ? reflect(p->d, n) : refract_dir(p->d, n, ior);
        if (m->roughness > 0.0f) {
            const float u1 = prng(R), u2 = prng(R);
            dir = nrm(add(dir, scl(hemisphere_dir(n, u1, u2), m->roughness)));
        }
        const v3 gN = geom_normal(P, h->prim, n);
        if (refl && dot(dir, gN) <= 0.0f) dir = reflect(p->d, gN);
        const v3 col = ld3(m->color);
        p->beta = mul(mul(p->beta, col), mask);
        p->tp = mul(p->tp, col);
        p->bsdf_pdf = 0.0f;
        const float di = dot(p->d, n), dout = dot(dir, n);
        if (di * dout > 0.0f) p->medium = crossed_medium(P, m, di < 0.0f, p->medium);
        p->o = surface_origin(hit, dir, gN);
        p->d = dir;
    }
}

static void twin_surface(const prep_t *P, rng_t *R, path_t *p, const isect_t *h, v3 n) {
    const or_object *m = obj_of(P, h->prim);
    const int eye = p->passes < 0;
    if (m->emission > 0.0f) {
        const v3 E = scl(ld3(m->color), m->emission);
        if (!eye && p->bsdf_pdf > 0.0f) {
            const float cl = -dot(p->d, n);
            if (cl > 0.0f) {
                const v3 dd = sub(h->pos, p->o);
                const float r2 = dd.x * dd.x + dd.y * dd.y + dd.z * dd.z;
                const float pl = r2 / (cl * P->sampled_area[h->prim]);
                const float pb = p->bsdf_pdf;
                const float pb2 = pb * pb, pl2 = pl * pl;
                add_L(p, mul(p->beta, scl(E, pb2 / (pb2 + pl2))));
            }
        } else {
            add_L(p, mul(p->beta, E));
        }
        p->alive = 0;
        return;
    }
    if (dot(p->d, n) > 0.0f) {
        if (m->transparency > 0.001f) {
            if (!eye && p->passes + 1 <= P->max_depth) {
                p->passes = p->passes + 1;
                twin_transparency(P, p, h, n, m->transparency, 1);
                return;
            }
            p->alive = 0;
            return;
        } else if (m->refraction < 0.001f && m->glossy > 0.001f) {
            p->alive = 0;
            return;
        }
    }
    if (!eye) {
        if (p->passes + 1 >= P->max_depth) {
            if (1.0f - m->glossy - m->refraction > 0.00001f && P->color_sampling)
                add_L(p, mul(mul(p->beta, mul(ld3(m->color), albedo_tex(P, h))), trace_light(P, R, h->pos, n)));
            p->alive = 0;
            return;
        }
        p->passes = p->passes + 1;
        if (p->passes >= MIN_RR_DEPTH) {
            float q = maxc(p->tp);
            q = fminf(RR_MAX_PROB, fmaxf(RR_MIN_PROB, q));
            if (prng(R) >= q) { p->alive = 0; return; }
            p->beta = scl(p->beta, 1.0f / q);
        }
        if (m->transparency > 0.001f) {
            twin_transparency(P, p, h, n, m->transparency, 0);
            return;
        }
        twin_mix(P, R, p, h, n);
        return;
    }
    p->passes = 0;
    if (m->transparency > 0.01f) {
        if (prng(R) < m->transparency) {
            twin_transparency(P, p, h, n, 1.0f, 1);
            /* the reference evaluates the Mix branch first on the same
             * BSDFParam (rayrenderer.cpp:2889-2894), and its refraction lobe
             * leaves the dispersion band it drew there (bsdf.cpp:402-407
             * set chromaChannel, nothing restores it): the transparency
             * path carries a uniformly drawn band, without the 3x mask */
            if (m->refraction > 0.00001f && m->chroma_dispersion > 0.0f)
                p->chroma = (int)fminf(2.0f, floorf(prng(R) * 3.0f));
            return;
        }
    }
    twin_mix(P, R, p, h, n);
}

/* ======================================================================== */
/* camera + pixel loop                                                      */
/* ======================================================================== */
static void camera_ray(const prep_t *P, rng_t *R, int x, int y, int aa, v3 *o, v3 *d) {
    const float dx = ((float)x + 0.5f - P->half_w) * P->view_scale;
    const float dy = -((float)y + 0.5f - P->half_h) * P->view_scale;
    const float jx = lds1(R), jy = lds1(R);
    const float jxo = aa ? (jx - 0.5f) : 0.0f;
    const float jyo = aa ? (jy - 0.5f) : 0.0f;
    if (P->dof >= 0.001f && P->aperture > 0.0f) {
        const float px = dx + jxo * P->view_scale;
        const float py = dy - jyo * P->view_scale;
        const v3 focal = mk(px * P->dof, py * P->dof, -P->dof);
        const float du = lds1(R), dv = lds1(R);
        float ox, oy;
        if (P->blades >= 3) {
            const int N = P->blades;
            const float u0 = du * (float)N;
            const int wedge = (int)fminf((float)(N - 1), floorf(u0));
            float s = u0 - (float)wedge, t = dv;
            if (s + t > 1.0f) { s = 1.0f - s; t = 1.0f - t; }
            const float step = 2.0f * PI_F / (float)N;
            const float a0 = P->ap_rot + step * (float)wedge;
            const float a1 = a0 + step;
            ox = (cosf(a0) * s + cosf(a1) * t) * P->aperture;
            oy = (sinf(a0) * s + sinf(a1) * t) * P->aperture;
        } else {
            const float a = 2.0f * du - 1.0f, b = 2.0f * dv - 1.0f;
            float r, phi;
            if (a == 0.0f && b == 0.0f) { r = 0.0f; phi = 0.0f; }
            else if (a * a > b * b) { r = a; phi = (float)(M_PI / 4.0) * (b / a); }
            else { r = b; phi = (float)(M_PI / 2.0) - (float)(M_PI / 4.0) * (a / b); }
            ox = r * cosf(phi) * P->aperture;
            oy = r * sinf(phi) * P->aperture;
        }
        *o = mk(ox, oy, 0.0f);
        *d = nrm(sub(focal, *o));
    } else {
        *o = mk(0, 0, 0);
        *d = nrm(mk(dx + jxo * P->view_scale, dy - jyo * P->view_scale, -1.0f));
    }
}

static v3 render_sample(const prep_t *P, const or_settings *S, int mode, int x, int y, int s) {
    rng_t R;
    memset(&R, 0, sizeof R);
    R.x = x;
    R.y = y;
    R.sample = s;
    R.dim = 0;
    R.pix = (uint32_t)y * S->width + (uint32_t)x;
    R.key0 = S->seed;
    R.key1 = S->job;
    R.blk_id = 0xFFFFFFFFu;
    v3 o, d;
    camera_ray(P, &R, x, y, S->antialias, &o, &d);
    v3 c;
    if (mode == 0 || P->shader != 5) {
        /* shading systems 0-3 have no lobe choice: one restatement */
        c = trace_eye(P, &R, o, d);
    } else {
        path_t p;
        memset(&p, 0, sizeof p);
        p.o = o;
        p.d = d;
        p.beta = mk(1, 1, 1);
        p.tp = mk(1, 1, 1);
        p.L = mk(0, 0, 0);
        p.passes = -1;
        p.medium = P->d->global_medium;
        p.chroma = -1;
        p.alive = 1;
        const medium_t *gm = medium_at(P, P->d->global_medium);
        int eye_hit = 0;
        if (gm && gm->active) eye_hit = nearest(P, o, d).hit;
        int guard = 0;
        while (p.alive && guard++ < 4 * P->max_depth + 64) twin_step(P, &R, &p);
        c = mk(fmaxf(p.L.x, 0.0f), fmaxf(p.L.y, 0.0f), fmaxf(p.L.z, 0.0f));
        if (gm && gm->active && !eye_hit && is_zero(p.L)) c = P->back;
    }
    if (S->firefly_clamp > 0.0f) {
        c.x = fminf(c.x, S->firefly_clamp);
        c.y = fminf(c.y, S->firefly_clamp);
        c.z = fminf(c.z, S->firefly_clamp);
    }
    return c;
}

/* ------------------------------------------------------------------------ */
/* scene preparation                                                         */
/* ------------------------------------------------------------------------ */
static int prepare(prep_t *P, const or_scene_desc *d, const or_settings *S) {
    memset(P, 0, sizeof *P);
    P->d = d;
    if (d->ntex) {
        P->tex_lin = (float **)calloc(d->ntex, sizeof(float *));
        for (uint32_t k = 0; k < d->ntex; k++) {
            const size_t m = 3 * (size_t)d->tex[k].w * d->tex[k].h;
            P->tex_lin[k] = (float *)malloc(m * sizeof(float));
            for (size_t i = 0; i < m; i++)
                P->tex_lin[k][i] = d->tex[k].srgb ? srgb_to_linear(d->tex[k].rgb[i]) : d->tex[k].rgb[i];
        }
    }
    const uint32_t n = d->ntris;
    P->nodes = (or_node *)calloc(n ? 2 * (size_t)n : 1, sizeof(or_node));
    P->perm = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    P->nnodes = or_bvh_build(d->v9, n, P->nodes, P->perm);
    P->gn = (v3 *)calloc(n ? n : 1, sizeof(v3));
    P->ti_a = (float *)calloc(n ? n : 1, sizeof(float));
    P->area = (float *)calloc(n ? n : 1, sizeof(float));
    P->sampled_area = (float *)calloc(n ? n : 1, sizeof(float));
    for (uint32_t t = 0; t < n; t++) {
        const v3 a = tri_v(P, t, 0), b = tri_v(P, t, 1), c = tri_v(P, t, 2);
        const v3 pd = cross(sub(b, a), sub(c, b));  /* precalc (raycommon.cpp:128-149) */
        P->ti_a[t] = 1.0f / len(cross(sub(a, b), sub(a, c)));
        P->gn[t] = nrm(pd);
        P->area[t] = 0.5f * len(cross(sub(b, a), sub(c, a)));
    }
    /* lights: emissive objects, in object order (transformScene DFS order) */
    uint32_t *cnt = (uint32_t *)calloc(d->nobj ? d->nobj : 1, sizeof(uint32_t));
    for (uint32_t t = 0; t < n; t++) cnt[d->obj_of_tri[t]]++;
    P->light_obj = (uint32_t *)calloc(d->nobj ? d->nobj : 1, sizeof(uint32_t));
    P->light_first = (uint32_t *)calloc(d->nobj ? d->nobj : 1, sizeof(uint32_t));
    P->light_count = (uint32_t *)calloc(d->nobj ? d->nobj : 1, sizeof(uint32_t));
    P->light_tris = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    uint32_t nl = 0, nt = 0;
    for (uint32_t o = 0; o < d->nobj; o++) {
        if (!(d->objs[o].emission > 0.0f) || cnt[o] == 0) continue;
        P->light_obj[nl] = o;
        P->light_first[nl] = nt;
        P->light_count[nl] = cnt[o];
        for (uint32_t t = 0; t < n; t++)
            if (d->obj_of_tri[t] == o) {
                P->light_tris[nt++] = t;
                P->sampled_area[t] = (float)cnt[o] * P->area[t];
            }
        nl++;
    }
    P->nlights = nl;
    free(cnt);
    /* media (HomogeneousMedium::prepare, medium.cpp:18-57) */
    P->media = (medium_t *)calloc(d->nmedia ? d->nmedia : 1, sizeof(medium_t));
    for (uint32_t i = 0; i < d->nmedia; i++) prep_medium(&P->media[i], &d->media[i]);
    /* a heat-haze global medium (where tracePath recurses without end, and
     * which the library refuses) renders as no medium */
    if (d->global_medium >= 0 && P->media[d->global_medium].haze) {
        P->media[d->global_medium].haze = 0;
        P->media[d->global_medium].active = 0;
    }
    /* envmap CDF (scene.cpp:400-463) */
    P->has_env = d->env.rgb != NULL && d->env.w > 0 && d->env.h > 0;
    if (P->has_env) {
        const int W = (int)d->env.w, H = (int)d->env.h;
        P->env_cond = (float *)calloc((size_t)H * (W + 1), sizeof(float));
        P->env_marg = (float *)calloc((size_t)H + 1, sizeof(float));
        float *roww = (float *)calloc((size_t)H, sizeof(float));
        for (int y = 0; y < H; y++) {
            const float v = ((float)y + 0.5f) / (float)H;
            const float st = sinf(PI_F * v);
            float rs = 0.0f;
            P->env_cond[(size_t)y * (W + 1)] = 0.0f;
            for (int x = 0; x < W; x++) {
                rs += lum(env_texel(P, x, y)) * st;
                P->env_cond[(size_t)y * (W + 1) + x + 1] = rs;
            }
            roww[y] = rs;
        }
        for (int y = 0; y < H; y++)
            if (roww[y] > 0.0f) {
                const float inv = 1.0f / roww[y];
                for (int x = 1; x <= W; x++) P->env_cond[(size_t)y * (W + 1) + x] *= inv;
            }
        P->env_marg[0] = 0.0f;
        for (int y = 0; y < H; y++) P->env_marg[y + 1] = P->env_marg[y] + roww[y];
        P->env_total = P->env_marg[H];
        if (P->env_total > 0.0f) {
            const float inv = 1.0f / P->env_total;
            for (int y = 0; y <= H; y++) P->env_marg[y] *= inv;
        }
        free(roww);
    }
    /* camera (initRenderThreadContext, rayrenderer.cpp:97-132) */
    const float W = (float)S->width, H = (float)S->height;
    P->half_w = W * 0.5f;
    P->half_h = H * 0.5f;
    const float fov = fminf(fmaxf(d->cam.fov_deg, 1.0f), 179.0f);
    const float th = tanf(fov * (float)(M_PI / 180.0) * 0.5f);
    P->view_scale = (2.0f * th) / H;
    P->dof = d->cam.depth_of_field;
    P->aperture = d->cam.aperture > 0.0f ? 1.0f / d->cam.aperture : 0.0f;
    P->blades = d->cam.aperture_blades;
    P->ap_rot = d->cam.aperture_rotation_deg * (float)(M_PI / 180.0);
    P->max_depth = (int)S->max_depth;
    P->color_sampling = S->color_sampling != 0;
    P->back = mk(S->back_color[0], S->back_color[1], S->back_color[2]);
    P->shader = S->shader;
    P->world = mk(S->world_color[0], S->world_color[1], S->world_color[2]);
    P->ao_samples = (int)(S->ao_samples ? S->ao_samples : (S->spp ? S->spp : 1));
    return 0;
}

static void release(prep_t *P) {
    free(P->nodes); free(P->perm); free(P->gn); free(P->ti_a); free(P->area); free(P->sampled_area);
    free(P->light_obj); free(P->light_first); free(P->light_count); free(P->light_tris);
    free(P->media); free(P->env_cond); free(P->env_marg);
    if (P->tex_lin) {
        for (uint32_t k = 0; k < P->d->ntex; k++) free(P->tex_lin[k]);
        free(P->tex_lin);
    }
}

typedef struct {
    const prep_t *P;
    const or_settings *S;
    int mode;
    float *sum, *sumsq;
    const uint8_t *active_tile; /* nullable: per 32x32 tile */
    atomic_int next_row;
} render_job;

static void *render_worker(void *arg) {
    render_job *j = (render_job *)arg;
    const or_settings *S = j->S;
    const int tiles_x = (int)((S->width + 31) / 32);
    for (;;) {
        const int y = atomic_fetch_add(&j->next_row, 1);
        if (y >= (int)S->height) break;
        for (int x = 0; x < (int)S->width; x++) {
            if (j->active_tile && !j->active_tile[(y / 32) * tiles_x + x / 32]) continue;
            v3 s = mk(0, 0, 0), s2 = mk(0, 0, 0);
            for (uint32_t k = 0; k < S->spp; k++) {
                const v3 c = render_sample(j->P, S, j->mode, x, y, (int)(S->sample_start + k));
                s = add(s, c);
                s2 = add(s2, mul(c, c));
            }
            const size_t i = (size_t)y * S->width + (size_t)x;
            j->sum[3 * i] += s.x; j->sum[3 * i + 1] += s.y; j->sum[3 * i + 2] += s.z;
            if (j->sumsq) {
                j->sumsq[3 * i] += s2.x; j->sumsq[3 * i + 1] += s2.y; j->sumsq[3 * i + 2] += s2.z;
            }
        }
    }
    return NULL;
}

int or_render(const or_scene_desc *desc, const or_settings *S, int mode, int nthreads,
              const uint8_t *active_tile, float *sum_rgb, float *sumsq_rgb) {
    prep_t P;
    prepare(&P, desc, S);
    render_job j;
    j.P = &P; j.S = S; j.mode = mode; j.sum = sum_rgb; j.sumsq = sumsq_rgb; j.active_tile = active_tile;
    atomic_init(&j.next_row, 0);
    if (nthreads <= 1) render_worker(&j);
    else {
        pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
        for (int i = 0; i < nthreads; i++) pthread_create(&th[i], NULL, render_worker, &j);
        for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
        free(th);
    }
    release(&P);
    return 0;
}

/* One sample's radiance for a given pixel (diagnostics / tests). */
void or_render_one(const or_scene_desc *desc, const or_settings *S, int mode, int x, int y, int s,
                   float out[3]) {
    prep_t P;
    prepare(&P, desc, S);
    const v3 c = render_sample(&P, S, mode, x, y, s);
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
    release(&P);
}

/* ======================================================================== */
/* guide buffers + A-Trous denoiser (rayrenderer.cpp:819-844, 2536-2772)    */
/* ======================================================================== */
/* Primary-hit AOVs per pixel (accumulatePixelSamples :819-844), traced along
 * the pixel-centre camera ray: normal*0.5+0.5 (a = 1 on a hit), material
 * colour (a = 1), depth 1 - clamp(sqrt(dist / view_far)); on a miss normal
 * 0.5 (a = 0), backColor (a = 0), depth 0.  Outputs are RGBA float. */
void or_aov(const or_scene_desc *desc, const or_settings *S, float *normal_rgba, float *albedo_rgba,
            float *depth) {
    prep_t P;
    prepare(&P, desc, S);
    for (uint32_t y = 0; y < S->height; y++)
        for (uint32_t x = 0; x < S->width; x++) {
            const size_t i = (size_t)y * S->width + x;
            const float dx = ((float)x + 0.5f - P.half_w) * P.view_scale;
            const float dy = -((float)y + 0.5f - P.half_h) * P.view_scale;
            const v3 d = nrm(mk(dx, dy, -1.0f));
            const isect_t h = nearest(&P, mk(0, 0, 0), d);
            if (h.hit) {
                const v3 n = interp_normal(&P, &h);
                const or_object *m = obj_of(&P, h.prim);
                normal_rgba[4 * i] = n.x * 0.5f + 0.5f;
                normal_rgba[4 * i + 1] = n.y * 0.5f + 0.5f;
                normal_rgba[4 * i + 2] = n.z * 0.5f + 0.5f;
                normal_rgba[4 * i + 3] = 1.0f;
                albedo_rgba[4 * i] = m->color[0];
                albedo_rgba[4 * i + 1] = m->color[1];
                albedo_rgba[4 * i + 2] = m->color[2];
                albedo_rgba[4 * i + 3] = 1.0f;
                float dd = len(h.pos) / desc->cam.view_far;
                dd = sqrtf(dd);
                depth[i] = 1.0f - fminf(fmaxf(dd, 0.0f), 1.0f);
            } else {
                normal_rgba[4 * i] = normal_rgba[4 * i + 1] = normal_rgba[4 * i + 2] = 0.5f;
                normal_rgba[4 * i + 3] = 0.0f;
                albedo_rgba[4 * i] = S->back_color[0];
                albedo_rgba[4 * i + 1] = S->back_color[1];
                albedo_rgba[4 * i + 2] = S->back_color[2];
                albedo_rgba[4 * i + 3] = S->back_color[3];
                depth[i] = 0.0f;
            }
        }
    release(&P);
}

static inline float lum3(const float *c) { return 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2]; }

/* atrousPass (rayrenderer.cpp:2536-2608) */
static void atrous(const float *src, float *dst, const float *nrm_rgba, const float *depth, int w, int h,
                   int step, float sc, float sn, float sd) {
    static const float k[5] = {1.0f / 16.0f, 1.0f / 4.0f, 3.0f / 8.0f, 1.0f / 4.0f, 1.0f / 16.0f};
    const float inv_c2 = 1.0f / (2.0f * sc * sc + 1e-8f);
    const float dscale = 1.0f / (sd * (float)step + 1e-8f);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const size_t c = (size_t)y * w + x;
            const float *cc = src + 4 * c;
            const float cn[3] = {nrm_rgba[4 * c] * 2.0f - 1.0f, nrm_rgba[4 * c + 1] * 2.0f - 1.0f,
                                 nrm_rgba[4 * c + 2] * 2.0f - 1.0f};
            const float cd = depth[c];
            const float cl = lum3(cc);
            float sx = 0.0f, sy = 0.0f, sz = 0.0f, ws = 0.0f;
            for (int ky = 0; ky < 5; ky++) {
                const int ny = y + (ky - 2) * step;
                if (ny < 0 || ny >= h) continue;
                for (int kx = 0; kx < 5; kx++) {
                    const int nx = x + (kx - 2) * step;
                    if (nx < 0 || nx >= w) continue;
                    const size_t s = (size_t)ny * w + nx;
                    const float *cs = src + 4 * s;
                    const float sn3[3] = {nrm_rgba[4 * s] * 2.0f - 1.0f, nrm_rgba[4 * s + 1] * 2.0f - 1.0f,
                                          nrm_rgba[4 * s + 2] * 2.0f - 1.0f};
                    const float ld = lum3(cs) - cl;
                    const float wc = expf(-(ld * ld) * inv_c2);
                    float nd = sn3[0] * cn[0] + sn3[1] * cn[1] + sn3[2] * cn[2];
                    if (nd < 0.0f) nd = 0.0f;
                    const float wn = powf(nd, sn);
                    const float wd = expf(-fabsf(depth[s] - cd) * dscale);
                    const float wt = k[kx] * k[ky] * wc * wn * wd;
                    sx += cs[0] * wt;
                    sy += cs[1] * wt;
                    sz += cs[2] * wt;
                    ws += wt;
                }
            }
            float *o = dst + 4 * c;
            if (ws > 1e-8f) { o[0] = sx / ws; o[1] = sy / ws; o[2] = sz / ws; }
            else { o[0] = cc[0]; o[1] = cc[1]; o[2] = cc[2]; }
            o[3] = cc[3];
        }
}

/* denoiseImage (rayrenderer.cpp:2649-2772): firefly pre-pass + albedo
 * demodulation, `levels` A-Trous passes, remodulation and blend. */
void or_denoise(int w, int h, const float *noisy, const float *nrm_rgba, const float *albedo_rgba,
                const float *depth, int levels, float sigma_c, float sigma_n, float sigma_d, float intensity,
                float *out) {
    const size_t n = (size_t)w * h;
    float *a = (float *)calloc(4 * n, sizeof(float));
    float *b = (float *)calloc(4 * n, sizeof(float));
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const size_t i = (size_t)y * w + x;
            float c[4] = {noisy[4 * i], noisy[4 * i + 1], noisy[4 * i + 2], noisy[4 * i + 3]};
            float mx = 0.0f;
            for (int dy = -1; dy <= 1; dy++) {
                const int ny = y + dy;
                if (ny < 0 || ny >= h) continue;
                for (int dx = -1; dx <= 1; dx++) {
                    if (dx == 0 && dy == 0) continue;
                    const int nx = x + dx;
                    if (nx < 0 || nx >= w) continue;
                    const float l = lum3(noisy + 4 * ((size_t)ny * w + nx));
                    if (l > mx) mx = l;
                }
            }
            const float cl = lum3(c);
            const float cap = 1.5f * mx + 0.01f;
            if (cl > cap) {
                const float s = cap / cl;
                c[0] *= s; c[1] *= s; c[2] *= s;
            }
            const float *al = albedo_rgba + 4 * i;
            if (al[3] > 0.5f) {
                a[4 * i] = fminf(c[0] / fmaxf(al[0], 0.3f), 3.0f);
                a[4 * i + 1] = fminf(c[1] / fmaxf(al[1], 0.3f), 3.0f);
                a[4 * i + 2] = fminf(c[2] / fmaxf(al[2], 0.3f), 3.0f);
                a[4 * i + 3] = 1.0f;
            } else {
                a[4 * i] = c[0]; a[4 * i + 1] = c[1]; a[4 * i + 2] = c[2]; a[4 * i + 3] = c[3];
            }
        }
    float *src = a, *dst = b;
    if (levels < 1) levels = 1;
    for (int l = 0; l < levels; l++) {
        atrous(src, dst, nrm_rgba, depth, w, h, 1 << l, sigma_c, sigma_n, sigma_d);
        float *t = src; src = dst; dst = t;
    }
    const float t = fminf(fmaxf(intensity, 0.0f), 1.0f);
    const float s = 1.0f - t;
    for (size_t i = 0; i < n; i++) {
        const float *f = src + 4 * i, *nz = noisy + 4 * i, *al = albedo_rgba + 4 * i;
        float r[3];
        if (al[3] > 0.5f) { r[0] = f[0] * al[0]; r[1] = f[1] * al[1]; r[2] = f[2] * al[2]; }
        else { r[0] = f[0]; r[1] = f[1]; r[2] = f[2]; }
        out[4 * i] = nz[0] * s + r[0] * t;
        out[4 * i + 1] = nz[1] * s + r[1] * t;
        out[4 * i + 2] = nz[2] * s + r[2] * t;
        out[4 * i + 3] = 1.0f;
    }
    free(a);
    free(b);
}

/* ======================================================================== */
/* adaptive sampling driver (rayrenderer.cpp:1011-1252)                     */
/* ======================================================================== */
/* computeTileNoise (:1038-1087): mean over the tile's pixels (skipping those
 * whose largest mean channel is < 1e-3) of the largest per-channel relative
 * standard error of the mean; 1e9 before two samples. */
float or_tile_noise(const float *sum, const float *sumsq, int W, int H, int tx, int ty, int n) {
    if (n <= 1) return 1e9f;
    const float invN = 1.0f / (float)n;
    double acc = 0.0;
    int cnt = 0;
    for (int y = ty * 32; y < ty * 32 + 32 && y < H; y++)
        for (int x = tx * 32; x < tx * 32 + 32 && x < W; x++) {
            const size_t i = 3 * ((size_t)y * W + x);
            const float mr = sum[i] * invN, mg = sum[i + 1] * invN, mb = sum[i + 2] * invN;
            if (fmaxf(fmaxf(mr, mg), mb) < 1e-3f) continue;
            const float vr = fmaxf(0.0f, sumsq[i] * invN - mr * mr);
            const float vg = fmaxf(0.0f, sumsq[i + 1] * invN - mg * mg);
            const float vb = fmaxf(0.0f, sumsq[i + 2] * invN - mb * mb);
            const float rr = sqrtf(vr * invN) / fmaxf(mr, 1e-3f);
            const float rg = sqrtf(vg * invN) / fmaxf(mg, 1e-3f);
            const float rb = sqrtf(vb * invN) / fmaxf(mb, 1e-3f);
            acc += (double)fmaxf(fmaxf(rr, rg), rb);
            cnt++;
        }
    if (cnt == 0) return 0.0f;
    return (float)(acc / (double)cnt);
}

/* renderAdaptive (:1162-1252): pass 0 gives every tile min(base, spp)
 * samples; each refinement pass adds min(base, spp - spent) samples to the
 * tiles still above the threshold.  nspp receives the per-tile counts. */
int or_render_adaptive(const or_scene_desc *desc, const or_settings *S, int mode, int nthreads, uint32_t base,
                       float threshold, float *sum_rgb, float *sumsq_rgb, uint32_t *nspp) {
    const int W = (int)S->width, H = (int)S->height;
    const int tx = (W + 31) / 32, ty = (H + 31) / 32, nt = tx * ty;
    const uint32_t target = S->spp > 0 ? S->spp : 1;
    if (base < 1) base = 1;
    or_settings s = *S;
    const uint32_t first = base < target ? base : target;
    s.sample_start = 0;
    s.spp = first;
    or_render(desc, &s, mode, nthreads, NULL, sum_rgb, sumsq_rgb);
    for (int t = 0; t < nt; t++) nspp[t] = first;
    uint32_t spent = first;
    uint8_t *active = (uint8_t *)calloc((size_t)nt, 1);
    while (spent < target) {
        int any = 0;
        for (int t = 0; t < nt; t++) {
            active[t] = 0;
            if (nspp[t] >= target) continue;
            if (or_tile_noise(sum_rgb, sumsq_rgb, W, H, t % tx, t / tx, (int)nspp[t]) > threshold) {
                active[t] = 1;
                any = 1;
            }
        }
        if (!any) break;
        const uint32_t pc = base < target - spent ? base : target - spent;
        s.sample_start = spent;
        s.spp = pc;
        or_render(desc, &s, mode, nthreads, active, sum_rgb, sumsq_rgb);
        for (int t = 0; t < nt; t++) if (active[t]) nspp[t] += pc;
        spent += pc;
    }
    free(active);
    return 0;
}

/* ======================================================================== */
/* probes of single functions (tests only)                                  */
/* ======================================================================== */
void or_phase_sample(float g, const float wo[3], int n, const float *u2, float *dirs) {
    medium_t m;
    memset(&m, 0, sizeof m);
    m.g = g;
    const v3 w = mk(wo[0], wo[1], wo[2]);
    for (int i = 0; i < n; i++) {
        const v3 d = sample_phase(&m, w, u2[2 * i], u2[2 * i + 1]);
        dirs[3 * i] = d.x;
        dirs[3 * i + 1] = d.y;
        dirs[3 * i + 2] = d.z;
    }
}

void or_phase_pdf(float g, const float wo[3], int n, const float *wi, float *pdf) {
    medium_t m;
    memset(&m, 0, sizeof m);
    m.g = g;
    const v3 w = mk(wo[0], wo[1], wo[2]);
    for (int i = 0; i < n; i++) pdf[i] = phase_pdf(&m, w, mk(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]));
}

void or_env_probe(const or_scene_desc *desc, int n, const float *u2, float *dirs, float *pdfs, float *rad, int nq,
                  const float *qdirs, float *qpdfs, float *qrad) {
    or_settings S;
    memset(&S, 0, sizeof S);
    S.width = S.height = 1;
    S.spp = S.max_depth = 1;
    prep_t P;
    prepare(&P, desc, &S);
    for (int i = 0; i < n; i++) {
        v3 d;
        float pdf;
        sample_env_dir(&P, u2[2 * i], u2[2 * i + 1], &d, &pdf);
        const v3 L = sample_environment(&P, d);
        dirs[3 * i] = d.x;
        dirs[3 * i + 1] = d.y;
        dirs[3 * i + 2] = d.z;
        pdfs[i] = pdf;
        rad[3 * i] = L.x;
        rad[3 * i + 1] = L.y;
        rad[3 * i + 2] = L.z;
    }
    for (int i = 0; i < nq; i++) {
        const v3 d = mk(qdirs[3 * i], qdirs[3 * i + 1], qdirs[3 * i + 2]);
        qpdfs[i] = env_dir_pdf(&P, d);
        const v3 L = sample_environment(&P, d);
        qrad[3 * i] = L.x;
        qrad[3 * i + 1] = L.y;
        qrad[3 * i + 2] = L.z;
    }
    release(&P);
}

/* ---- medium probes (tests only) ------------------------------------------ */
static rng_t probe_rng(uint32_t seed, uint32_t i) {
    rng_t R;
    memset(&R, 0, sizeof R);
    R.pix = i;
    R.sample = 0;
    R.key0 = seed;
    R.key1 = 0x9E3779B9u;
    R.blk_id = 0xFFFFFFFFu;
    return R;
}

void or_medium_eval(const or_medium *src, int op, int n, const float *pts, float *out) {
    medium_t m;
    prep_medium(&m, src);
    for (int i = 0; i < n; i++) {
        const v3 p = mk(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        if (op == 0) out[i] = density_at(&m, p);
        else if (op == 1) {
            const v3 e = emission_at(&m, p);
            out[3 * i] = e.x;
            out[3 * i + 1] = e.y;
            out[3 * i + 2] = e.z;
        } else {
            out[i] = ior_at(&m, p);
        }
    }
}

void or_delta_track_probe(const or_medium *src, const float o[3], const float d[3], float maxT, int n,
                          uint32_t seed, float *t_out) {
    medium_t m;
    prep_medium(&m, src);
    for (int i = 0; i < n; i++) {
        rng_t R = probe_rng(seed, (uint32_t)i);
        float t = 0.0f, den = 0.0f;
        t_out[i] = delta_track(&m, &R, ld3(o), ld3(d), maxT, &t, &den) ? t : -1.0f;
    }
}

void or_emission_integral_probe(const or_medium *src, const float o[3], const float d[3], float maxT, int n,
                                uint32_t seed, float *out) {
    medium_t m;
    prep_medium(&m, src);
    for (int i = 0; i < n; i++) {
        rng_t R = probe_rng(seed, (uint32_t)i);
        const v3 e = emission_integral(&m, &R, ld3(o), ld3(d), maxT);
        out[3 * i] = e.x;
        out[3 * i + 1] = e.y;
        out[3 * i + 2] = e.z;
    }
}

void or_bend_probe(const or_medium *src, const float o[3], const float d[3], float length, float *out6) {
    medium_t m;
    prep_medium(&m, src);
    v3 bo, bd;
    bend_ray(&m, ld3(o), ld3(d), length, &bo, &bd);
    out6[0] = bo.x; out6[1] = bo.y; out6[2] = bo.z;
    out6[3] = bd.x; out6[4] = bd.y; out6[5] = bd.z;
}

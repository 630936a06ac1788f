/*
 * rg_oracle.c — CPU restatement of RayGen's BVH build, ray/triangle tests and
 * BVH traversal (TEST INFRASTRUCTURE: the checker, never the product).
 *
 * Compiled with -O3 -ffp-contract=off and no fast-math: every float operation
 * is evaluated in the order written, in binary32, so results are bit-exact
 * against the HIP kernels built with the same contract setting.
 *
 * Reference files restated (jingwood/raygen-renderer, src/raygen/):
 *   bvh.cpp:22-25     SAH constants
 *   bvh.cpp:76-98     TriangleBVH::build
 *   bvh.cpp:100-231   TriangleBVH::buildRecursive
 *   bvh.cpp:233-285   TriangleBVH::intersectClosest
 *   bvh.h:63-101      TriangleBVH::intersectAny
 *   raycommon.cpp:151-198  RenderMeshTriangle::intersectsRay(ray, maxt, t, hit)
 *   raycommon.cpp:200-235  RenderMeshTriangle::intersectsRay(ray, info)
 *   raycommon.cpp:21-92    Halton LDS + hashPixel
 * The ugm vector library the reference includes is not vendored; dot() and
 * cross() are taken in their textbook left-to-right form.
 */
#include "rg_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* vector helpers: explicit component order                                  */
/* ------------------------------------------------------------------------ */
static inline void v_sub(float r[3], const float a[3], const float b[3]) {
    r[0] = a[0] - b[0];
    r[1] = a[1] - b[1];
    r[2] = a[2] - b[2];
}
static inline void v_cross(float r[3], const float a[3], const float b[3]) {
    float x = a[1] * b[2] - a[2] * b[1];
    float y = a[2] * b[0] - a[0] * b[2];
    float z = a[0] * b[1] - a[1] * b[0];
    r[0] = x;
    r[1] = y;
    r[2] = z;
}
static inline float v_dot(const float a[3], const float b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11 "Parallel random numbers: as easy as */
/* 1, 2, 3").  Used to draw every non-LDS random number of the renderer.     */
/* ------------------------------------------------------------------------ */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ------------------------------------------------------------------------ */
/* Halton LDS with per-pixel Cranley-Patterson rotation (raycommon.cpp:21-92) */
/* ------------------------------------------------------------------------ */
static const int LDS_PRIMES[16] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53};

uint32_t or_hash_pixel(int x, int y, uint32_t salt) {
    uint32_t v = (uint32_t)x * 73856093u ^ (uint32_t)y * 19349663u ^ salt * 83492791u;
    v ^= v >> 16; v *= 0x7feb352du;
    v ^= v >> 15; v *= 0x846ca68bu;
    v ^= v >> 16;
    return v;
}

float or_radical_inverse(uint32_t i, int base) {
    const float inv_base = 1.0f / (float)base;
    float inv_base_n = 1.0f;
    uint64_t rev = 0;
    while (i > 0) {
        uint64_t next = i / (uint32_t)base;
        uint64_t digit = i - next * (uint64_t)base;
        rev = rev * (uint64_t)base + digit;
        inv_base_n *= inv_base;
        i = (uint32_t)next;
    }
    float v = (float)rev * inv_base_n;
    return v < 1.0f ? v : 0.9999999f;
}

/* Value of ldsNext1D() for (pixel, sample) at dimension `dim` (< 16). */
float or_lds_sample(int x, int y, int sample_idx, int dim) {
    const uint32_t scramble = or_hash_pixel(x, y, 0);
    const float v = or_radical_inverse((uint32_t)sample_idx + 1u, LDS_PRIMES[dim]);
    const uint32_t h = or_hash_pixel(dim, 0, scramble);
    const float offset = (float)(h >> 8) * (1.0f / (float)(1u << 24));
    float w = v + offset;
    if (w >= 1.0f) w -= 1.0f;
    return w;
}

/* ------------------------------------------------------------------------ */
/* Möller–Trumbore, closest variant (raycommon.cpp:200-235)                  */
/* ------------------------------------------------------------------------ */
int or_mt_closest(const float v1[3], const float v2[3], const float v3[3],
                  const float o[3], const float d[3], float tmin,
                  float *t, float *u, float *v) {
    float e1[3], e2[3], p[3], tv[3], q[3];
    v_sub(e1, v2, v1);
    v_sub(e2, v3, v1);
    v_cross(p, d, e2);
    const float det = v_dot(e1, p);
    if (fabsf(det) < 1e-8f) return 0;
    const float inv_det = 1.0f / det;
    v_sub(tv, o, v1);
    const float uu = v_dot(tv, p) * inv_det;
    if (uu < 0.0f || uu > 1.0f) return 0;
    v_cross(q, tv, e1);
    const float vv = v_dot(d, q) * inv_det;
    if (vv < 0.0f || uu + vv > 1.0f) return 0;
    const float tt = v_dot(e2, q) * inv_det;
    if (tt < tmin) return 0;  /* reference: t < 0 (tmin = 0) */
    if (tt >= *t) return 0;   /* strict: first found wins exact ties */
    *t = tt;
    *u = uu;
    *v = vv;
    return 1;
}

/* Same test, reporting the candidate t without the `>= current` reject. */
static int mt_closest_raw(const float v1[3], const float v2[3], const float v3[3],
                          const float o[3], const float d[3], float tmin,
                          float *t, float *u, float *v) {
    float big = INFINITY;
    int r = or_mt_closest(v1, v2, v3, o, d, tmin, &big, u, v);
    *t = big;
    return r;
}

/* ------------------------------------------------------------------------ */
/* Möller–Trumbore, any-hit variant (raycommon.cpp:151-198)                  */
/* ------------------------------------------------------------------------ */
int or_mt_any(const float v1[3], const float v2[3], const float v3[3],
              const float o[3], const float d[3], float maxt, float *t_out) {
    const float EPS = 1e-6f;
    float e1[3], e2[3], h[3], s[3], q[3];
    v_sub(e1, v2, v1);
    v_sub(e2, v3, v1);
    v_cross(h, d, e2);
    const float a = v_dot(e1, h);
    if (fabsf(a) < EPS) return 0;
    const float f = 1.0f / a;
    v_sub(s, o, v1);
    const float u = f * v_dot(s, h);
    if (u < 0.0f || u > 1.0f) return 0;
    v_cross(q, s, e1);
    const float v = f * v_dot(d, q);
    if (v < 0.0f || u + v > 1.0f) return 0;
    const float t = f * v_dot(e2, q);
    if (t > EPS && t < maxt) {
        if (t_out) *t_out = t;
        return 1;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Binned SAH build (bvh.cpp:22-25, 76-231)                                  */
/* ------------------------------------------------------------------------ */
#define OR_BINS 16
#define OR_LEAF 4u
#define OR_TRAV_COST 1.0f
#define OR_ISECT_COST 1.5f

typedef struct { float mn[3], mx[3]; } box3;

static inline void box_empty(box3 *b) {
    for (int k = 0; k < 3; k++) { b->mn[k] = FLT_MAX; b->mx[k] = -FLT_MAX; }
}
static inline void box_grow_box(box3 *d, const box3 *s) {
    for (int k = 0; k < 3; k++) {
        d->mn[k] = fminf(d->mn[k], s->mn[k]);
        d->mx[k] = fmaxf(d->mx[k], s->mx[k]);
    }
}
static inline void box_grow_pt(box3 *d, const float p[3]) {
    for (int k = 0; k < 3; k++) {
        d->mn[k] = fminf(d->mn[k], p[k]);
        d->mx[k] = fmaxf(d->mx[k], p[k]);
    }
}
static inline float box_area(const box3 *b) {
    const float dx = b->mx[0] - b->mn[0];
    const float dy = b->mx[1] - b->mn[1];
    const float dz = b->mx[2] - b->mn[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

typedef struct {
    const box3 *pbox;    /* per original triangle */
    uint32_t *perm;      /* prims (original ids), permuted in place */
    float (*cent)[3];    /* centroids, permuted alongside */
    or_node *nodes;
    int64_t nnodes;
} build_ctx;

static void build_rec(build_ctx *c, uint32_t node_idx, uint32_t first, uint32_t count) {
    box3 pb, cb;
    box_empty(&pb);
    box_empty(&cb);
    for (uint32_t i = first; i < first + count; i++) {
        box_grow_box(&pb, &c->pbox[c->perm[i]]);
        box_grow_pt(&cb, c->cent[i]);
    }
    or_node *nd = &c->nodes[node_idx];
    for (int k = 0; k < 3; k++) { nd->bmin[k] = pb.mn[k]; nd->bmax[k] = pb.mx[k]; }

    if (count <= OR_LEAF) {
        nd->first_or_left = first;
        nd->count = (uint16_t)count;
        nd->axis = 0;
        return;
    }

    float cext[3];
    for (int k = 0; k < 3; k++) cext[k] = cb.mx[k] - cb.mn[k];
    int axis = 0;
    if (cext[1] > cext[0]) axis = 1;
    if (cext[2] > ((axis == 0) ? cext[0] : cext[1])) axis = 2;

    const float cmin = cb.mn[axis];
    const float cext_axis = cext[axis];
    if (cext_axis < 1e-12f) {
        nd->first_or_left = first;
        nd->count = (uint16_t)(count < 0xFFFFu ? count : 0xFFFFu);
        nd->axis = 0;
        return;
    }
    const float scale = (float)OR_BINS / cext_axis;

    box3 bbox[OR_BINS];
    int bcnt[OR_BINS];
    for (int b = 0; b < OR_BINS; b++) { box_empty(&bbox[b]); bcnt[b] = 0; }
    for (uint32_t i = first; i < first + count; i++) {
        int b = (int)((c->cent[i][axis] - cmin) * scale);
        if (b < 0) b = 0;
        if (b >= OR_BINS) b = OR_BINS - 1;
        box_grow_box(&bbox[b], &c->pbox[c->perm[i]]);
        bcnt[b]++;
    }

    float larea[OR_BINS - 1], rarea[OR_BINS - 1];
    int lcnt[OR_BINS - 1], rcnt[OR_BINS - 1];
    {
        box3 lb; box_empty(&lb);
        int lc = 0;
        for (int i = 0; i < OR_BINS - 1; i++) {
            if (bcnt[i] > 0) box_grow_box(&lb, &bbox[i]);
            lc += bcnt[i];
            larea[i] = (lc > 0) ? box_area(&lb) : 0.0f;
            lcnt[i] = lc;
        }
        box3 rb; box_empty(&rb);
        int rc = 0;
        for (int i = OR_BINS - 1; i > 0; i--) {
            if (bcnt[i] > 0) box_grow_box(&rb, &bbox[i]);
            rc += bcnt[i];
            rarea[i - 1] = (rc > 0) ? box_area(&rb) : 0.0f;
            rcnt[i - 1] = rc;
        }
    }

    const float parea = box_area(&pb);
    const float inv_parea = (parea > 0.0f) ? 1.0f / parea : 0.0f;
    const float leaf_cost = (float)count * OR_ISECT_COST;
    float best_cost = leaf_cost;
    int best_bin = -1;
    for (int i = 0; i < OR_BINS - 1; i++) {
        if (lcnt[i] == 0 || rcnt[i] == 0) continue;
        const float cost = OR_TRAV_COST
            + OR_ISECT_COST * inv_parea
              * (larea[i] * (float)lcnt[i] + rarea[i] * (float)rcnt[i]);
        if (cost < best_cost) { best_cost = cost; best_bin = i; }
    }
    if (best_bin < 0) {
        nd->first_or_left = first;
        nd->count = (uint16_t)(count < 0xFFFFu ? count : 0xFFFFu);
        nd->axis = 0;
        return;
    }

    uint32_t mid = first;
    for (uint32_t i = first; i < first + count; i++) {
        int b = (int)((c->cent[i][axis] - cmin) * scale);
        if (b < 0) b = 0;
        if (b >= OR_BINS) b = OR_BINS - 1;
        if (b <= best_bin) {
            uint32_t tp = c->perm[i]; c->perm[i] = c->perm[mid]; c->perm[mid] = tp;
            float tc[3];
            memcpy(tc, c->cent[i], sizeof tc);
            memcpy(c->cent[i], c->cent[mid], sizeof tc);
            memcpy(c->cent[mid], tc, sizeof tc);
            mid++;
        }
    }
    const uint32_t lcount = mid - first;
    const uint32_t rcount = count - lcount;
    if (lcount == 0 || rcount == 0) {
        nd->first_or_left = first;
        nd->count = (uint16_t)(count < 0xFFFFu ? count : 0xFFFFu);
        nd->axis = 0;
        return;
    }
    const uint32_t left = (uint32_t)c->nnodes;
    c->nnodes += 2;
    memset(&c->nodes[left], 0, 2 * sizeof(or_node));
    nd = &c->nodes[node_idx];
    nd->first_or_left = left;
    nd->count = 0;
    nd->axis = (uint16_t)axis;
    build_rec(c, left, first, lcount);
    build_rec(c, left + 1, mid, rcount);
}

int64_t or_bvh_build(const float *v9, uint32_t n, or_node *nodes, uint32_t *perm) {
    if (n == 0) return 0;
    box3 *pbox = (box3 *)malloc(sizeof(box3) * n);
    float (*cent)[3] = malloc(sizeof(float[3]) * n);
    for (uint32_t i = 0; i < n; i++) {
        const float *t = v9 + 9 * (size_t)i;
        /* BoundingBox::fromTriangle: componentwise min/max of the vertices */
        box_empty(&pbox[i]);
        box_grow_pt(&pbox[i], t);
        box_grow_pt(&pbox[i], t + 3);
        box_grow_pt(&pbox[i], t + 6);
        perm[i] = i;
        for (int k = 0; k < 3; k++) cent[i][k] = (pbox[i].mn[k] + pbox[i].mx[k]) * 0.5f;
    }
    build_ctx c = {pbox, perm, cent, nodes, 1};
    memset(&nodes[0], 0, sizeof(or_node));
    build_rec(&c, 0, 0, n);
    free(pbox);
    free(cent);
    return c.nnodes;
}

/* ------------------------------------------------------------------------ */
/* Traversal                                                                 */
/* ------------------------------------------------------------------------ */
static inline float inv_dir(float d) { return (fabsf(d) > 1e-20f) ? 1.0f / d : 1e30f; }

static inline int slab(const or_node *n, const float o[3], const float inv[3],
                       float tcur) {
    const float t1x = (n->bmin[0] - o[0]) * inv[0];
    const float t2x = (n->bmax[0] - o[0]) * inv[0];
    const float t1y = (n->bmin[1] - o[1]) * inv[1];
    const float t2y = (n->bmax[1] - o[1]) * inv[1];
    const float t1z = (n->bmin[2] - o[2]) * inv[2];
    const float t2z = (n->bmax[2] - o[2]) * inv[2];
    const float tmin = fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fminf(t1z, t2z));
    const float tmax = fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fmaxf(t1z, t2z));
    return !(tmax < 0.0f || tmin > tmax || tmin > tcur);
}

#define OR_STACK 128

static void closest_one(const or_node *nodes, int64_t nnodes, const float *v9,
                        const uint32_t *perm, const or_ray *r, or_hit *h,
                        uint32_t *cnt, uint8_t *tie) {
    h->prim = 0xFFFFFFFFu;
    h->t = r->tmax;
    h->u = 0.0f;
    h->v = 0.0f;
    uint32_t pops = 0, tests = 0;
    float tie_t = NAN;
    if (nnodes > 0) {
        const float inv[3] = {inv_dir(r->d[0]), inv_dir(r->d[1]), inv_dir(r->d[2])};
        const int sign[3] = {inv[0] < 0.0f, inv[1] < 0.0f, inv[2] < 0.0f};
        uint32_t stack[OR_STACK];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const or_node *n = &nodes[stack[--sp]];
            pops++;
            if (!slab(n, r->o, inv, h->t)) continue;
            if (n->count > 0) {
                for (uint32_t i = 0; i < n->count; i++) {
                    const uint32_t p = perm[n->first_or_left + i];
                    const float *tv = v9 + 9 * (size_t)p;
                    float tt = h->t, uu, vv;
                    tests++;
                    if (or_mt_closest(tv, tv + 3, tv + 6, r->o, r->d, r->tmin, &tt, &uu, &vv)) {
                        h->t = tt; h->u = uu; h->v = vv; h->prim = p;
                    } else if (tie) {
                        float t2, u2, v2;
                        if (mt_closest_raw(tv, tv + 3, tv + 6, r->o, r->d, r->tmin, &t2, &u2, &v2)
                            && t2 == h->t)
                            tie_t = t2;
                    }
                }
            } else {
                const uint32_t l = n->first_or_left;
                if (sign[n->axis]) { stack[sp++] = l; stack[sp++] = l + 1; }
                else { stack[sp++] = l + 1; stack[sp++] = l; }
            }
        }
    }
    if (cnt) { cnt[0] = pops; cnt[1] = tests; }
    if (tie) *tie = (h->prim != 0xFFFFFFFFu && tie_t == h->t) ? 1 : 0;
}

static int any_one(const or_node *nodes, int64_t nnodes, const float *v9,
                   const uint32_t *perm, const or_ray *r) {
    if (nnodes <= 0) return 0;
    const float inv[3] = {inv_dir(r->d[0]), inv_dir(r->d[1]), inv_dir(r->d[2])};
    uint32_t stack[OR_STACK];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const or_node *n = &nodes[stack[--sp]];
        if (!slab(n, r->o, inv, r->tmax)) continue;
        if (n->count > 0) {
            for (uint32_t i = 0; i < n->count; i++) {
                const float *tv = v9 + 9 * (size_t)perm[n->first_or_left + i];
                if (or_mt_any(tv, tv + 3, tv + 6, r->o, r->d, r->tmax, NULL)) return 1;
            }
        } else {
            stack[sp++] = n->first_or_left + 1;
            stack[sp++] = n->first_or_left;
        }
    }
    return 0;
}

/* Work distribution mirrors the reference's tile scheduler
 * (rayrenderer.cpp:717-809): fixed-size chunks claimed by fetch_add. */
typedef struct {
    const or_node *nodes; int64_t nnodes; const float *v9; const uint32_t *perm;
    const or_ray *rays; uint64_t nrays;
    or_hit *hits; uint32_t *counts; uint8_t *ties; uint8_t *occ;
    int mode;
    atomic_uint_fast64_t next;
} trav_job;

#define OR_CHUNK 4096u

static void *trav_worker(void *arg) {
    trav_job *j = (trav_job *)arg;
    for (;;) {
        uint64_t s = atomic_fetch_add(&j->next, OR_CHUNK);
        if (s >= j->nrays) break;
        uint64_t e = s + OR_CHUNK < j->nrays ? s + OR_CHUNK : j->nrays;
        for (uint64_t i = s; i < e; i++) {
            if (j->mode == 0)
                closest_one(j->nodes, j->nnodes, j->v9, j->perm, &j->rays[i], &j->hits[i],
                            j->counts ? j->counts + 2 * i : NULL, j->ties ? j->ties + i : NULL);
            else
                j->occ[i] = (uint8_t)any_one(j->nodes, j->nnodes, j->v9, j->perm, &j->rays[i]);
        }
    }
    return NULL;
}

static void run_trav(trav_job *j, int nthreads) {
    atomic_init(&j->next, 0);
    if (nthreads <= 1) { trav_worker(j); return; }
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int i = 0; i < nthreads; i++) pthread_create(&th[i], NULL, trav_worker, j);
    for (int i = 0; i < nthreads; i++) pthread_join(th[i], NULL);
    free(th);
}

void or_trace_closest(const or_node *nodes, int64_t nnodes, const float *v9,
                      const uint32_t *perm, const or_ray *rays, uint64_t nrays,
                      or_hit *hits, uint32_t *counts, uint8_t *ties, int nthreads) {
    trav_job j;
    memset(&j, 0, sizeof j);
    j.nodes = nodes; j.nnodes = nnodes; j.v9 = v9; j.perm = perm;
    j.rays = rays; j.nrays = nrays; j.hits = hits; j.counts = counts; j.ties = ties;
    j.mode = 0;
    run_trav(&j, nthreads);
}

void or_trace_any(const or_node *nodes, int64_t nnodes, const float *v9,
                  const uint32_t *perm, const or_ray *rays, uint64_t nrays,
                  uint8_t *occluded, int nthreads) {
    trav_job j;
    memset(&j, 0, sizeof j);
    j.nodes = nodes; j.nnodes = nnodes; j.v9 = v9; j.perm = perm;
    j.rays = rays; j.nrays = nrays; j.occ = occluded;
    j.mode = 1;
    run_trav(&j, nthreads);
}

void or_closest1(const or_node *nodes, int64_t nnodes, const float *v9, const uint32_t *perm,
                 const or_ray *r, or_hit *h) {
    closest_one(nodes, nnodes, v9, perm, r, h, NULL, NULL);
}

/* intersectAny with the reference's shadow predicates (rayrenderer.cpp:1956-1959,
 * 1499-1503, 2172-2176, 2391-2393): pred 0 = every triangle, 1 = area-light
 * NEE, 2 = envmap NEE and point lights, 3 = volume lights, 4 = calcAO. */
static int occludes(const or_object *m, uint8_t pred) {
    if (pred == 1) return m->transparency < 0.01f && m->emission <= 0.0f;
    if (pred == 2) return !(m->emission > 0.0f) && (m->transparency < 0.01f || m->refraction > 0.1f);
    if (pred == 3) return !(m->emission > 0.0f) && m->transparency < 0.01f && m->refraction <= 0.1f;
    if (pred == 4) return m->transparency < 0.01f;
    return 1;
}

int or_any1_pred(const or_node *nodes, int64_t nnodes, const float *v9, const uint32_t *perm,
                 const or_ray *r, const uint32_t *obj_of_tri, const or_object *objs, uint8_t pred) {
    if (nnodes <= 0) return 0;
    const float inv[3] = {inv_dir(r->d[0]), inv_dir(r->d[1]), inv_dir(r->d[2])};
    uint32_t stack[OR_STACK];
    int sp = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        const or_node *n = &nodes[stack[--sp]];
        if (!slab(n, r->o, inv, r->tmax)) continue;
        if (n->count > 0) {
            for (uint32_t i = 0; i < n->count; i++) {
                const uint32_t p = perm[n->first_or_left + i];
                const float *tv = v9 + 9 * (size_t)p;
                if (or_mt_any(tv, tv + 3, tv + 6, r->o, r->d, r->tmax, NULL) &&
                    occludes(&objs[obj_of_tri[p]], pred))
                    return 1;
            }
        } else {
            stack[sp++] = n->first_or_left + 1;
            stack[sp++] = n->first_or_left;
        }
    }
    return 0;
}

void or_brute_closest(const float *v9, uint32_t ntris, const or_ray *rays,
                      uint64_t nrays, or_hit *hits, uint8_t *ties) {
    for (uint64_t r = 0; r < nrays; r++) {
        const or_ray *ry = &rays[r];
        or_hit h = {0xFFFFFFFFu, ry->tmax, 0.0f, 0.0f};
        int ntie = 0;
        for (uint32_t i = 0; i < ntris; i++) {
            const float *tv = v9 + 9 * (size_t)i;
            float t2, u2, v2;
            if (!mt_closest_raw(tv, tv + 3, tv + 6, ry->o, ry->d, ry->tmin, &t2, &u2, &v2)) continue;
            if (t2 < h.t) { h.t = t2; h.u = u2; h.v = v2; h.prim = i; ntie = 0; }
            else if (t2 == h.t && h.prim != 0xFFFFFFFFu) ntie = 1;
        }
        hits[r] = h;
        if (ties) ties[r] = (uint8_t)ntie;
    }
}

void or_brute_any(const float *v9, uint32_t ntris, const or_ray *rays,
                  uint64_t nrays, uint8_t *occluded) {
    for (uint64_t r = 0; r < nrays; r++) {
        const or_ray *ry = &rays[r];
        uint8_t occ = 0;
        for (uint32_t i = 0; i < ntris && !occ; i++) {
            const float *tv = v9 + 9 * (size_t)i;
            occ = (uint8_t)or_mt_any(tv, tv + 3, tv + 6, ry->o, ry->d, ry->tmax, NULL);
        }
        occluded[r] = occ;
    }
}

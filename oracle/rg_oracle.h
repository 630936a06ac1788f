/*
 * rg_oracle.h — CPU restatement of RayGen's BSDF path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked or
 * loaded by tests/, by __graft_entry__.smoke() and by bench.py's CPU-baseline
 * leg, never by the product library under raygen-renderer_amd/.
 *
 * Every function restates the reference algorithm it names (file:line in
 * jingwood/raygen-renderer, src/raygen/...).  Build with strict IEEE FP
 * (-ffp-contract=off, no fast-math) so that results are reproducible bit for
 * bit against the HIP kernels compiled with the same contract setting.
 *
 * Parity pins (what checks each function — see tests/):
 *   or_philox4x32_10     Random123 published known-answer vectors.
 *   or_mt_closest/any    hand-built triangles with closed-form hits, edge,
 *                        vertex, parallel and behind-origin cases, det
 *                        thresholds 1e-8 / 1e-6.
 *   or_bvh_build         structural invariants (every prim in one leaf, child
 *                        boxes inside parent, contiguous children, leaf size,
 *                        SAH choice re-derived by exhaustive bin search on
 *                        small inputs).
 *   or_trace_closest     brute-force O(N) loop with the same intersector
 *                        (identical except flagged exact-t ties).
 *   or_trace_any         brute-force any-hit loop (order independent).
 *   or_lds_*             exact rational radical inverse, hash closed form.
 *   or_render_*          white-furnace closed form, MIS two-strategy sum,
 *                        recursive (reference) vs single-lobe estimators.
 */
#ifndef RG_ORACLE_H
#define RG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same 32-byte node as BVHNode (bvh.h:22-30). */
typedef struct {
    float bmin[3];
    uint32_t first_or_left;
    float bmax[3];
    uint16_t count; /* >0 leaf prim count, 0 internal */
    uint16_t axis;  /* split axis of an internal node */
} or_node;

/* Ray as passed over the boundary: origin, tmin, direction, tmax (32 B). */
typedef struct {
    float o[3];
    float tmin;
    float d[3];
    float tmax;
} or_ray;

/* Closest-hit record (16 B): original triangle index (~0u on miss), t, u, v. */
typedef struct {
    uint32_t prim;
    float t;
    float u;
    float v;
} or_hit;

/* --- random numbers ------------------------------------------------------ */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* --- Halton/Cranley-Patterson sampler (raycommon.cpp:21-92) ------------- */
uint32_t or_hash_pixel(int x, int y, uint32_t salt);
float or_radical_inverse(uint32_t i, int base);
float or_lds_sample(int x, int y, int sample_idx, int dim);

/* --- triangle intersectors (raycommon.cpp:151-235) ----------------------- */
/* closest variant: returns 1 and updates *t,*u,*v when the hit is in
 * [0, *t) (strict reject of t >= current). */
int or_mt_closest(const float v1[3], const float v2[3], const float v3[3],
                  const float o[3], const float d[3], float tmin,
                  float *t, float *u, float *v);
/* any-hit variant: 1 when 1e-6 < t < maxt. */
int or_mt_any(const float v1[3], const float v2[3], const float v3[3],
              const float o[3], const float d[3], float maxt, float *t_out);

/* --- BVH (bvh.cpp:76-231) ------------------------------------------------ */
/* v9: n triangles x (v1,v2,v3).  nodes must hold 2n entries, perm n.
 * Returns the node count (0 for n == 0). */
int64_t or_bvh_build(const float *v9, uint32_t n, or_node *nodes, uint32_t *perm);

/* --- traversal (bvh.cpp:233-285, bvh.h:63-101) --------------------------- */
/* counts (nullable): per-ray [node pops, triangle tests] as uint32 pairs.
 * ties (nullable): per-ray 1 when another triangle produced exactly the
 * winning t (the winner then depends on visit order). */
void or_trace_closest(const or_node *nodes, int64_t nnodes, const float *v9,
                      const uint32_t *perm, const or_ray *rays, uint64_t nrays,
                      or_hit *hits, uint32_t *counts, uint8_t *ties, int nthreads);
void or_trace_any(const or_node *nodes, int64_t nnodes, const float *v9,
                  const uint32_t *perm, const or_ray *rays, uint64_t nrays,
                  uint8_t *occluded, int nthreads);
void or_brute_closest(const float *v9, uint32_t ntris, const or_ray *rays,
                      uint64_t nrays, or_hit *hits, uint8_t *ties);
void or_brute_any(const float *v9, uint32_t ntris, const or_ray *rays,
                  uint64_t nrays, uint8_t *occluded);

#ifdef __cplusplus
}
#endif
#endif

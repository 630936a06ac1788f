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

/* single-ray helpers used by the path tracer (same arithmetic as above) */
void or_closest1(const or_node *nodes, int64_t nnodes, const float *v9, const uint32_t *perm,
                 const or_ray *r, or_hit *h);

/* --- path tracer (rg_oracle_pt.c) ---------------------------------------- */
/* Material of a scene object (material.h:20-70 fields used by -d 5). */
typedef struct {
    float color[3];
    float glossy, roughness, transparency, refraction, refraction_ratio;
    float metallic, anisotropy, aniso_rotation, emission;
    int32_t interior_medium; /* index into media, -1 none */
    uint32_t visible;
    float chroma_dispersion; /* Material::chromaDispersion (material.h:36-39) */
    uint32_t pad;
} or_object;

/* HomogeneousMedium authoring fields (medium.h:37-214), positions in render
 * space (after HomogeneousMedium::bake, medium.cpp:59-104). */
typedef struct {
    float p[3];
    float radius, t;
} or_path_point;
#define OR_MAX_PATH_POINTS 16
typedef struct {
    float sigma_a[3], sigma_s[3], sigma_e[3];
    float g, density;
    int32_t emission_mode;              /* 0 constant, 1 cone, 2 path */
    float cone_origin[3], cone_axis[3];
    float cone_length, cone_radius;
    float cone_inner[3], cone_outer[3];
    float cone_intensity, cone_peak_axial, cone_peak_sharpness;
    int32_t cone_samples;
    or_path_point path[OR_MAX_PATH_POINTS];
    uint32_t n_path;
    float path_inner[3], path_outer[3];
    float path_intensity, path_falloff_power;
    int32_t path_samples;
    int32_t density_field;              /* 0 none, 1 fBm */
    float noise_frequency;
    int32_t noise_octaves;
    float noise_gain, noise_lacunarity, noise_amplitude, noise_bias;
    float noise_offset[3];
    int32_t heat_haze;
    float ior_amplitude, ior_frequency;
    int32_t ior_octaves;
    float ior_gain, ior_lacunarity;
    int32_t ior_march_steps;
    float ior_offset[3], ior_falloff;
} or_medium;

/* An emissive volume for NEE: medium index and the owning object's location
 * in render space (rayrenderer.cpp:180-206, 2083-2096). */
typedef struct {
    uint32_t medium;
    float pos[3];
} or_volume_light;

/* Camera at the render-space origin looking down -Z (scene.h:115-135). */
typedef struct {
    float fov_deg, exposure, depth_of_field, aperture;
    int32_t aperture_blades;
    float aperture_rotation_deg, view_far, pad;
} or_camera;

/* Equirectangular environment (linear RGB float, row 0 = top). */
typedef struct {
    uint32_t w, h;
    const float *rgb;
    float intensity, rotation_deg;
} or_env;

/* Material texture (texture.h:19-50): RGB floats as stored (LDR byte/255),
 * row 0 = top; srgb marks an LDR sRGB image. */
typedef struct {
    uint32_t w, h;
    const float *rgb;
    int32_t srgb;
    uint32_t pad;
} or_texture;

/* Point light source (transformScene, rayrenderer.cpp:333-353): render-
 * space position and direction, the object's colour, emission and
 * spotRange (0 = omni). */
typedef struct {
    float pos[3], normal[3], color[3];
    float emission, spot_range;
    uint32_t pad;
} or_point_light;

typedef struct {
    const float *v9;            /* render-space vertices */
    const float *n9;            /* render-space vertex normals */
    const uint32_t *obj_of_tri; /* object index per triangle */
    uint32_t ntris;
    const or_object *objs;
    uint32_t nobj;
    const or_medium *media;
    uint32_t nmedia;
    int32_t global_medium;      /* -1 none */
    or_env env;
    or_camera cam;
    /* albedo textures (nullable): texture table, texture per object (-1
     * none), texTiling per object (2 floats), first UV set per triangle (6) */
    const or_texture *tex;
    uint32_t ntex;
    const int32_t *tex_of_obj;
    const float *tiling;
    const float *uv6;
    /* point light sources (nullable) */
    const or_point_light *plights;
    uint32_t nplights;
    /* emissive volumes for NEE (nullable) */
    const or_volume_light *vlights;
    uint32_t nvlights;
} or_scene_desc;

typedef struct {
    uint32_t width, height, spp, max_depth;
    int32_t antialias;
    float firefly_clamp;
    uint32_t seed, job;
    uint32_t sample_start;
    int32_t color_sampling; /* RendererSettings::enableColorSampling (rayrenderer.h:106) */
    float back_color[4];    /* RendererSettings::backColor (rayrenderer.h:152) */
    int32_t shader;         /* RendererSettings::shaderProvider (-d; rayrenderer.cpp:59-77): 0-3 or 5 */
    float world_color[3];   /* RendererSettings::worldColor (rayrenderer.h:151) */
    uint32_t ao_samples;    /* RendererSettings::samples as calcAO reads it (0 = spp) */
} or_settings;

int or_any1_pred(const or_node *nodes, int64_t nnodes, const float *v9, const uint32_t *perm,
                 const or_ray *r, const uint32_t *obj_of_tri, const or_object *objs, uint8_t pred);

/* Adds, per pixel, the sum and sum of squares over samples
 * [sample_start, sample_start + spp) of the clamped per-sample radiance.
 * mode 0 = recursive restatement, 1 = single-lobe twin of the GPU path.
 * active_tile (nullable) masks 32x32 tiles. */
int or_render(const or_scene_desc *desc, const or_settings *s, int mode, int nthreads,
              const uint8_t *active_tile, float *sum_rgb, float *sumsq_rgb);
void or_render_one(const or_scene_desc *desc, const or_settings *s, int mode, int x, int y,
                   int sample, float out[3]);

/* guide buffers (RGBA normal/albedo, depth) and the A-Trous denoiser */
void or_aov(const or_scene_desc *desc, const or_settings *s, float *normal_rgba, float *albedo_rgba,
            float *depth);
void or_denoise(int w, int h, const float *noisy_rgba, const float *normal_rgba, const float *albedo_rgba,
                const float *depth, int levels, float sigma_c, float sigma_n, float sigma_d, float intensity,
                float *out_rgba);

/* adaptive sampling (computeTileNoise / renderAdaptive) */
float or_tile_noise(const float *sum_rgb, const float *sumsq_rgb, int w, int h, int tx, int ty, int n);
/* Probes of single restated functions, for the tests that pin them:
 * HomogeneousMedium::samplePhase / phasePdf (medium.cpp:496-525) for n
 * (u1, u2) pairs about wo; and the environment of a scene description:
 * sampleEnvmapDirection for n (u0, u1) pairs (directions, pdfs, radiance
 * seen along them) and envmapPdf + sampleEnvironment for nq query
 * directions (rayrenderer.cpp:1299-1573). */
void or_phase_sample(float g, const float wo[3], int n, const float *u2, float *dirs);
/* medium probes: op 0 densityAt, 1 emissionAt (3 floats), 2 iorAt at n
 * points; delta-tracking collision distances (-1: none before maxT) and
 * emission integrals for n independent generators; one bent ray */
void or_medium_eval(const or_medium *m, int op, int n, const float *pts, float *out);
void or_delta_track_probe(const or_medium *m, const float o[3], const float d[3], float maxT, int n,
                          uint32_t seed, float *t_out);
void or_emission_integral_probe(const or_medium *m, const float o[3], const float d[3], float maxT, int n,
                                uint32_t seed, float *out);
void or_bend_probe(const or_medium *m, const float o[3], const float d[3], float length, float *out6);
void or_phase_pdf(float g, const float wo[3], int n, const float *wi, float *pdf);
void or_env_probe(const or_scene_desc *desc, int n, const float *u2, float *dirs, float *pdfs, float *rad, int nq,
                  const float *qdirs, float *qpdfs, float *qrad);
int or_render_adaptive(const or_scene_desc *desc, const or_settings *s, int mode, int nthreads, uint32_t base,
                       float threshold, float *sum_rgb, float *sumsq_rgb, uint32_t *nspp);

/* Output stage (rg_oracle_post.c): HDR bloom in place
 * (RayRenderer::applyPostProcess, rayrenderer.cpp:598-697) and the
 * Reinhard + gamma tonemap (applyTonemapGamma, :2614-2647); RGBA floats. */
void or_bloom(float *hdr_rgba, int w, int h, float threshold, float strength, float radius, float size_aspect,
              float curve);
void or_tonemap(const float *hdr_rgba, int w, int h, float *ldr_rgba);

#ifdef __cplusplus
}
#endif
#endif

// rg_pt.h — device-resident scene and path-state layouts of the wavefront
// path tracer, shared by the host driver (rg_render.cpp) and the kernels
// (rg_pt.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rg_internal.h"

namespace rg {

struct RootBox {
    float lo[3], hi[3];
};

// Device view of a built scene's acceleration structure.
struct DevBvh {
    const void *nodes;  // Node4 or Node8 array (the kernel's width template)
    const float4 *tris;
    RootBox root;
    int has_tree;
    int width;          // 4 or 8
    uint32_t top;       // leading nodes (the top levels) the persistent traversal copies to LDS
    int leaf_batch = 16;  // lanes holding a leaf that send a wave to its leaf phase
};

// One SceneObject's material (material.h:20-70); same layout as rg_object.
struct DevObject {
    float color[3];
    float glossy, roughness, transparency, refraction, refraction_ratio;
    float metallic, anisotropy, aniso_rotation, emission;
    int32_t interior_medium;
    uint32_t visible;
    float chroma_dispersion;
    uint32_t pad;
};
static_assert(sizeof(DevObject) == sizeof(rg_object), "DevObject must mirror rg_object");

// Per-triangle shading record (original triangle order), 64 B:
// vertex normals, ti.normalizedpd, object index, triCount*area of its
// object (the light pdf's sampled area), ti.a and area
// (raycommon.cpp:128-149; rayrenderer.cpp:2030-2036).
struct DevTriShade {
    float n[9];
    float gn[3];
    uint32_t obj;
    float sampled_area;
    float ti_a;
    float area;
};
static_assert(sizeof(DevTriShade) == 64, "DevTriShade must be 64 bytes");

// HomogeneousMedium after prepare() (medium.cpp:18-57): the homogeneous
// coefficients, the fBm bound for delta tracking, the normalised cone axis,
// the heat-haze test, and the authored fields (positions in render space).
struct DevMedium {
    float sigma_t[3], sigma_s[3], sigma_e[3];
    float hero, g;
    int32_t active;
    int32_t het;          // fBm density field
    int32_t haze;         // isHeatHaze
    float t_max_hero;     // sigma_t_hero x the fBm peak (delta-tracking rate)
    float cone_axis[3];
    rg_medium src;
};

// An emissive volume for NEE (same layout as rg_volume_light).
struct DevVolumeLight {
    uint32_t medium;
    float pos[3];
};
static_assert(sizeof(DevVolumeLight) == sizeof(rg_volume_light), "DevVolumeLight must mirror rg_volume_light");

// Area light = an emissive object's triangles (transformScene,
// rayrenderer.cpp:333-354).
struct DevLight {
    uint32_t obj, first, count, pad;
};

// Point light (tracePointLight, rayrenderer.cpp:2205-2269); same layout as
// rg_point_light, spot_range replaced by the cone's cosine (0 with no cone).
struct DevPointLight {
    float pos[3], normal[3], color[3];
    float emission, spot_cos;
    int32_t spot;
};
static_assert(sizeof(DevPointLight) == sizeof(rg_point_light), "DevPointLight must mirror rg_point_light");

// NEE shadow-ray slots per path event: area light, point light,
// environment, emissive volume.
constexpr uint32_t kNeeSlots = 4;

// Everything a shade kernel reads, passed by value.
struct PTScene {
    const float *v9;
    const DevTriShade *shade;
    const DevObject *objs;
    const DevLight *lights;
    const uint32_t *light_tris;
    uint32_t nlights;
    const DevPointLight *plights;
    uint32_t nplights;
    const DevVolumeLight *vlights;
    uint32_t nvlights;
    const DevMedium *media;
    int32_t global_medium;
    // equirect environment
    const float *env_rgb;
    const float *env_marg;
    const float *env_cond;
    uint32_t env_w, env_h;
    float env_total, env_intensity, env_cos, env_sin;
    int32_t has_env;
    // albedo textures: per object (texture id or -1, texTiling u, v as
    // float bits), per texture (w, h, first texel), linear RGBA texels,
    // per-triangle UVs
    const int4 *otex;
    const uint4 *tex_desc;
    const float4 *texels;
    const float *uv6;
    int32_t has_tex;
    // camera (initRenderThreadContext, rayrenderer.cpp:97-132)
    float view_scale, half_w, half_h, aperture, dof, ap_rot;
    int32_t blades;
    // settings
    int32_t max_depth;
    int32_t antialias;
    uint32_t width, height;
    uint32_t seed, job;
    // RendererSettings::enableColorSampling and backColor (rgba)
    int32_t color_sampling;
    float back[4];
    // path events a generation runs (max_depth + 2, twice that when a
    // heat-haze medium can add an event without a pass)
    int32_t events;
    // the scene needs the shade kernels built with the extended media
    // (fBm density, cone / path emission, heat haze, volume-light NEE)
    int32_t ext;
    // shading system (5 = BSDF path tracer, 0-3 = the eye-ray providers),
    // calcAO's sample count (the render's spp) and worldColor
    int32_t shader;
    uint32_t ao_samples;
    float world[3];
};

// Structure-of-arrays path state, capacity = pixels x samples in flight.
// PTState::count slots of the extension and shadow-ray traversals' work-
// sharing counters: each is zero when its kernel starts (the generation's
// first kernel clears both, and each traversal kernel clears the other's).
constexpr int kExtendWork = 4, kShadowWork = 5;

struct PTState {
    rg_ray *ray;        // current ray
    rg_hit *hit;        // its closest hit
    float4 *beta;       // estimator weight, w = bsdfSampledPdf
    float4 *tp;         // BSDFParam::throughput (Russian roulette)
    float4 *L;          // radiance so far (without the last event's NEE), w = alive flag
    int4 *meta;         // passes, medium word (medium, dispersion band), next LDS dim, next Philox draw
    // NEE slots are slot-major: (path p, slot k) at k * slot_stride + p, so
    // a wave's paths write and read one slot's records contiguously
    rg_ray *sray;       // NEE shadow rays per (path, k), k < kNeeSlots (tmin = predicate)
    uint32_t *ssrc;     // queue of the (path, k) pairs with a shadow ray: k * slot_stride + path
    uint8_t *socc;      // occluded flag per (path, k)
    float4 *scontrib;   // unoccluded contribution per (path, k)
    uint32_t *nshadow;  // shadow rays per path this event
    uint32_t *q[2];     // live-path queues
    uint32_t *count;    // [0], [1] path-queue lengths, [2] shadow-ray queue length,
                        // [kExtendWork], [kShadowWork] the single-job traversals' work-sharing counters
    unsigned long long *stats; // [0] extension rays, [1] shadow rays, [2] path samples
    uint32_t sample_base;
    uint32_t slot_stride;  // path capacity of the slot arrays
};

// Kernel-side scene view of one job: scene buffers, camera and settings.
PTScene make_ptscene(const Scene &scn, const rg_settings &st);
// RG_E_INVALID unless the shading system is 0-3 or 5 and worldColor finite.
int check_shader(const rg_settings &st);
// Path events a generation runs: max_depth + 2 (every event but the eye
// hit's adds a pass, and the last pass ends the path), twice that when the
// scene has a heat-haze medium (a bend adds an event without a pass, and at
// most one bend follows each other event, its outer medium not being heat
// haze).
int path_events(const Scene &s, uint32_t max_depth);
// The scene has a medium with a density field, a cone or path emission
// profile or heat haze, or emissive-volume lights.
bool scene_uses_ext(const Scene &s);
// Validates shading inputs and uploads what changed since the last upload.
int prepare_scene(Scene *s, void *stream);

// One generation of S samples per pixel (the active tiles' when d_tile_mask
// is set) from sample_base on; with accumulate, its samples are added to the
// sums at the end.
int launch_generation(const Scene &scn, const PTScene &sc, const PTState &st, uint32_t S, uint32_t sample_base,
                      const uint8_t *d_tile_mask, float *d_sum, float *d_sumsq, float clamp_max, void *stream,
                      bool accumulate = true);
// Shading systems 0-3: one kernel computes every sample's value of the
// generation (eye ray, provider, transparency continuation, the global
// medium when one is set) into the path state's radiance, as the wavefront
// leaves it; the same accumulation follows.
int launch_preview(const Scene &scn, const PTScene &sc, const PTState &st, uint32_t S, uint32_t sample_base,
                   const uint8_t *d_tile_mask, float *d_sum, float *d_sumsq, float clamp_max, void *stream,
                   bool accumulate);
// Adds the finished generation's samples [s0, s0 + S) of the pixels of the
// tiles in d_tile_mask (all pixels when null) to the sums.
int launch_accumulate(const PTScene &sc, const PTState &st, uint32_t s0, uint32_t S, const uint8_t *d_tile_mask,
                      float *d_sum, float *d_sumsq, float clamp_max, void *stream);
int launch_resolve(const float *d_sum, const uint32_t *d_nspp, uint32_t width, uint32_t height, float exposure,
                   float *d_hdr, void *stream);
// Fills n 32-bit words with value.
int launch_fill_u32(void *p, uint32_t value, uint64_t n, void *stream);
// The per-triangle shading records (precalc, raycommon.cpp:128-149) from
// the device triangles, vertex normals and object ids; scale[o] = the
// object's triangle count when it emits, else 0 (sampled_area = scale x
// area).
int launch_tri_shade(const float *d_v9, const float *d_n9, const uint32_t *d_obj, const float *d_scale, uint32_t n,
                     DevTriShade *d_out, void *stream);

}  // namespace rg

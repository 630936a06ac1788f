/*
 * rg.h — C ABI of the MI355X-native RayGen BSDF path-tracing hot path.
 *
 * Plain C, no C++ types, no torch types.  Status codes are int (0 = ok,
 * negative = RG_E_*); the text of the last error on the calling thread is
 * returned by rg_last_error().  Caller owns host buffers; an rg_ctx owns its
 * device memory and HIP stream.  One driving host thread per rg_ctx.
 *
 * Each entry point names the reference interface it replaces
 * (jingwood/raygen-renderer, src/raygen/...).
 */
#ifndef RG_H
#define RG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define RG_OK 0
#define RG_E_INVALID -1     /* bad argument */
#define RG_E_NODEVICE -2    /* no HIP device (or a host-only context) */
#define RG_E_HIP -3         /* HIP runtime error */
#define RG_E_NOMEM -4       /* host or device allocation failed */
#define RG_E_CANCELLED -5   /* rg_cancel() observed */
#define RG_E_UNSUPPORTED -6 /* option outside the built path */
#define RG_E_BUDGET -7      /* request exceeds the context's memory budget */
#define RG_E_IO -8          /* file could not be read or parsed */
#define RG_E_BUSY -9        /* rg_server_submit: refused by the server's queue limit */
#define RG_E_INTERNAL -10   /* an internal consistency check failed */
#define RG_TIMEOUT 1        /* rg_server_wait: not finished within the timeout */

typedef struct rg_ctx rg_ctx;
typedef struct rg_scene rg_scene;

/* Ray, 32 B: origin, tmin, direction (need not be unit), tmax.
 * Replaces ugm::Ray + the initial RayTriangleIntersectionInfo::t
 * (raycommon.h:168-180; the reference's default tmax is 1e6). */
typedef struct {
    float o[3];
    float tmin;
    float d[3];
    float tmax;
} rg_ray;

/* Closest hit, 16 B: prim = index of the triangle in the order it was handed
 * to rg_scene_create (0xFFFFFFFF on miss); t; barycentrics u, v of
 * RenderMeshTriangle::intersectsRay (raycommon.cpp:200-235). */
typedef struct {
    uint32_t prim;
    float t;
    float u;
    float v;
} rg_hit;

/* Build parameters of TriangleBVH::build (bvh.cpp:22-25): bins = 16,
 * leaf = 4; width = 2 keeps the binary layout, 4 collapses to 4-wide nodes
 * (the traversal order and results are identical either way).  builder
 * picks where the tree is built; every builder gives the same nodes and
 * permutation. */
#define RG_BUILDER_AUTO 0   /* the GPU for device contexts from 2^14 triangles, else the host */
#define RG_BUILDER_HOST 1   /* the host's cores */
#define RG_BUILDER_DEVICE 2 /* the context's GPU (bins <= 32) */
typedef struct {
    uint32_t bins;
    uint32_t leaf;
    uint32_t width;
    uint32_t builder;
} rg_bvh_params;

/* 32-byte binary node, layout of BVHNode (bvh.h:22-30); used by
 * rg_bvh_export2 so callers can inspect the tree. */
typedef struct {
    float bmin[3];
    uint32_t first_or_left;
    float bmax[3];
    uint16_t count;
    uint16_t axis;
} rg_bvh2_node;

/* Occlusion predicates of the reference's intersectAny call sites. */
#define RG_PRED_ALL 0  /* every triangle occludes (microbenchmark)          */
#define RG_PRED_AREA 1 /* transparency<0.01 && emission<=0 (rayrenderer.cpp:1956-1959) */
#define RG_PRED_ENV 2  /* !(emission>0) && (transparency<0.01 || refraction>0.1) (:1499-1503; point lights :2211-2215) */
#define RG_PRED_VOLUME 3 /* !(emission>0) && transparency<0.01 && refraction<=0.1 (volume lights, :2172-2176) */
#define RG_PRED_OPAQUE 4 /* transparency<0.01 (calcAO, :2391-2393) */

/* ---- context ------------------------------------------------------------ */
/* device >= 0: HIP device ordinal.  device == -1: host-only context (scene
 * building and BVH export work, every device call returns RG_E_NODEVICE). */
int rg_create(int device, rg_ctx **out);
void rg_destroy(rg_ctx *ctx);
const char *rg_last_error(void);
/* Launch subsequent work on `hip_stream` (a hipStream_t; NULL = the
 * context's own stream). */
int rg_set_stream(rg_ctx *ctx, void *hip_stream);
int rg_synchronize(rg_ctx *ctx);
/* Page-locked host memory (no reference counterpart; for callers that want
 * rg_render's images delivered by DMA straight into their buffers instead of
 * through the HIP runtime's staging copies of pageable memory).  Free with
 * rg_host_free. */
int rg_host_alloc(uint64_t bytes, void **out);
void rg_host_free(void *p);
/* Cooperative cancel (RayRenderer::cancelRequested, rayrenderer.h:427):
 * async-safe; observed between launches.  Like the reference, which clears
 * the flag when a render starts (rayrenderer.cpp:537), rg_render and
 * rg_render_batch clear it on entry, so it cancels the render in flight
 * (call it from another thread or from the progress callback); a request
 * server fails its queued and running requests and clears it. */
void rg_cancel(rg_ctx *ctx);

/* Progress of rg_render (RendererSettings::progressCallback,
 * rayrenderer.h:317, called at rayrenderer.cpp:804-806 and :1155-1156): fn
 * receives the fraction of the sample budget whose paths have finished,
 * non-decreasing, on the thread that called rg_render (a generation behind
 * the GPU, which keeps running), and 1.0 when the render is done.  NULL
 * removes it. */
typedef void (*rg_progress_fn)(float progress, void *user);
int rg_set_progress(rg_ctx *ctx, rg_progress_fn fn, void *user);

/* ---- scene geometry ----------------------------------------------------- */
/* Replaces RayRenderer::transformScene's triangle list (rayrenderer.cpp:157-363):
 * v9 = n_tris x (v1, v2, v3) in render (camera) space.  The data is copied. */
int rg_scene_create(rg_ctx *ctx, const float *v9, uint32_t n_tris, rg_scene **out);
void rg_scene_destroy(rg_scene *scene);
/* Replaces TriangleBVH::build (bvh.cpp:76-231): the binned-SAH binary tree,
 * its 4-wide (or 8-wide) collapse and, for device contexts, the triangle
 * records in HBM.  Built on the GPU (device contexts, RG_BUILDER_AUTO from
 * 2^14 triangles, or RG_BUILDER_DEVICE) or on the host's cores.  params may
 * be NULL (16 bins, leaf 4, width 4, RG_BUILDER_AUTO). */
int rg_bvh_build(rg_ctx *ctx, rg_scene *scene, const rg_bvh_params *params);
/* Copies the binary tree out: nodes (capacity cap) and the primitive
 * permutation (n_tris entries, nullable).  *n_nodes receives the count. */
int rg_bvh_export2(const rg_scene *scene, rg_bvh2_node *nodes, uint64_t cap,
                   uint32_t *perm, uint64_t *n_nodes);
/* Device-side footprint of the scene's tree (triangle records + wide nodes,
 * and after a GPU build also its binary nodes and permutation), in bytes. */
int rg_scene_device_bytes(const rg_scene *scene, uint64_t *bytes);
/* Bounding box of the geometry: lo xyz, hi xyz. */
int rg_scene_bounds(const rg_scene *scene, float box[6]);

/* ---- traversal ---------------------------------------------------------- */
/* Replaces TriangleBVH::intersectClosest (bvh.cpp:233-285) for a batch.
 * d_rays/d_hits are DEVICE pointers (HBM-resident batch). */
int rg_trace_closest(rg_ctx *ctx, const rg_scene *scene, const rg_ray *d_rays,
                     rg_hit *d_hits, uint64_t n);
/* Replaces TriangleBVH::intersectAny<Pred> (bvh.h:63-101) with maxT =
 * ray.tmax; pred is one of RG_PRED_*.  d_occluded: one byte per ray. */
int rg_trace_any(rg_ctx *ctx, const rg_scene *scene, const rg_ray *d_rays,
                 uint8_t *d_occluded, uint64_t n, int pred);
/* Host-buffer conveniences (copy in, trace, copy out; synchronous). */
int rg_trace_closest_host(rg_ctx *ctx, const rg_scene *scene, const rg_ray *rays,
                          rg_hit *hits, uint64_t n);
int rg_trace_any_host(rg_ctx *ctx, const rg_scene *scene, const rg_ray *rays,
                      uint8_t *occluded, uint64_t n, int pred);

/* ---- workload generators (traversal microbenchmark) --------------------- */
/* Level-`level` subdivided icosahedron (20*4^level triangles) on the unit
 * sphere, radially displaced by 1 + amp*(2*fbm(p)-1) with 3-octave value
 * noise (medium.cpp:125-185 basis) keyed by `seed`.  out_v9 holds
 * 20*4^level*9 floats. */
int rg_gen_icosphere(uint32_t level, float amp, uint32_t seed, float *out_v9);
/* Incoherent rays on the device: origins uniform in scale x box (centred),
 * directions uniform on the unit sphere, Philox4x32-10 keyed by `key`,
 * counter = ray index + first.  tmin = 0, tmax = +inf. */
int rg_gen_rays_sphere(rg_ctx *ctx, const float box[6], float scale, uint32_t key,
                       uint64_t first, rg_ray *d_rays, uint64_t n);

/* ---- path tracing, shading system 5 (-d 5) ------------------------------ */
/* Material of one scene object: the Material fields (material.h:20-70) that
 * RayBSDFShaderProvider reads.  Defaults (rg_object_default) follow
 * material.h: color 0.8, roughness 0.5, refraction_ratio 1.45, rest 0. */
typedef struct {
    float color[3];
    float glossy, roughness, transparency, refraction, refraction_ratio;
    float metallic, anisotropy, aniso_rotation, emission;
    int32_t interior_medium; /* index into the scene's media, -1 none */
    uint32_t visible;
    /* Material::chromaDispersion (material.h:36-39): a path's first
     * dispersive refraction picks one of three bands, index of refraction
     * refraction_ratio * (1 - s), * 1, * (1 + s) (bsdf.cpp:394-420) */
    float chroma_dispersion;
    uint32_t pad;
} rg_object;

/* One control point of a path-emission tube (HomogeneousMedium::PathSample,
 * medium.h:54-62): position (render space), tube radius, axial parameter. */
typedef struct {
    float p[3];
    float radius, t;
} rg_path_point;
#define RG_MAX_PATH_POINTS 16

/* HomogeneousMedium authoring fields (medium.h:37-214).  Positions and axes
 * (cone_origin, cone_axis, path points) are in render space: the loader
 * bakes them as HomogeneousMedium::bake does (medium.cpp:59-104).
 * rg_medium_default gives the reference's defaults. */
typedef struct {
    float sigma_a[3], sigma_s[3], sigma_e[3];
    float g, density;
    /* emission profile (EmissionMode): 0 constant sigma_e, 1 cone, 2 path */
    int32_t emission_mode;
    float cone_origin[3], cone_axis[3];
    float cone_length, cone_radius;
    float cone_inner[3], cone_outer[3];
    float cone_intensity, cone_peak_axial, cone_peak_sharpness;
    int32_t cone_samples;                 /* coneEmissionSamples */
    rg_path_point path[RG_MAX_PATH_POINTS];
    uint32_t n_path;
    float path_inner[3], path_outer[3];
    float path_intensity, path_falloff_power;
    int32_t path_samples;                 /* pathEmissionSamples */
    /* density field (DensityFieldMode): 0 none, 1 fBm value noise */
    int32_t density_field;
    float noise_frequency;
    int32_t noise_octaves;
    float noise_gain, noise_lacunarity, noise_amplitude, noise_bias;
    float noise_offset[3];
    /* heat haze: the ray bends by the gradient of n = 1 + amplitude * fbm */
    int32_t heat_haze;
    float ior_amplitude, ior_frequency;
    int32_t ior_octaves;
    float ior_gain, ior_lacunarity;
    int32_t ior_march_steps;
    float ior_offset[3], ior_falloff;
} rg_medium;

/* An emissive volume that NEE samples (RayRenderer's
 * emissiveVolumeSources, rayrenderer.cpp:180-206): a medium with a cone
 * profile of positive intensity, or constant non-zero emission, inside an
 * object; pos = that object's own location through the view matrix (the
 * point a constant-emission volume is sampled at, :2089-2096). */
typedef struct {
    uint32_t medium;
    float pos[3];
} rg_volume_light;

/* Camera (scene.h Camera): render space has the eye at the origin looking
 * down -Z, as after RayRenderer's view transform (rayrenderer.cpp:247). */
typedef struct {
    float fov_deg, exposure, depth_of_field, aperture;
    int32_t aperture_blades;
    float aperture_rotation_deg, view_far, pad;
} rg_camera;

/* Equirectangular environment map: w x h linear RGB floats, row 0 = top
 * (Scene::envmap with envmapIntensity / envmapRotation). */
typedef struct {
    uint32_t w, h;
    const float *rgb;
    float intensity, rotation_deg;
} rg_env;

/* RendererSettings (rayrenderer.h:98-153) subset of the -d 5 path plus the
 * seeding this build adds (the reference seeds from time()). */
typedef struct {
    uint32_t width, height;   /* -r WxH */
    uint32_t spp;             /* -s N */
    uint32_t max_depth;       /* MAX_TRACE_DEPTH (reference constant 32) */
    int32_t antialias;        /* -enaa */
    float firefly_clamp;      /* fireflyClamp, 0 disables */
    uint32_t seed, job;       /* Philox key of the path's random draws */
    int32_t adaptive;         /* enableAdaptiveSampling */
    uint32_t adaptive_base;   /* adaptiveBaseSamples */
    float adaptive_threshold; /* adaptiveThreshold (mean tile rSEM) */
    int32_t denoise;          /* enableDenoise (A-Trous) */
    uint32_t denoise_levels;
    float sigma_color, sigma_normal, sigma_depth, denoise_intensity;
    uint32_t samples_in_flight; /* paths per launch wave = pixels x this; 0 = auto */
    /* enableColorSampling (-encs, rayrenderer.h:106; read at bsdf.cpp:69,
     * 186 and rayrenderer.cpp:2844): 0 = diffuse albedo 1, glossy F0 from the
     * untextured colour, depth-cap light term 0 */
    int32_t color_sampling;
    /* cullBackFace (-cb, rayrenderer.h:110): accepted and kept; the
     * reference's shading system 5 never reads it (rayrenderer.cpp:82 only
     * stores it), so it changes nothing here either */
    int32_t cull_back_face;
    /* backColor (-bc, rayrenderer.h:152; default (1, 0.95, 0.9, 0) * 0.2):
     * the radiance of an eye ray that sees neither geometry nor environment
     * (rayrenderer.cpp:1273, 1296) and the albedo guide of a background pixel
     * (:841); alpha is kept for the guide only */
    float back_color[4];
    /* shaderProvider (-d, rayrenderer.cpp:59-77): 5 = the BSDF path tracer
     * (default); 0 = RaySimpleShaderProvider, 1 = RayAmbientOcclusion-
     * ShaderProvider, 2 = LambertShaderProvider, 3 = LambertWithAO-
     * ShaderProvider (rayrenderer.h:446-475, lambert.cpp:14-67).  Systems
     * 0-3 take one eye ray per sample and no bounces; their calcAO casts
     * spp rays per sample (settings.samples, rayrenderer.cpp:2385). */
    int32_t shader;
    /* worldColor (-wc, rayrenderer.h:151; default (1, 0.95, 0.9) * 0.1):
     * the ambient term lambertTraceLights adds (rayrenderer.cpp:2334) */
    float world_color[3];
} rg_settings;

/* Point light: an emissive scene object without meshes, which
 * transformScene lists as a point light source (rayrenderer.cpp:333-353).
 * pos = the object's origin and normal = its direction
 * normalize(cos(ay) sin(ax), cos(ay) cos(ax), sin(ay)) after the
 * normal matrix, both in render space; color, emission and spot_range are
 * its material's color, emission and spotRange (material.h:50-52; 0 = omni,
 * else the cone angle read through RADIAN_TO_DEGREE as the reference does,
 * rayrenderer.cpp:2228-2232). */
typedef struct {
    float pos[3], normal[3], color[3];
    float emission, spot_range;
    uint32_t pad;
} rg_point_light;

/* One material texture (Texture, texture.h:19-50): w x h texels of RGB
 * floats as stored in the image (LDR: byte / 255, HDR: linear radiance), row
 * 0 = top.  srgb != 0 marks an LDR image authored in sRGB: its texels are
 * converted to linear as Texture::sample does (texture.cpp:199-235). */
typedef struct {
    uint32_t w, h;
    const float *rgb;
    int32_t srgb;
    uint32_t pad;
} rg_texture;

void rg_object_default(rg_object *obj);
void rg_medium_default(rg_medium *m);
void rg_settings_default(rg_settings *s);

/* Per-triangle shading inputs of RayRenderer::transformObject
 * (rayrenderer.cpp:237-363): n9 = vertex normals in render space,
 * obj_of_tri = owning object index; objs = one material per object.  An
 * object with emission > 0 is an area light made of its triangles. */
int rg_scene_set_shading(rg_scene *scene, const float *n9, const uint32_t *obj_of_tri,
                         const rg_object *objs, uint32_t n_obj);
/* Albedo textures (Material::texture + texTiling, material.h:54-60, sampled
 * by the diffuse and glossy lobes and the depth-cap light term, bsdf.cpp:67-73,
 * 183-188; rayrenderer.cpp:2843-2850): tex_of_obj[n_obj] names each object's
 * texture (-1 none), tiling_of_obj[n_obj * 2] its texTiling, uv6[n_tris * 6]
 * every triangle's first UV set (RenderMeshTriangle uv1..uv3).  Call after
 * rg_scene_set_shading; n_tex == 0 removes all textures. */
int rg_scene_set_textures(rg_scene *scene, const rg_texture *tex, uint32_t n_tex, const int32_t *tex_of_obj,
                          const float *tiling_of_obj, const float *uv6);
/* Media referenced by rg_object.interior_medium; global_medium = the
 * scene's globalMedium (-1 none).  At most 32767 media.  A heat-haze global
 * medium is refused (RG_E_UNSUPPORTED): there the reference's tracePath
 * recurses without end (rayrenderer.cpp:1645-1663 with no BSDFParam). */
int rg_scene_set_media(rg_scene *scene, const rg_medium *media, uint32_t n, int32_t global_medium);
/* The scene's emissive volumes for NEE (the diffuse and glossy lobes and
 * in-medium scattering, bsdf.cpp:92-103, 318-329; rayrenderer.cpp:1756-1768):
 * one picked uniformly per call, its shadow ray blocked by RG_PRED_VOLUME
 * triangles.  n == 0 removes them. */
int rg_scene_set_volume_lights(rg_scene *scene, const rg_volume_light *v, uint32_t n);
/* Copies them out: n first, then v (n records) where non-NULL. */
int rg_scene_export_volume_lights(const rg_scene *scene, rg_volume_light *v, uint32_t *n);
/* env NULL or w == 0 removes the environment. */
int rg_scene_set_env(rg_scene *scene, const rg_env *env);
/* Point light sources of RayRenderer::traceLight (rayrenderer.cpp:2205-2290;
 * the pointLightSources list built at :333-353): the diffuse lobe's direct
 * term and the depth-cap light term add one light's tracePointLight (the
 * only one, or one picked uniformly), occluded by the triangles that block
 * environment rays (RG_PRED_ENV).  n == 0 removes them. */
int rg_scene_set_point_lights(rg_scene *scene, const rg_point_light *lights, uint32_t n);
int rg_scene_set_camera(rg_scene *scene, const rg_camera *cam);

/* Replaces RayRenderer::render() (rayrenderer.cpp:478-596) for shader 5:
 * samples every pixel (adaptive per 32x32 tile when enabled), optionally
 * denoises, and writes the linear HDR image (mean * exposure, >= 0; RGBA
 * float, W*H*4) to hdr_rgba.  sum_rgb / sumsq_rgb (nullable, W*H*3) receive
 * the per-pixel sums of the clamped sample radiance and of its square;
 * nspp (nullable, one per 32x32 tile, row-major) the samples taken.  All
 * output pointers are host memory.  Blocks until done. */
int rg_render(rg_ctx *ctx, rg_scene *scene, const rg_settings *settings, float *hdr_rgba,
              float *sum_rgb, float *sumsq_rgb, uint32_t *nspp);

/* ---- output stage ------------------------------------------------------- */
/* Bloom knobs of RendererSettings (rayrenderer.h:137-144; -blth -blst -blrd
 * -blcv).  Defaults (rg_bloom_default) are the reference's. */
typedef struct {
    float threshold;   /* bloomThreshold: linear luma where bloom starts (1.0) */
    float strength;    /* bloomStrength: gain on the blurred halo (1.0) */
    float radius;      /* bloomRadius: halo sigma as a fraction of the width (0.03) */
    float size_aspect; /* bloomSizeAspect: glow buffer size / image size (0.15) */
    float curve;       /* bloomCurve: knee sharpness, 1 = linear (1.0) */
    float pad;
} rg_bloom;
void rg_bloom_default(rg_bloom *b);

/* Replaces RayRenderer::applyPostProcess + applyTonemapGamma, the stage that
 * ends render() and reapplyPostProcess() (rayrenderer.cpp:598-697,
 * 2614-2647, 701-712).  rg_render returns the pre-bloom HDR image; this
 * call takes it on: when bloom is non-NULL (enableRenderingPostProcess,
 * -enpp) the energy-based HDR bloom is added to hdr_rgba in place, then, when
 * ldr_rgba is non-NULL, the luminance-Reinhard + 1/2.2-gamma image is
 * written there (RGBA floats in [0, 1], alpha copied).  Host buffers of
 * W*H*4 floats; blocks until done. */
int rg_postprocess(rg_ctx *ctx, uint32_t width, uint32_t height, float *hdr_rgba, const rg_bloom *bloom,
                   float *ldr_rgba);
/* The same on DEVICE images, queued on the context's stream (asynchronous). */
int rg_postprocess_device(rg_ctx *ctx, uint32_t width, uint32_t height, float *d_hdr_rgba, const rg_bloom *bloom,
                          float *d_ldr_rgba);

/* Image writers of the render CLI's output stage (src/main.cpp:400-410,
 * 479-485, saveImage): the format follows the extension of path.  ".hdr"
 * writes Radiance RGBE of the linear HDR image; ".jpg"/".jpeg" (baseline
 * JPEG, 4:4:4, quality 1-100, <= 0 = 90), ".png" (8-bit RGB), ".ppm" and
 * ".bmp" write an LDR image (RGBA floats in [0, 1], e.g. rg_postprocess's
 * ldr_rgba) as bytes round(v * 255).  RG_E_UNSUPPORTED for another
 * extension, RG_E_IO when the file cannot be written. */
int rg_image_write(const char *path, uint32_t width, uint32_t height, const float *rgba, int quality);

/* rg_postprocess_device without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as rg_postprocess_device (same kernels, the
 * bloomed HDR and LDR images are bit-identical), but the two glow buffers
 * live in a fixed per-GPU scratch array of 2 x 1,048,576 glow pixels
 * (32 MiB, held until the process ends) and the Gaussian taps are passed as
 * kernel arguments (at most 127 per side).  A glow buffer larger than the
 * scratch (image size x size_aspect, per axis) or a wider blur fails with
 * RG_E_BUDGET.  Calls that overlap on one GPU (two contexts or streams at
 * once) share the scratch and race; calls queued on one stream do not. */
int rg_postprocess_device_noalloc(rg_ctx *ctx, uint32_t width, uint32_t height, float *d_hdr_rgba,
                                  const rg_bloom *bloom, float *d_ldr_rgba);

/* ---- scene ingest ------------------------------------------------------- */
/* Replaces RendererSceneLoader::load + RayRenderer::transformScene
 * (sceneloader.cpp:847-882, 505-783; rayrenderer.cpp:157-363): reads the
 * reference's scene JSON (comments and trailing commas allowed) with its
 * objects, location/angle/scale transforms, plane/cube/sphere primitives,
 * binary .mesh files, Wavefront .obj files with their .mtl libraries (one
 * child object per OBJ object, ObjFileReader, objreader.cpp:93-590;
 * sceneloader.cpp:405-478), inline and `_materials` materials, object and
 * global `medium`, `mainCamera`, an .hdr `envmap`, albedo textures and
 * point lights (emissive objects without meshes), and returns a scene in
 * render space with shading, media, environment, textures, point lights and
 * camera set (BVH not built).  Parts outside this path (FBX, .toba,
 * cubemaps, non-homogeneous media) are skipped with a warning
 * (rg_scene_warnings). */
int rg_scene_load_json(rg_ctx *ctx, const char *path, rg_scene **out);
/* Newline-separated warnings of the load into buf (capacity cap); returns the
 * full length. */
int rg_scene_warnings(const rg_scene *scene, char *buf, uint64_t cap);
int rg_scene_get_camera(const rg_scene *scene, rg_camera *cam);
/* Copies the flattened scene out: n_tris and n_obj first (any array
 * argument may be NULL), then v9/n9 (n_tris*9 floats), obj_of_tri, objs. */
int rg_scene_export(const rg_scene *scene, float *v9, float *n9, uint32_t *obj_of_tri, rg_object *objs,
                    uint32_t *n_tris, uint32_t *n_obj);
/* Copies the media table and environment out: n_media, global_medium and env
 * (env->rgb is left NULL) first, then media (n_media records) and env_rgb
 * (env->width * env->height * 3 floats) where non-NULL. */
int rg_scene_export_media(const rg_scene *scene, rg_medium *media, uint32_t *n_media, int32_t *global_medium,
                          rg_env *env, float *env_rgb);

/* Copies the point lights out: n first, then lights (n records) where
 * non-NULL. */
int rg_scene_export_point_lights(const rg_scene *scene, rg_point_light *lights, uint32_t *n);

/* Copies the textures out: n_tex first, then, where non-NULL, tex (w, h,
 * srgb of each; rgb left NULL), texels (the textures' RGB floats as handed
 * in, concatenated), tex_of_obj (n_obj), tiling_of_obj (n_obj * 2) and uv6
 * (n_tris * 6; zeros when the scene has no textures). */
int rg_scene_export_textures(const rg_scene *scene, uint32_t *n_tex, rg_texture *tex, float *texels,
                             int32_t *tex_of_obj, float *tiling_of_obj, float *uv6);

/* Counters of the last rg_render on this context: [0] extension (closest-hit)
 * rays, [1] NEE shadow (any-hit) rays, [2] path samples. */
int rg_render_stats(rg_ctx *ctx, uint64_t out[3]);

/* ---- many jobs on one GPU ------------------------------------------------
 * The reference renders one job per RayRenderer::render call
 * (rayrenderer.cpp:478-596, driven by main.cpp:162-490).  The entry points
 * below run many such jobs together: the admitted jobs' samples are packed
 * into shared wavefront generations (each job contributes chunks of S
 * samples of every pixel, S being rg_render's samples in flight for that
 * job), the bounce loop runs over one compacted queue of all their paths,
 * and each job's image is what rg_render returns for the same scene and
 * settings (same per-sample random streams, same per-pixel summation
 * order).  Jobs render a fixed sample count: settings.adaptive and
 * settings.denoise must be 0. */
typedef struct {
    uint32_t scene;        /* index into the scene list of the batch / server */
    uint32_t pad;
    rg_settings settings;  /* width, height, spp, max_depth, antialias,
                              firefly_clamp, seed, job, samples_in_flight */
} rg_job;

/* Engine knobs; 0 picks the default.  rg_autotune measures candidates. */
typedef struct {
    uint32_t pool_paths;       /* path slots of one generation (default 2^25; a
                                  generation always takes at least one chunk) */
    uint32_t stack;            /* traversal stack entries 32/64/128/256
                                  (default: smallest that fits every BVH) */
    uint32_t lanes;            /* streams with their own path state that consecutive
                                  generations rotate over, 1 to 4 (default 3) */
    uint32_t blocks_per_cu;    /* cap on persistent traversal blocks per compute unit
                                  (default 0: as many as the device holds) */
    uint32_t iters_per_graph;  /* nonzero: replay each generation's bounce loop from a
                                  captured hipGraph (default 0: direct launches) */
    uint32_t pad;
} rg_tune;

/* Renders n_jobs jobs over the scenes (BVH built, shading set) and blocks
 * until all are done.  hdr_out[i] (W*H*4 floats) and sum_out[i] (W*H*3) are
 * host buffers (the arrays and their entries may be NULL).  stats (nullable):
 * [0] extension rays, [1] shadow rays, [2] samples, [3] generations. */
int rg_render_batch(rg_ctx *ctx, rg_scene *const *scenes, uint32_t n_scenes, const rg_job *jobs, uint32_t n_jobs,
                    const rg_tune *tune, float *const *hdr_out, float *const *sum_out, uint64_t stats[4]);

/* Times every candidate of a small grid of generation size / stack / lanes /
 * blocks per CU / graph settings on the probe jobs (NULL: one 512x384, 8 spp,
 * depth 6 job per scene) and returns the fastest. */
int rg_autotune(rg_ctx *ctx, rg_scene *const *scenes, uint32_t n_scenes, const rg_job *probe, uint32_t n_probe,
                rg_tune *out);

/* Request server: keeps the scenes resident, admits requests first come,
 * first served as a job slot and hbm_budget bytes of framebuffer space
 * (sums, sums of squares, image) allow, and renders every admitted request
 * together with the others on a thread of its own: a request admitted while
 * generations run joins the next generation. */
typedef struct rg_server rg_server;
typedef struct {
    int32_t status;     /* RG_OK or the failure */
    uint32_t pad;
    uint64_t samples;   /* W*H*spp */
    double submit_s, admit_s, done_s;  /* library steady clock, seconds */
} rg_request_info;

int rg_server_create(rg_ctx *ctx, rg_scene *const *scenes, uint32_t n_scenes, uint64_t hbm_budget,
                     const rg_tune *tune, rg_server **out);
/* Queues one request; hdr_out / sum_out (nullable host buffers) are written
 * before rg_server_wait reports it done. */
int rg_server_submit(rg_server *srv, const rg_job *job, float *hdr_out, float *sum_out, uint64_t *ticket);
/* Waits up to timeout_s (< 0: forever) for a request; RG_TIMEOUT if still
 * running.  A finished ticket is forgotten after it is reported. */
int rg_server_wait(rg_server *srv, uint64_t ticket, double timeout_s, rg_request_info *info);
/* [0] requests done, [1] samples done, [2] extension rays, [3] shadow rays,
 * [4] generations, [5] framebuffer bytes in use, [6] budget, [7] graph replays */
int rg_server_stats(rg_server *srv, uint64_t out[8]);
int rg_server_get_tune(rg_server *srv, rg_tune *out);
/* Overload policy of the request server (the reference has no server; this
 * bounds the latency of what it accepts): with a nonzero limit,
 * rg_server_submit refuses with RG_E_BUSY a request whose samples (W*H*spp)
 * would take the samples submitted and not yet done past
 * max_outstanding_samples; a request is always accepted when nothing is
 * outstanding.  0 (the default) removes the limit. */
int rg_server_set_queue_limit(rg_server *srv, uint64_t max_outstanding_samples);
/* [0] samples outstanding, [1] the limit, [2] requests refused so far. */
int rg_server_queue(rg_server *srv, uint64_t out[3]);
int rg_server_destroy(rg_server *srv);

#ifdef __cplusplus
}
#endif
#endif

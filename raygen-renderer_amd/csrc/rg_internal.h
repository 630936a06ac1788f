// rg_internal.h — shared host-side definitions of the rg library.
#pragma once

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/rg.h"

namespace rg {

// ---------------------------------------------------------------------------
// Device layouts
// ---------------------------------------------------------------------------

// 4-wide node, 128 B (one L2 line on gfx950).  Child boxes are stored SoA so
// one node fetch is 8 x 16 B vector loads.  child[i]:
//   bit 31 clear : index of a child Node4
//   bit 31 set   : leaf; bits 0..30 = first triangle in the permuted triangle
//                  array (the leaf runs until a triangle flagged "last")
//   0xFFFFFFFF   : empty slot
// order[oct] lists the 4 slots (2 bits each, first visited in bits 0-1) in the
// depth-first order the binary tree would visit them for a ray whose
// direction-sign octant is oct (bit k set when 1/d_k < 0); rank[oct] is its
// inverse, each slot's visit position (2 bits per slot, slot 0 in bits 0-1).
struct alignas(128) Node4 {
    float lox[4], loy[4], loz[4];
    float hix[4], hiy[4], hiz[4];
    uint32_t child[4];
    uint8_t order[8];
    uint8_t rank[8];
};
static_assert(sizeof(Node4) == 128, "Node4 must be 128 bytes");

// 8-wide node, 256 B (two L2 lines): the same SoA boxes for 8 children, 8
// child refs (same encoding as Node4), and per octant a 24-bit visit order
// (3 bits per slot, first visited in bits 0-2).
struct alignas(128) Node8 {
    float lox[8], loy[8], loz[8];
    float hix[8], hiy[8], hiz[8];
    uint32_t child[8];
    uint32_t order[8];
};
static_assert(sizeof(Node8) == 256, "Node8 must be 256 bytes");

// Default path budget of the single-job path: samples_in_flight = 0 runs
// max(1, kPathsInFlight / pixels) samples of every pixel per launch wave
// (capped by spp, and by the adaptive pass size).  The batched engine sums
// each pixel's samples in groups of the same size, so both stay
// bit-identical.
constexpr uint32_t kPathsInFlight = 1u << 25;

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;

// Triangle record, 48 B = 3 x float4, in BVH (permuted) order:
//   a = (v1.xyz, bits(original index)), b = (v2-v1, bits(flags)), c = (v3-v1, 0)
// flags bit0 = last triangle of its leaf; c.w holds the occluder bits
// (bit p set when predicate RG_PRED_p counts the triangle as an occluder).  Edges are the exact binary32
// differences RenderMeshTriangle::intersectsRay forms (raycommon.cpp:410-411).
struct alignas(16) Tri48 {
    float a[4];
    float b[4];
    float c[4];
};
static_assert(sizeof(Tri48) == 48, "Tri48 must be 48 bytes");

// Binary node, identical to BVHNode (bvh.h:22-30).
struct Node2 {
    float bmin[3];
    uint32_t first_or_left;
    float bmax[3];
    uint16_t count;
    uint16_t axis;
};
static_assert(sizeof(Node2) == 32, "Node2 must be 32 bytes");

// ---------------------------------------------------------------------------
// Host objects
// ---------------------------------------------------------------------------
// Device buffers of the wavefront path tracer, grown on demand and reused
// across rg_render calls on one context.
struct PTBuffers {
    uint64_t cap = 0;       // path slots
    uint64_t npix_cap = 0;  // pixels
    uint64_t ntile_cap = 0; // 32x32 tiles
    void *ray = nullptr, *hit = nullptr, *beta = nullptr, *tp = nullptr, *L = nullptr, *meta = nullptr;
    void *sray = nullptr, *ssrc = nullptr, *socc = nullptr, *scontrib = nullptr, *nshadow = nullptr;
    void *q0 = nullptr, *q1 = nullptr;
    void *count = nullptr;
    void *sum = nullptr, *sumsq = nullptr, *hdr = nullptr;
    void *tile_mask = nullptr, *tile_nspp = nullptr, *tile_noise = nullptr, *n_active = nullptr;
    void *aov = nullptr, *dn_a = nullptr, *dn_b = nullptr;
    void *stats = nullptr;
};

// Device buffers of the output stage (bloom glow ping-pong, Gaussian taps,
// staging images of the host-buffer entry point), grown on demand.
struct PostBuffers {
    void *glow_a = nullptr, *glow_b = nullptr, *taps = nullptr, *img_hdr = nullptr, *img_ldr = nullptr;
    uint64_t glow_a_cap = 0, glow_b_cap = 0, taps_cap = 0, img_hdr_cap = 0, img_ldr_cap = 0;
    // Gaussian taps now on the device and the sigma they were built for
    // (the host copy is the source of the asynchronous upload)
    std::vector<float> taps_host;
    float taps_sigma = -1.0f;
    int taps_r = 0;
};

struct Ctx {
    int device = -1;
    void *stream = nullptr;        // hipStream_t in use
    void *own_stream = nullptr;    // stream created by rg_create
    std::atomic<bool> cancel{false};
    rg_progress_fn progress = nullptr;  // rg_render's progress callback
    void *progress_user = nullptr;
    void *progress_ev[2] = {nullptr, nullptr};  // hipEvent_t after the last two generations
    void *lane_streams[4] = {nullptr, nullptr, nullptr, nullptr};  // the batched engines' lanes, kept
    PTBuffers pt;
    PostBuffers post;
    // work-sharing counters of rg_trace_closest / rg_trace_any, one per
    // launch in turn, so launches on different streams do not share one
    // unless kTraceSlots launches apart
    static constexpr uint32_t kTraceSlots = 64;
    uint64_t *trace_work = nullptr;
    uint32_t trace_slot = 0;
    // the GPU BVH builder's scratch, kept between builds (grown on demand)
    // and its page-locked word for the per-level counters
    void *build_ws = nullptr;
    uint64_t build_ws_cap = 0;
    void *build_host = nullptr;
    uint64_t last_stats[3] = {0, 0, 0};
};

struct Scene {
    int device = -1;                  // device holding the scene's buffers (-1: none)
    uint32_t n_tris = 0;
    std::vector<float> v9;            // as handed in
    float box[6] = {0, 0, 0, 0, 0, 0};
    // binary tree (reference layout) + permutation; after a GPU build they
    // stay on the device (d_nodes2, d_perm) until a host reader asks
    std::vector<Node2> nodes2;
    std::vector<uint32_t> perm;
    bool host_tree = true;            // nodes2 / perm hold the current tree
    uint32_t n_nodes2 = 0;            // binary nodes
    uint64_t n_wide = 0;              // wide nodes
    float root_box[6] = {0, 0, 0, 0, 0, 0};  // the binary root's bounds
    // wide tree
    std::vector<Node4> nodes4;      // width 2 or 4
    std::vector<Node8> nodes8;      // width 8
    bool root_is_leaf = false;
    uint32_t root_leaf_first = 0;
    int width = 4;
    uint32_t stack_need = 0;       // deepest traversal stack the wide tree can need
    bool built = false;
    // device copies
    void *d_nodes4 = nullptr;        // the wide nodes on the device (Node4 or Node8 by width)
    void *d_nodes2 = nullptr;        // binary nodes of a GPU build
    void *d_perm = nullptr;          // and its permutation
    void *d_tris = nullptr;
    void *d_tri_flags = nullptr;  // per-triangle occluder mask bits (predicates)
    uint64_t device_bytes = 0;
    // per-triangle predicate mask: bit p set when predicate p treats it as occluder
    std::vector<uint8_t> pred_mask;

    // shading inputs (rg_scene_set_*)
    bool has_shading = false;
    std::vector<float> n9;
    std::vector<uint32_t> obj_of_tri;
    std::vector<rg_object> objects;
    std::vector<rg_medium> media;
    int32_t global_medium = -1;
    uint32_t env_w = 0, env_h = 0;
    std::vector<float> env_rgb;
    float env_intensity = 1.0f, env_rotation_deg = 0.0f;
    rg_camera camera{75.0f, 1.0f, 0.0f, 1.8f, 0, 0.0f, 50.0f, 0.0f};
    std::vector<rg_point_light> point_lights;
    std::vector<rg_volume_light> volume_lights;
    // device copies for the path tracer (uploaded lazily)
    bool pt_dirty = true;
    bool tris_dirty = false;
    void *d_v9 = nullptr, *d_shade = nullptr, *d_objs = nullptr, *d_lights = nullptr;
    void *d_light_tris = nullptr, *d_media = nullptr, *d_env_rgb = nullptr, *d_env_marg = nullptr;
    void *d_env_cond = nullptr, *d_plights = nullptr, *d_vlights = nullptr;
    uint32_t nlights = 0;
    float env_total = 0.0f;
    // albedo textures (rg_scene_set_textures): the texels as handed in, their
    // linear device form, per-object binding and per-triangle UVs
    std::vector<rg_texture> tex;          // rgb pointers unused after the copy
    std::vector<float> tex_rgb;           // RGB as handed in, concatenated
    std::vector<float> tex_linear;        // RGBA linear, concatenated (device layout)
    std::vector<uint32_t> tex_offset;     // first texel of each texture
    std::vector<int32_t> tex_of_obj;
    std::vector<float> tex_tiling;        // 2 per object
    std::vector<float> uv6;               // 6 per triangle
    void *d_texels = nullptr, *d_tex_desc = nullptr, *d_otex = nullptr, *d_uv6 = nullptr;
    // loader diagnostics
    std::vector<std::string> warnings;
    bool has_camera = false;
};

// error reporting (thread-local text)
int set_error(int code, const std::string &msg);

// rg_bvh.cpp
void build_bvh2(const float *v9, uint32_t n, uint32_t bins, uint32_t leaf,
                std::vector<Node2> &nodes, std::vector<uint32_t> &perm);
// The first kTreeletBytes of a wide tree's nodes are its top levels in
// breadth-first order (21 4-wide or 10 8-wide nodes); the traversal kernels
// keep them in LDS.
constexpr uint32_t kTreeletBytes = 2688;
void collapse_bvh4(const std::vector<Node2> &nodes2, std::vector<Node4> &out);
void collapse_bvh8(const std::vector<Node2> &nodes2, std::vector<Node8> &out);
void make_tris(const float *v9, const std::vector<Node2> &nodes2,
               const std::vector<uint32_t> &perm, const std::vector<uint8_t> &pred_mask,
               std::vector<Tri48> &out);
uint32_t stack_need4(const std::vector<Node4> &nodes);
uint32_t stack_need8(const std::vector<Node8> &nodes);

// rg_build.hip: the same binary tree, wide tree and triangle records built
// on the device (d_nodes2, d_perm, d_nodes4, d_tris, n_nodes2, n_wide,
// stack_need and root_box of the scene); the host copies are left empty.
bool gpu_build_supported(uint32_t bins, uint32_t leaf);
// The context provides the stream and the builder's scratch workspace.
int gpu_build_bvh(Ctx &c, Scene &s, uint32_t bins, uint32_t leaf, int width);
// Frees the builder's workspace of a context.
void release_build(Ctx &c);
// Rewrites the triangle records of a GPU-built scene (occluder bits changed).
int gpu_make_tris(Scene &s, void *stream);
// Fills nodes2 / perm from the device copies of a GPU build.
int download_host_tree(Scene &s, void *stream);
// Smallest scene RG_BUILDER_AUTO builds on the GPU.
constexpr uint32_t kGpuBuildMin = 1u << 14;

// rg_trace.hip (device side)
int dev_malloc(void **p, size_t bytes);
void dev_free(void *p);
int dev_upload(void *dst, const void *src, size_t bytes, void *stream);
int dev_download(void *dst, const void *src, size_t bytes, void *stream);
int dev_sync(void *stream);
int dev_set_device(int device);
void dev_clear_error();
// work: a device counter the launch may clear and use to share the rays
// among its waves (null: even split); launches that share one must not
// overlap.
int launch_closest(const Scene &s, const rg_ray *rays, rg_hit *hits, uint64_t n, uint64_t *work, void *stream);
int launch_any(const Scene &s, const rg_ray *rays, uint8_t *occ, uint64_t n, int pred, uint64_t *work,
               void *stream);
int launch_gen_rays(const float box[6], float scale, uint32_t key, uint64_t first,
                    rg_ray *rays, uint64_t n, void *stream);

}  // namespace rg

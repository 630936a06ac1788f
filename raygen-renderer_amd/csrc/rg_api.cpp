// rg_api.cpp — the extern "C" surface declared in include/rg.h.
#include <cmath>
#include <cstring>
#include <new>
#include <string>

#include "rg_internal.h"

namespace rg {

static thread_local std::string g_last_error;

int set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

}  // namespace rg

using namespace rg;

namespace {

int need_device(const Ctx *c) {
    if (!c) return set_error(RG_E_INVALID, "null context");
    if (c->device < 0) return set_error(RG_E_NODEVICE, "host-only context has no device");
    return RG_OK;
}

void free_device(Scene *s) {
    dev_free(s->d_nodes4);
    dev_free(s->d_nodes2);
    dev_free(s->d_perm);
    dev_free(s->d_tris);
    s->d_nodes4 = s->d_nodes2 = s->d_perm = s->d_tris = nullptr;
    s->device_bytes = 0;
}

// stream helpers live in rg_trace.hip; declared here to keep rg_internal.h
// free of HIP types
}  // namespace

namespace rg {
int dev_stream_create(void **s);
void dev_stream_destroy(void *s);
void release_pt(Ctx *c);
void release_engine(Ctx *c);
void release_pt_scene(Scene *s);
void release_post(Ctx *c);
int post_scratch_init(int device);
}  // namespace rg

extern "C" {

const char *rg_last_error(void) { return g_last_error.c_str(); }

int rg_create(int device, rg_ctx **out) {
    if (!out) return set_error(RG_E_INVALID, "out is null");
    *out = nullptr;
    Ctx *c = new (std::nothrow) Ctx();
    if (!c) return set_error(RG_E_NOMEM, "context allocation failed");
    c->device = device;
    if (device >= 0) {
        int rc = dev_set_device(device);
        if (rc) { delete c; return set_error(RG_E_NODEVICE, "no HIP device " + std::to_string(device) + ": " + g_last_error); }
        rc = dev_stream_create(&c->own_stream);
        if (rc) { delete c; return rc; }
        c->stream = c->own_stream;
        rc = dev_malloc((void **)&c->trace_work, Ctx::kTraceSlots * sizeof(uint64_t));
        if (rc) { dev_stream_destroy(c->own_stream); delete c; return rc; }
        // the no-allocation entry point's scratch address, once per device;
        // a failure is reported by that entry point, not here
        (void)post_scratch_init(device);
    }
    *out = reinterpret_cast<rg_ctx *>(c);
    return RG_OK;
}

void rg_destroy(rg_ctx *ctx) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return;
    if (c->device >= 0) {
        dev_set_device(c->device);
        release_pt(c);
        release_post(c);
        release_engine(c);
        release_build(*c);
        dev_free(c->trace_work);
    }
    if (c->own_stream) dev_stream_destroy(c->own_stream);
    if (c->device >= 0) dev_clear_error();
    delete c;
}

int rg_set_stream(rg_ctx *ctx, void *hip_stream) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (int rc = need_device(c)) return rc;
    c->stream = hip_stream ? hip_stream : c->own_stream;
    return RG_OK;
}

int rg_synchronize(rg_ctx *ctx) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (int rc = need_device(c)) return rc;
    return dev_sync(c->stream);
}

void rg_cancel(rg_ctx *ctx) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (c) c->cancel.store(true);
}

int rg_set_progress(rg_ctx *ctx, rg_progress_fn fn, void *user) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c) return set_error(RG_E_INVALID, "null context");
    c->progress = fn;
    c->progress_user = user;
    return RG_OK;
}

int rg_scene_create(rg_ctx *ctx, const float *v9, uint32_t n_tris, rg_scene **out) {
    if (!ctx || !out) return set_error(RG_E_INVALID, "null argument");
    if (n_tris > 0 && !v9) return set_error(RG_E_INVALID, "v9 is null");
    if (n_tris >= 0x7FFFFFFFu) return set_error(RG_E_INVALID, "too many triangles");
    *out = nullptr;
    Scene *s = new (std::nothrow) Scene();
    if (!s) return set_error(RG_E_NOMEM, "scene allocation failed");
    s->device = reinterpret_cast<Ctx *>(ctx)->device;
    s->n_tris = n_tris;
    try {
        s->v9.assign(v9, v9 + 9 * (size_t)n_tris);
    } catch (...) {
        delete s;
        return set_error(RG_E_NOMEM, "geometry copy failed");
    }
    // bounds in registers (a loop through s->box kept every step in memory)
    float lo0 = INFINITY, lo1 = INFINITY, lo2 = INFINITY, hi0 = -INFINITY, hi1 = -INFINITY, hi2 = -INFINITY;
    const float *p = s->v9.data();
    for (size_t i = 0; i < (size_t)n_tris * 3; i++, p += 3) {
        lo0 = p[0] < lo0 ? p[0] : lo0;
        lo1 = p[1] < lo1 ? p[1] : lo1;
        lo2 = p[2] < lo2 ? p[2] : lo2;
        hi0 = p[0] > hi0 ? p[0] : hi0;
        hi1 = p[1] > hi1 ? p[1] : hi1;
        hi2 = p[2] > hi2 ? p[2] : hi2;
    }
    s->box[0] = lo0; s->box[1] = lo1; s->box[2] = lo2;
    s->box[3] = hi0; s->box[4] = hi1; s->box[5] = hi2;
    *out = reinterpret_cast<rg_scene *>(s);
    return RG_OK;
}

void rg_scene_destroy(rg_scene *scene) {
    Scene *s = reinterpret_cast<Scene *>(scene);
    if (!s) return;
    // the scene may outlive the context it was created on: it keeps its own
    // device id
    if (s->device >= 0) {
        dev_set_device(s->device);
        free_device(s);
        release_pt_scene(s);
        dev_clear_error();
    }
    delete s;
}

int rg_bvh_build(rg_ctx *ctx, rg_scene *scene, const rg_bvh_params *params) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    Scene *s = reinterpret_cast<Scene *>(scene);
    if (!c || !s) return set_error(RG_E_INVALID, "null argument");
    rg_bvh_params p = {16, 4, 4, RG_BUILDER_AUTO};
    if (params) p = *params;
    if (p.bins < 2 || p.bins > 256) return set_error(RG_E_INVALID, "bins must be in [2, 256]");
    if (p.leaf < 1 || p.leaf > 0xFFFF) return set_error(RG_E_INVALID, "leaf must be in [1, 65535]");
    if (p.width != 2 && p.width != 4 && p.width != 8) return set_error(RG_E_UNSUPPORTED, "width must be 2, 4 or 8");
    if (p.builder > RG_BUILDER_DEVICE) return set_error(RG_E_INVALID, "unknown builder");
    if (p.builder == RG_BUILDER_DEVICE) {
        if (c->device < 0) return set_error(RG_E_NODEVICE, "the device builder needs a device context");
        if (!gpu_build_supported(p.bins, p.leaf)) return set_error(RG_E_UNSUPPORTED, "the device builder takes bins <= 32");
    }
    const bool on_gpu = c->device >= 0 && s->n_tris > 0 && gpu_build_supported(p.bins, p.leaf) &&
                        (p.builder == RG_BUILDER_DEVICE || (p.builder == RG_BUILDER_AUTO && s->n_tris >= kGpuBuildMin));
    s->width = p.width == 8 ? 8 : 4;
    s->built = false;
    s->nodes2.clear();
    s->perm.clear();
    s->nodes4.clear();
    s->nodes8.clear();
    if (c->device >= 0) {
        if (int rc = dev_set_device(c->device)) return rc;
        free_device(s);
        s->device = c->device;
    }
    if (on_gpu) {
        if (int rc = gpu_build_bvh(*c, *s, p.bins, p.leaf, s->width)) return rc;
        s->host_tree = false;
        s->built = true;
        s->tris_dirty = false;
        // the traversal's wide nodes and triangle records, plus the binary
        // tree and permutation kept for exports and triangle rewrites
        s->device_bytes = s->n_wide * (s->width == 8 ? sizeof(Node8) : sizeof(Node4)) +
                          (uint64_t)s->n_tris * sizeof(Tri48) + (uint64_t)s->n_nodes2 * sizeof(Node2) +
                          (uint64_t)s->n_tris * 4;
        return RG_OK;
    }
    try {
        build_bvh2(s->v9.data(), s->n_tris, p.bins, p.leaf, s->nodes2, s->perm);
        if (p.width == 8) collapse_bvh8(s->nodes2, s->nodes8);
        else collapse_bvh4(s->nodes2, s->nodes4);
    } catch (const std::bad_alloc &) {
        return set_error(RG_E_NOMEM, "BVH build ran out of host memory");
    }
    s->host_tree = true;
    s->n_nodes2 = (uint32_t)s->nodes2.size();
    s->n_wide = p.width == 8 ? s->nodes8.size() : s->nodes4.size();
    for (int k = 0; k < 3; k++) {
        s->root_box[k] = s->nodes2.empty() ? 0.0f : s->nodes2[0].bmin[k];
        s->root_box[3 + k] = s->nodes2.empty() ? 0.0f : s->nodes2[0].bmax[k];
    }
    s->stack_need = p.width == 8 ? stack_need8(s->nodes8) : stack_need4(s->nodes4);
    s->built = true;
    if (c->device < 0) return RG_OK;

    std::vector<Tri48> tris;
    make_tris(s->v9.data(), s->nodes2, s->perm, s->pred_mask, tris);
    const void *hn = s->width == 8 ? (const void *)s->nodes8.data() : (const void *)s->nodes4.data();
    const size_t b4 = s->width == 8 ? s->nodes8.size() * sizeof(Node8) : s->nodes4.size() * sizeof(Node4);
    const size_t bt = tris.size() * sizeof(Tri48);
    if (int rc = dev_malloc(&s->d_nodes4, b4)) return rc;
    if (int rc = dev_malloc(&s->d_tris, bt)) { free_device(s); return rc; }
    if (int rc = dev_upload(s->d_nodes4, hn, b4, c->stream)) { free_device(s); return rc; }
    if (int rc = dev_upload(s->d_tris, tris.data(), bt, c->stream)) { free_device(s); return rc; }
    s->device_bytes = b4 + bt;
    s->tris_dirty = false;
    return RG_OK;
}

int rg_bvh_export2(const rg_scene *scene, rg_bvh2_node *nodes, uint64_t cap, uint32_t *perm,
                   uint64_t *n_nodes) {
    Scene *s = const_cast<Scene *>(reinterpret_cast<const Scene *>(scene));
    if (!s || !n_nodes) return set_error(RG_E_INVALID, "null argument");
    if (!s->built) return set_error(RG_E_INVALID, "BVH not built");
    *n_nodes = s->n_nodes2;
    if (!nodes && !perm) return RG_OK;
    if (nodes && cap < s->n_nodes2) return set_error(RG_E_INVALID, "node capacity too small");
    if (!s->host_tree) {
        if (int rc = dev_set_device(s->device)) return rc;
        if (int rc = download_host_tree(*s, nullptr)) return rc;
    }
    if (nodes) std::memcpy(nodes, s->nodes2.data(), s->nodes2.size() * sizeof(Node2));
    if (perm) std::memcpy(perm, s->perm.data(), s->perm.size() * sizeof(uint32_t));
    return RG_OK;
}

int rg_scene_device_bytes(const rg_scene *scene, uint64_t *bytes) {
    const Scene *s = reinterpret_cast<const Scene *>(scene);
    if (!s || !bytes) return set_error(RG_E_INVALID, "null argument");
    *bytes = s->device_bytes;
    return RG_OK;
}

int rg_scene_bounds(const rg_scene *scene, float box[6]) {
    const Scene *s = reinterpret_cast<const Scene *>(scene);
    if (!s || !box) return set_error(RG_E_INVALID, "null argument");
    std::memcpy(box, s->box, sizeof s->box);
    return RG_OK;
}

int rg_trace_closest(rg_ctx *ctx, const rg_scene *scene, const rg_ray *d_rays, rg_hit *d_hits,
                     uint64_t n) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    const Scene *s = reinterpret_cast<const Scene *>(scene);
    if (int rc = need_device(c)) return rc;
    if (!s || (n && (!d_rays || !d_hits))) return set_error(RG_E_INVALID, "null argument");
    if (!s->built) return set_error(RG_E_INVALID, "BVH not built");
    uint64_t *work = c->trace_work + c->trace_slot++ % Ctx::kTraceSlots;
    return launch_closest(*s, d_rays, d_hits, n, work, c->stream);
}

int rg_trace_any(rg_ctx *ctx, const rg_scene *scene, const rg_ray *d_rays, uint8_t *d_occluded,
                 uint64_t n, int pred) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    const Scene *s = reinterpret_cast<const Scene *>(scene);
    if (int rc = need_device(c)) return rc;
    if (!s || (n && (!d_rays || !d_occluded))) return set_error(RG_E_INVALID, "null argument");
    if (!s->built) return set_error(RG_E_INVALID, "BVH not built");
    if (pred < RG_PRED_ALL || pred > RG_PRED_OPAQUE) return set_error(RG_E_INVALID, "unknown predicate");
    uint64_t *work = c->trace_work + c->trace_slot++ % Ctx::kTraceSlots;
    return launch_any(*s, d_rays, d_occluded, n, pred, work, c->stream);
}

int rg_trace_closest_host(rg_ctx *ctx, const rg_scene *scene, const rg_ray *rays, rg_hit *hits,
                          uint64_t n) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (int rc = need_device(c)) return rc;
    if (n == 0) return RG_OK;
    if (!rays || !hits) return set_error(RG_E_INVALID, "null argument");
    void *dr = nullptr, *dh = nullptr;
    int rc = dev_malloc(&dr, n * sizeof(rg_ray));
    if (!rc) rc = dev_malloc(&dh, n * sizeof(rg_hit));
    if (!rc) rc = dev_upload(dr, rays, n * sizeof(rg_ray), c->stream);
    if (!rc) rc = rg_trace_closest(ctx, scene, (const rg_ray *)dr, (rg_hit *)dh, n);
    if (!rc) rc = dev_download(hits, dh, n * sizeof(rg_hit), c->stream);
    dev_free(dr);
    dev_free(dh);
    return rc;
}

int rg_trace_any_host(rg_ctx *ctx, const rg_scene *scene, const rg_ray *rays, uint8_t *occluded,
                      uint64_t n, int pred) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (int rc = need_device(c)) return rc;
    if (n == 0) return RG_OK;
    if (!rays || !occluded) return set_error(RG_E_INVALID, "null argument");
    void *dr = nullptr, *dh = nullptr;
    int rc = dev_malloc(&dr, n * sizeof(rg_ray));
    if (!rc) rc = dev_malloc(&dh, n);
    if (!rc) rc = dev_upload(dr, rays, n * sizeof(rg_ray), c->stream);
    if (!rc) rc = rg_trace_any(ctx, scene, (const rg_ray *)dr, (uint8_t *)dh, n, pred);
    if (!rc) rc = dev_download(occluded, dh, n, c->stream);
    dev_free(dr);
    dev_free(dh);
    return rc;
}

int rg_gen_rays_sphere(rg_ctx *ctx, const float box[6], float scale, uint32_t key, uint64_t first,
                       rg_ray *d_rays, uint64_t n) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (int rc = need_device(c)) return rc;
    if (!box || (n && !d_rays)) return set_error(RG_E_INVALID, "null argument");
    return launch_gen_rays(box, scale, key, first, d_rays, n, c->stream);
}

}  // extern "C"

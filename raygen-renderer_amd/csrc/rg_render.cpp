// rg_render.cpp — host driver of the wavefront path tracer: scene shading
// inputs, device upload, the sample loop (uniform or per-tile adaptive,
// RayRenderer::render / renderAdaptive, rayrenderer.cpp:478-596, 1162-1252)
// and the HDR resolve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "rg_internal.h"
#include "rg_pt.h"

namespace rg {

int launch_tile_noise(const float *d_sum, const float *d_sumsq, const uint32_t *d_nspp, uint32_t width,
                      uint32_t height, uint32_t target, float threshold, uint8_t *d_mask, uint32_t *d_nactive,
                      void *stream);
int launch_tiles_add(uint32_t *d_nspp, const uint8_t *d_mask, uint32_t ntiles, uint32_t add, void *stream);
int launch_denoise(const rg_settings &s, uint32_t width, uint32_t height, const PTScene &sc, const Scene &scn,
                   float *d_hdr, void *d_aov, void *d_a, void *d_b, void *stream);
int launch_aov(const Scene &scn, const PTScene &sc, void *d_aov, void *stream);

namespace {

int hip_ok(hipError_t e, const char *what) {
    if (e == hipSuccess) return RG_OK;
    return set_error(RG_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void free_pt_scene(Scene *s) {
    for (void **p : {&s->d_v9, &s->d_shade, &s->d_objs, &s->d_lights, &s->d_light_tris, &s->d_media,
                     &s->d_env_rgb, &s->d_env_marg, &s->d_env_cond, &s->d_texels, &s->d_tex_desc, &s->d_otex,
                     &s->d_uv6, &s->d_plights, &s->d_vlights}) {
        dev_free(*p);
        *p = nullptr;
    }
}

template <class T>
int upload_vec(void **dst, const std::vector<T> &v, void *stream) {
    dev_free(*dst);
    *dst = nullptr;
    if (v.empty()) return RG_OK;
    if (int rc = dev_malloc(dst, v.size() * sizeof(T))) return rc;
    return dev_upload(*dst, v.data(), v.size() * sizeof(T), stream);
}

// Per-triangle records, lights, media and envmap CDF; the arithmetic is the
// reference's (precalc raycommon.cpp:128-149, prepare medium.cpp:18-57,
// buildEnvmapCDF scene.cpp:400-463).
int upload_pt_scene(Scene *s, void *stream) {
    const uint32_t n = s->n_tris;
    const size_t n_obj = s->objects.size();
    std::vector<uint32_t> cnt(n_obj, 0);
    for (uint32_t t = 0; t < n; t++) cnt[s->obj_of_tri[t]]++;
    // per object: its triangle count when it emits (sampledArea = count x
    // area), else 0; emissive objects' triangles in object then index order
    std::vector<float> scale(n_obj, 0.0f);
    std::vector<uint32_t> first(n_obj, 0);
    std::vector<DevLight> lights;
    uint32_t n_light_tris = 0;
    for (uint32_t o = 0; o < n_obj; o++) {
        if (!(s->objects[o].emission > 0.0f) || cnt[o] == 0) continue;
        scale[o] = (float)cnt[o];
        first[o] = n_light_tris;
        lights.push_back(DevLight{o, n_light_tris, cnt[o], 0});
        n_light_tris += cnt[o];
    }
    std::vector<uint32_t> light_tris(n_light_tris);
    if (n_light_tris) {
        std::vector<uint32_t> at(first);
        for (uint32_t t = 0; t < n; t++) {
            const uint32_t o = s->obj_of_tri[t];
            if (scale[o] > 0.0f) light_tris[at[o]++] = t;
        }
    }
    std::vector<DevMedium> media(s->media.size());
    for (size_t i = 0; i < s->media.size(); i++) {
        const rg_medium &m = s->media[i];
        DevMedium &q = media[i];
        const float den = fmaxf(0.0f, m.density);
        for (int k = 0; k < 3; k++) {
            q.sigma_s[k] = m.sigma_s[k] * den;
            q.sigma_e[k] = m.sigma_e[k] * den;
            q.sigma_t[k] = m.sigma_a[k] * den + q.sigma_s[k];
        }
        q.hero = (q.sigma_t[0] + q.sigma_t[1] + q.sigma_t[2]) * (1.0f / 3.0f);
        q.g = m.g;
        q.src = m;
        q.het = m.density_field == 1 ? 1 : 0;
        const float peak = q.het ? fmaxf(0.0f, m.noise_amplitude * 1.0f + m.noise_bias) : 1.0f;
        q.t_max_hero = q.het ? q.hero * peak : q.hero;
        const float al2 = m.cone_axis[0] * m.cone_axis[0] + m.cone_axis[1] * m.cone_axis[1] +
                          m.cone_axis[2] * m.cone_axis[2];
        if (al2 > 1e-12f) {
            const float inv = 1.0f / sqrtf(al2);
            for (int k = 0; k < 3; k++) q.cone_axis[k] = m.cone_axis[k] * inv;
        } else {
            q.cone_axis[0] = q.cone_axis[1] = 0.0f;
            q.cone_axis[2] = -1.0f;
        }
        q.haze = (m.heat_haze && m.ior_amplitude > 0.0f && m.ior_march_steps > 0) ? 1 : 0;
        // isActive (medium.h:214-223)
        const bool emits = q.sigma_e[0] != 0.0f || q.sigma_e[1] != 0.0f || q.sigma_e[2] != 0.0f;
        q.active = (q.hero > 0.0f || (m.emission_mode == 1 && m.cone_intensity > 0.0f) ||
                    (m.emission_mode == 2 && m.path_intensity > 0.0f && m.n_path >= 2) ||
                    (m.heat_haze && m.ior_amplitude > 0.0f) || emits) ? 1 : 0;
    }
    std::vector<float> marg, cond;
    s->env_total = 0.0f;
    if (s->env_w > 0 && s->env_h > 0) {
        const int W = (int)s->env_w, H = (int)s->env_h;
        cond.assign((size_t)H * (W + 1), 0.0f);
        marg.assign((size_t)H + 1, 0.0f);
        std::vector<float> roww(H, 0.0f);
        for (int y = 0; y < H; y++) {
            const float v = ((float)y + 0.5f) / (float)H;
            const float st = sinf(3.14159265358979323846f * v);
            float rs = 0.0f;
            for (int x = 0; x < W; x++) {
                const float *c = s->env_rgb.data() + 3 * ((size_t)y * W + x);
                const float l = fmaxf(0.0f, 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2]);
                rs += l * st;
                cond[(size_t)y * (W + 1) + x + 1] = rs;
            }
            roww[y] = rs;
        }
        for (int y = 0; y < H; y++)
            if (roww[y] > 0.0f) {
                const float inv = 1.0f / roww[y];
                for (int x = 1; x <= W; x++) cond[(size_t)y * (W + 1) + x] *= inv;
            }
        for (int y = 0; y < H; y++) marg[y + 1] = marg[y] + roww[y];
        s->env_total = marg[H];
        if (s->env_total > 0.0f) {
            const float inv = 1.0f / s->env_total;
            for (int y = 0; y <= H; y++) marg[y] *= inv;
        }
    }
    s->nlights = (uint32_t)lights.size();
    // point lights: the spot cone's cosine, cosf(RADIAN_TO_DEGREE(spotRange
    // / 2)) as the reference computes it (rayrenderer.cpp:2230), with
    // RADIAN_TO_DEGREE(x) = x * 180 / pi
    std::vector<DevPointLight> plights(s->point_lights.size());
    for (size_t i = 0; i < plights.size(); i++) {
        const rg_point_light &l = s->point_lights[i];
        DevPointLight &q = plights[i];
        std::memcpy(q.pos, l.pos, sizeof q.pos);
        std::memcpy(q.normal, l.normal, sizeof q.normal);
        std::memcpy(q.color, l.color, sizeof q.color);
        q.emission = l.emission;
        q.spot = l.spot_range > 0.0f ? 1 : 0;
        q.spot_cos = q.spot ? cosf(l.spot_range * 0.5f * (float)(180.0 / M_PI)) : 0.0f;
    }
    // texture tables (kept empty when no object has a texture)
    std::vector<int32_t> otex;
    std::vector<uint32_t> tex_desc;
    if (!s->tex.empty()) {
        otex.resize(4 * s->objects.size());
        for (size_t o = 0; o < s->objects.size(); o++) {
            const int32_t t = o < s->tex_of_obj.size() ? s->tex_of_obj[o] : -1;
            float tu = 1.0f, tv = 1.0f;
            if (2 * o + 1 < s->tex_tiling.size()) {
                tu = s->tex_tiling[2 * o];
                tv = s->tex_tiling[2 * o + 1];
            }
            otex[4 * o] = t;
            std::memcpy(&otex[4 * o + 1], &tu, 4);
            std::memcpy(&otex[4 * o + 2], &tv, 4);
            otex[4 * o + 3] = 0;
        }
        for (size_t k = 0; k < s->tex.size(); k++)
            tex_desc.insert(tex_desc.end(), {s->tex[k].w, s->tex[k].h, s->tex_offset[k], 0u});
    }
    // the triangles never change after rg_scene_create: a copy already on the
    // device (the GPU builder's, or an earlier upload) is current
    int rc = s->d_v9 ? RG_OK : upload_vec(&s->d_v9, s->v9, stream);
    // the shading records are made on the device from the triangles, the
    // vertex normals and the object ids
    void *d_n9 = nullptr, *d_obj = nullptr, *d_scale = nullptr;
    if (!rc) rc = upload_vec(&d_n9, s->n9, stream);
    if (!rc) rc = upload_vec(&d_obj, s->obj_of_tri, stream);
    if (!rc) rc = upload_vec(&d_scale, scale, stream);
    dev_free(s->d_shade);
    s->d_shade = nullptr;
    if (!rc && n) rc = dev_malloc(&s->d_shade, (size_t)n * sizeof(DevTriShade));
    if (!rc)
        rc = launch_tri_shade((const float *)s->d_v9, (const float *)d_n9, (const uint32_t *)d_obj,
                              (const float *)d_scale, n, (DevTriShade *)s->d_shade, stream);
    if (!rc) rc = dev_sync(stream);
    dev_free(d_n9);
    dev_free(d_obj);
    dev_free(d_scale);
    if (!rc) rc = upload_vec(&s->d_otex, otex, stream);
    if (!rc) rc = upload_vec(&s->d_tex_desc, tex_desc, stream);
    if (!rc) rc = upload_vec(&s->d_texels, s->tex.empty() ? std::vector<float>() : s->tex_linear, stream);
    if (!rc) rc = upload_vec(&s->d_uv6, s->tex.empty() ? std::vector<float>() : s->uv6, stream);
    if (!rc) rc = upload_vec(&s->d_objs, s->objects, stream);
    if (!rc) rc = upload_vec(&s->d_lights, lights, stream);
    if (!rc) rc = upload_vec(&s->d_light_tris, light_tris, stream);
    if (!rc) rc = upload_vec(&s->d_plights, plights, stream);
    if (!rc) rc = upload_vec(&s->d_vlights, s->volume_lights, stream);
    if (!rc) rc = upload_vec(&s->d_media, media, stream);
    if (!rc) rc = upload_vec(&s->d_env_rgb, s->env_rgb, stream);
    if (!rc) rc = upload_vec(&s->d_env_marg, marg, stream);
    if (!rc) rc = upload_vec(&s->d_env_cond, cond, stream);
    if (rc) return rc;
    s->pt_dirty = false;
    return RG_OK;
}

int reupload_tris(Scene *s, void *stream) {
    if (!s->host_tree) {
        s->tris_dirty = false;
        return gpu_make_tris(*s, stream);
    }
    std::vector<Tri48> tris;
    make_tris(s->v9.data(), s->nodes2, s->perm, s->pred_mask, tris);
    if (tris.empty()) return RG_OK;
    s->tris_dirty = false;
    return dev_upload(s->d_tris, tris.data(), tris.size() * sizeof(Tri48), stream);
}

void free_buffers(PTBuffers &b) {
    for (void **p : {&b.ray, &b.hit, &b.beta, &b.tp, &b.L, &b.meta, &b.sray, &b.ssrc, &b.socc, &b.scontrib,
                     &b.nshadow, &b.q0,
                     &b.q1, &b.count, &b.sum, &b.sumsq, &b.hdr, &b.tile_mask, &b.tile_nspp,
                     &b.tile_noise, &b.n_active, &b.aov, &b.dn_a, &b.dn_b, &b.stats}) {
        dev_free(*p);
        *p = nullptr;
    }
    b.cap = b.npix_cap = b.ntile_cap = 0;
}

int ensure_buffers(PTBuffers &b, uint64_t cap, uint64_t npix, uint64_t ntiles) {
    if (cap <= b.cap && npix <= b.npix_cap && ntiles <= b.ntile_cap) return RG_OK;
    cap = std::max(cap, b.cap);
    npix = std::max(npix, b.npix_cap);
    ntiles = std::max(ntiles, b.ntile_cap);
    free_buffers(b);
    struct A {
        void **p;
        uint64_t bytes;
    } allocs[] = {
        {&b.ray, cap * 32}, {&b.hit, cap * 16}, {&b.beta, cap * 16}, {&b.tp, cap * 16}, {&b.L, cap * 16},
        {&b.meta, cap * 16}, {&b.sray, cap * 32 * kNeeSlots}, {&b.ssrc, cap * 4 * kNeeSlots},
        {&b.socc, cap * kNeeSlots}, {&b.scontrib, cap * 16 * kNeeSlots}, {&b.nshadow, cap * 4},
        {&b.q0, cap * 4}, {&b.q1, cap * 4}, {&b.count, 64}, {&b.sum, npix * 12},
        {&b.sumsq, npix * 12}, {&b.hdr, npix * 16}, {&b.tile_mask, ntiles}, {&b.tile_nspp, ntiles * 4},
        {&b.tile_noise, ntiles * 4}, {&b.n_active, 64}, {&b.aov, npix * 36}, {&b.dn_a, npix * 16},
        {&b.dn_b, npix * 16}, {&b.stats, 64},
    };
    for (auto &a : allocs)
        if (int rc = dev_malloc(a.p, a.bytes)) {
            free_buffers(b);
            return rc;
        }
    b.cap = cap;
    b.npix_cap = npix;
    b.ntile_cap = ntiles;
    return RG_OK;
}

}  // namespace

int path_events(const Scene &s, uint32_t max_depth) {
    bool haze = false;
    for (const rg_medium &m : s.media) haze = haze || (m.heat_haze && m.ior_amplitude > 0.0f && m.ior_march_steps > 0);
    return (int)(max_depth + 2) * (haze ? 2 : 1);
}

bool scene_uses_ext(const Scene &s) {
    if (!s.volume_lights.empty()) return true;
    for (const rg_medium &m : s.media)
        if (m.density_field != 0 || m.emission_mode != 0 || m.heat_haze) return true;
    return false;
}

int check_shader(const rg_settings &st) {
    if (st.shader != 0 && st.shader != 1 && st.shader != 2 && st.shader != 3 && st.shader != 5)
        return set_error(RG_E_INVALID, "shader must be 0, 1, 2, 3 or 5 (the reference's providers)");
    for (float v : st.world_color)
        if (!std::isfinite(v)) return set_error(RG_E_INVALID, "world_color not finite");
    return RG_OK;
}

PTScene make_ptscene(const Scene &scn, const rg_settings &st) {
    const Scene *s = &scn;
    const uint32_t W = st.width, H = st.height;
    PTScene sc;
    std::memset(&sc, 0, sizeof sc);
    sc.v9 = (const float *)s->d_v9;
    sc.shade = (const DevTriShade *)s->d_shade;
    sc.objs = (const DevObject *)s->d_objs;
    sc.lights = (const DevLight *)s->d_lights;
    sc.light_tris = (const uint32_t *)s->d_light_tris;
    sc.nlights = s->nlights;
    sc.plights = (const DevPointLight *)s->d_plights;
    sc.nplights = (uint32_t)s->point_lights.size();
    sc.vlights = (const DevVolumeLight *)s->d_vlights;
    sc.nvlights = (uint32_t)s->volume_lights.size();
    sc.events = path_events(*s, st.max_depth);
    sc.ext = scene_uses_ext(*s) ? 1 : 0;
    sc.media = (const DevMedium *)s->d_media;
    sc.global_medium = s->global_medium;
    sc.env_rgb = (const float *)s->d_env_rgb;
    sc.env_marg = (const float *)s->d_env_marg;
    sc.env_cond = (const float *)s->d_env_cond;
    sc.env_w = s->env_w;
    sc.env_h = s->env_h;
    sc.env_total = s->env_total;
    sc.env_intensity = s->env_intensity;
    {
        const float yaw = s->env_rotation_deg * (float)(M_PI / 180.0);
        sc.env_cos = cosf(yaw);
        sc.env_sin = sinf(yaw);
    }
    sc.has_env = (s->env_w > 0 && s->env_h > 0) ? 1 : 0;
    sc.has_tex = s->tex.empty() ? 0 : 1;
    sc.otex = (const int4 *)s->d_otex;
    sc.tex_desc = (const uint4 *)s->d_tex_desc;
    sc.texels = (const float4 *)s->d_texels;
    sc.uv6 = (const float *)s->d_uv6;
    const rg_camera &cam = s->camera;
    sc.half_w = (float)W * 0.5f;
    sc.half_h = (float)H * 0.5f;
    {
        const float fov = fminf(fmaxf(cam.fov_deg, 1.0f), 179.0f);
        const float th = tanf(fov * (float)(M_PI / 180.0) * 0.5f);
        sc.view_scale = (2.0f * th) / (float)H;
    }
    sc.dof = cam.depth_of_field;
    sc.aperture = cam.aperture > 0.0f ? 1.0f / cam.aperture : 0.0f;
    sc.blades = cam.aperture_blades;
    sc.ap_rot = cam.aperture_rotation_deg * (float)(M_PI / 180.0);
    sc.max_depth = (int32_t)st.max_depth;
    sc.antialias = st.antialias ? 1 : 0;
    sc.width = W;
    sc.height = H;
    sc.seed = st.seed;
    sc.job = st.job;
    sc.color_sampling = st.color_sampling ? 1 : 0;
    sc.back[0] = st.back_color[0];
    sc.back[1] = st.back_color[1];
    sc.back[2] = st.back_color[2];
    sc.back[3] = st.back_color[3];
    sc.shader = st.shader;
    sc.ao_samples = std::max<uint32_t>(1, st.spp);
    sc.world[0] = st.world_color[0];
    sc.world[1] = st.world_color[1];
    sc.world[2] = st.world_color[2];
    return sc;
}

int prepare_scene(Scene *s, void *stream) {
    if (!s->built) return set_error(RG_E_INVALID, "BVH not built");
    if (!s->has_shading) return set_error(RG_E_INVALID, "rg_scene_set_shading was not called");
    for (const rg_object &o : s->objects)
        if (o.interior_medium >= (int32_t)s->media.size())
            return set_error(RG_E_INVALID, "interior medium index out of range");
    if (s->tris_dirty)
        if (int rc = reupload_tris(s, stream)) return rc;
    if (s->pt_dirty)
        if (int rc = upload_pt_scene(s, stream)) return rc;
    return RG_OK;
}

void release_pt(Ctx *c) {
    free_buffers(c->pt);
    for (void *&e : c->progress_ev) {
        if (e) (void)hipEventDestroy((hipEvent_t)e);
        e = nullptr;
    }
}
void release_pt_scene(Scene *s) { free_pt_scene(s); }

}  // namespace rg

using namespace rg;

extern "C" {

void rg_object_default(rg_object *o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->color[0] = o->color[1] = o->color[2] = 0.8f;
    o->roughness = 0.5f;
    o->refraction_ratio = 1.45f;
    o->interior_medium = -1;
    o->visible = 1;
}

void rg_medium_default(rg_medium *m) {
    if (!m) return;
    std::memset(m, 0, sizeof *m);
    // HomogeneousMedium's member initialisers (medium.h:66-205)
    m->density = 1.0f;
    m->cone_axis[2] = -1.0f;
    m->cone_length = 1.0f;
    m->cone_radius = 0.3f;
    const float inner[3] = {0.6f, 0.7f, 1.0f}, outer[3] = {1.0f, 0.4f, 0.05f}, pouter[3] = {0.3f, 0.4f, 0.6f};
    std::memcpy(m->cone_inner, inner, sizeof inner);
    std::memcpy(m->cone_outer, outer, sizeof outer);
    m->cone_intensity = 50.0f;
    m->cone_peak_axial = 0.15f;
    m->cone_peak_sharpness = 5.0f;
    m->cone_samples = 4;
    std::memcpy(m->path_inner, inner, sizeof inner);
    std::memcpy(m->path_outer, pouter, sizeof pouter);
    m->path_intensity = 1.0f;
    m->path_falloff_power = 2.0f;
    m->path_samples = 6;
    for (rg_path_point &q : m->path) q.radius = 0.1f;
    m->noise_frequency = 1.0f;
    m->noise_octaves = 4;
    m->noise_gain = 0.5f;
    m->noise_lacunarity = 2.0f;
    m->noise_amplitude = 1.0f;
    m->ior_amplitude = 0.005f;
    m->ior_frequency = 4.0f;
    m->ior_octaves = 3;
    m->ior_gain = 0.5f;
    m->ior_lacunarity = 2.0f;
    m->ior_march_steps = 16;
}

void rg_settings_default(rg_settings *s) {
    if (!s) return;
    std::memset(s, 0, sizeof *s);
    s->width = 800;
    s->height = 600;
    s->spp = 100;
    s->max_depth = 32;
    s->antialias = 1;
    s->firefly_clamp = 10.0f;
    s->seed = 0xC0DE0000u;
    s->adaptive_base = 4;
    s->adaptive_threshold = 0.02f;
    s->denoise_levels = 5;
    s->sigma_color = 0.4f;
    s->sigma_normal = 128.0f;
    s->sigma_depth = 0.1f;
    s->denoise_intensity = 1.0f;
    s->color_sampling = 1;
    s->back_color[0] = 1.0f * 0.2f;
    s->back_color[1] = 0.95f * 0.2f;
    s->back_color[2] = 0.9f * 0.2f;
    s->back_color[3] = 0.0f;
    s->shader = 5;
    s->world_color[0] = 1.0f * 0.1f;
    s->world_color[1] = 0.95f * 0.1f;
    s->world_color[2] = 0.9f * 0.1f;
}

int rg_scene_set_shading(rg_scene *scene, const float *n9, const uint32_t *obj_of_tri, const rg_object *objs,
                         uint32_t n_obj) {
    Scene *s = reinterpret_cast<Scene *>(scene);
    if (!s || (s->n_tris && (!n9 || !obj_of_tri)) || (n_obj && !objs))
        return set_error(RG_E_INVALID, "null argument");
    for (uint32_t t = 0; t < s->n_tris; t++)
        if (obj_of_tri[t] >= n_obj) return set_error(RG_E_INVALID, "object index out of range");
    for (uint32_t o = 0; o < n_obj; o++)
        if (objs[o].interior_medium >= 0 && (size_t)objs[o].interior_medium >= s->media.size() &&
            objs[o].interior_medium >= 1 << 20)
            return set_error(RG_E_INVALID, "interior medium index out of range");
    s->n9.assign(n9, n9 + 9 * (size_t)s->n_tris);
    s->obj_of_tri.assign(obj_of_tri, obj_of_tri + s->n_tris);
    s->objects.assign(objs, objs + n_obj);
    s->pred_mask.resize(s->n_tris);
    for (uint32_t t = 0; t < s->n_tris; t++) {
        const rg_object &m = objs[obj_of_tri[t]];
        uint8_t mask = 1u << RG_PRED_ALL;
        if (m.transparency < 0.01f && m.emission <= 0.0f) mask |= 1u << RG_PRED_AREA;
        if (!(m.emission > 0.0f) && (m.transparency < 0.01f || m.refraction > 0.1f)) mask |= 1u << RG_PRED_ENV;
        if (!(m.emission > 0.0f) && m.transparency < 0.01f && m.refraction <= 0.1f) mask |= 1u << RG_PRED_VOLUME;
        if (m.transparency < 0.01f) mask |= 1u << RG_PRED_OPAQUE;
        s->pred_mask[t] = mask;
    }
    s->has_shading = true;
    s->pt_dirty = true;
    s->tris_dirty = s->built;
    return RG_OK;
}

// sRGB -> linear companding of Texture::sample (texture.cpp:203-206)
static float srgb_to_linear(float c) {
    if (c <= 0.04045f) return c * (1.0f / 12.92f);
    return powf((c + 0.055f) * (1.0f / 1.055f), 2.4f);
}

int rg_scene_set_textures(rg_scene *scene, const rg_texture *tex, uint32_t n_tex, const int32_t *tex_of_obj,
                          const float *tiling_of_obj, const float *uv6) {
    Scene *s = reinterpret_cast<Scene *>(scene);
    if (!s) return set_error(RG_E_INVALID, "null scene");
    if (n_tex == 0) {
        s->tex.clear();
        s->tex_rgb.clear();
        s->tex_linear.clear();
        s->tex_offset.clear();
        s->tex_of_obj.clear();
        s->tex_tiling.clear();
        s->uv6.clear();
        s->pt_dirty = true;
        return RG_OK;
    }
    if (!s->has_shading) return set_error(RG_E_INVALID, "rg_scene_set_shading must come first");
    const size_t n_obj = s->objects.size();
    if (!tex || !tex_of_obj || !tiling_of_obj || (s->n_tris && !uv6)) return set_error(RG_E_INVALID, "null argument");
    uint64_t total = 0;
    for (uint32_t k = 0; k < n_tex; k++) {
        if (tex[k].w == 0 || tex[k].h == 0 || tex[k].w > 65536 || tex[k].h > 65536 || !tex[k].rgb)
            return set_error(RG_E_INVALID, "texture " + std::to_string(k) + " is empty or too large");
        total += (uint64_t)tex[k].w * tex[k].h;
    }
    if (total > (1ull << 30)) return set_error(RG_E_INVALID, "textures exceed 2^30 texels");
    for (size_t o = 0; o < n_obj; o++)
        if (tex_of_obj[o] < -1 || tex_of_obj[o] >= (int32_t)n_tex)
            return set_error(RG_E_INVALID, "texture index out of range");
    s->tex.assign(tex, tex + n_tex);
    s->tex_rgb.clear();
    s->tex_linear.clear();
    s->tex_offset.clear();
    s->tex_rgb.reserve(3 * total);
    s->tex_linear.reserve(4 * total);
    for (uint32_t k = 0; k < n_tex; k++) {
        const rg_texture &t = tex[k];
        s->tex[k].rgb = nullptr;
        s->tex_offset.push_back((uint32_t)(s->tex_linear.size() / 4));
        const size_t m = (size_t)t.w * t.h;
        s->tex_rgb.insert(s->tex_rgb.end(), t.rgb, t.rgb + 3 * m);
        for (size_t i = 0; i < m; i++) {
            for (int c = 0; c < 3; c++) {
                const float v = t.rgb[3 * i + c];
                s->tex_linear.push_back(t.srgb ? srgb_to_linear(v) : v);
            }
            s->tex_linear.push_back(1.0f);
        }
    }
    s->tex_of_obj.assign(tex_of_obj, tex_of_obj + n_obj);
    s->tex_tiling.assign(tiling_of_obj, tiling_of_obj + 2 * n_obj);
    s->uv6.assign(uv6, uv6 + 6 * (size_t)s->n_tris);
    s->pt_dirty = true;
    return RG_OK;
}

int rg_scene_export_textures(const rg_scene *scene, uint32_t *n_tex, rg_texture *tex, float *texels,
                             int32_t *tex_of_obj, float *tiling_of_obj, float *uv6) {
    const Scene *s = reinterpret_cast<const Scene *>(scene);
    if (!s || !n_tex) return set_error(RG_E_INVALID, "null argument");
    *n_tex = (uint32_t)s->tex.size();
    if (tex)
        for (size_t k = 0; k < s->tex.size(); k++) tex[k] = s->tex[k];
    if (texels && !s->tex_rgb.empty()) std::memcpy(texels, s->tex_rgb.data(), s->tex_rgb.size() * 4);
    const size_t n_obj = s->objects.size();
    if (tex_of_obj)
        for (size_t o = 0; o < n_obj; o++) tex_of_obj[o] = o < s->tex_of_obj.size() ? s->tex_of_obj[o] : -1;
    if (tiling_of_obj)
        for (size_t o = 0; o < 2 * n_obj; o++) tiling_of_obj[o] = o < s->tex_tiling.size() ? s->tex_tiling[o] : 1.0f;
    if (uv6) {
        if (s->uv6.size() == 6 * (size_t)s->n_tris) std::memcpy(uv6, s->uv6.data(), s->uv6.size() * 4);
        else std::memset(uv6, 0, 6 * (size_t)s->n_tris * 4);
    }
    return RG_OK;
}

int rg_scene_set_media(rg_scene *scene, const rg_medium *media, uint32_t n, int32_t global_medium) {
    Scene *s = reinterpret_cast<Scene *>(scene);
    if (!s || (n && !media)) return set_error(RG_E_INVALID, "null argument");
    if (global_medium >= (int32_t)n) return set_error(RG_E_INVALID, "global medium index out of range");
    // the path state keeps a path's medium index in 16 bits
    if (n > 32767) return set_error(RG_E_INVALID, "at most 32767 media");
    for (uint32_t i = 0; i < n; i++) {
        if (media[i].n_path > RG_MAX_PATH_POINTS)
            return set_error(RG_E_INVALID, "medium " + std::to_string(i) + ": more than 16 path points");
        if (media[i].emission_mode < 0 || media[i].emission_mode > 2 || media[i].density_field < 0 ||
            media[i].density_field > 1)
            return set_error(RG_E_INVALID, "medium " + std::to_string(i) + ": unknown emission or density mode");
    }
    if (global_medium >= 0 && media[global_medium].heat_haze && media[global_medium].ior_amplitude > 0.0f &&
        media[global_medium].ior_march_steps > 0)
        return set_error(RG_E_UNSUPPORTED, "a heat-haze global medium (tracePath recurses without end there)");
    s->media.assign(media, media + n);
    s->global_medium = global_medium < 0 ? -1 : global_medium;
    s->pt_dirty = true;
    return RG_OK;
}

int rg_scene_set_env(rg_scene *scene, const rg_env *env) {
    Scene *s = reinterpret_cast<Scene *>(scene);
    if (!s) return set_error(RG_E_INVALID, "null scene");
    if (!env || env->w == 0 || env->h == 0) {
        s->env_w = s->env_h = 0;
        s->env_rgb.clear();
    } else {
        if (!env->rgb) return set_error(RG_E_INVALID, "env rgb is null");
        s->env_w = env->w;
        s->env_h = env->h;
        s->env_rgb.assign(env->rgb, env->rgb + 3 * (size_t)env->w * env->h);
        s->env_intensity = env->intensity;
        s->env_rotation_deg = env->rotation_deg;
    }
    s->pt_dirty = true;
    return RG_OK;
}

int rg_scene_set_volume_lights(rg_scene *scene, const rg_volume_light *v, uint32_t n) {
    Scene *s = reinterpret_cast<Scene *>(scene);
    if (!s) return set_error(RG_E_INVALID, "null scene");
    if (n && !v) return set_error(RG_E_INVALID, "volume lights is null");
    for (uint32_t i = 0; i < n; i++)
        if (v[i].medium >= s->media.size())
            return set_error(RG_E_INVALID, "volume light " + std::to_string(i) + ": medium index out of range");
    s->volume_lights.assign(v, v + n);
    s->pt_dirty = true;
    return RG_OK;
}

int rg_scene_export_volume_lights(const rg_scene *scene, rg_volume_light *v, uint32_t *n) {
    const Scene *s = reinterpret_cast<const Scene *>(scene);
    if (!s || !n) return set_error(RG_E_INVALID, "null argument");
    *n = (uint32_t)s->volume_lights.size();
    if (v && !s->volume_lights.empty())
        std::memcpy(v, s->volume_lights.data(), s->volume_lights.size() * sizeof(rg_volume_light));
    return RG_OK;
}

int rg_scene_set_point_lights(rg_scene *scene, const rg_point_light *lights, uint32_t n) {
    Scene *s = reinterpret_cast<Scene *>(scene);
    if (!s) return set_error(RG_E_INVALID, "null scene");
    if (n && !lights) return set_error(RG_E_INVALID, "lights is null");
    for (uint32_t i = 0; i < n; i++)
        for (float v : {lights[i].pos[0], lights[i].pos[1], lights[i].pos[2], lights[i].normal[0],
                        lights[i].normal[1], lights[i].normal[2], lights[i].color[0], lights[i].color[1],
                        lights[i].color[2], lights[i].emission, lights[i].spot_range})
            if (!std::isfinite(v)) return set_error(RG_E_INVALID, "point light " + std::to_string(i) + " not finite");
    s->point_lights.assign(lights, lights + n);
    s->pt_dirty = true;
    return RG_OK;
}

int rg_scene_set_camera(rg_scene *scene, const rg_camera *cam) {
    Scene *s = reinterpret_cast<Scene *>(scene);
    if (!s || !cam) return set_error(RG_E_INVALID, "null argument");
    s->camera = *cam;
    return RG_OK;
}

int rg_render(rg_ctx *ctx, rg_scene *scene, const rg_settings *settings, float *hdr_rgba, float *sum_rgb,
              float *sumsq_rgb, uint32_t *nspp) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    Scene *s = reinterpret_cast<Scene *>(scene);
    if (!c || !s || !settings) return set_error(RG_E_INVALID, "null argument");
    const rg_settings &st = *settings;
    if (st.width == 0 || st.height == 0 || st.width > 16384 || st.height > 16384)
        return set_error(RG_E_INVALID, "resolution must be 1..16384 per axis");
    if (st.max_depth < 1 || st.max_depth > 64) return set_error(RG_E_INVALID, "max_depth must be 1..64");
    if (int rc = check_shader(st)) return rc;
    if (c->device < 0) return set_error(RG_E_NODEVICE, "host-only context has no device");
    // a render clears the cancel flag when it starts (rayrenderer.cpp:537)
    c->cancel.store(false);
    if (int rc = dev_set_device(c->device)) return rc;
    void *stream = c->stream;
    if (int rc = prepare_scene(s, stream)) return rc;

    const uint32_t W = st.width, H = st.height, npix = W * H;
    const uint32_t tx = (W + 31) / 32, ty = (H + 31) / 32, ntiles = tx * ty;
    const uint32_t spp = std::max<uint32_t>(1, st.spp);
    uint32_t S = st.samples_in_flight;
    if (S == 0) S = std::max<uint32_t>(1, kPathsInFlight / npix);
    S = std::min(S, spp);
    // adaptive passes fused into one generation (a window) when the paths fit:
    // the window's later passes are rendered before the tile decisions that
    // select them are known; the samples of a tile that stops inside the
    // window are not accumulated, so the sums and tile counts are those of
    // one generation per pass
    uint32_t window_max = 1;
    if (st.adaptive) {
        const uint32_t base = std::max<uint32_t>(1, st.adaptive_base);
        S = std::min(S, base);
        if (st.samples_in_flight == 0)
            window_max = std::max<uint32_t>(1, std::min<uint64_t>(8, kPathsInFlight / ((uint64_t)npix * base)));
    }
    const uint64_t slots = (uint64_t)npix * std::max<uint32_t>(S, st.adaptive ? window_max *
                                                                 std::max<uint32_t>(1, st.adaptive_base) : S);
    // the NEE slot index k * slots + p is 32-bit
    if (slots * kNeeSlots > 0xFFFFFFFFull) return set_error(RG_E_BUDGET, "too many samples in flight");
    if (int rc = ensure_buffers(c->pt, slots, npix, ntiles)) return rc;
    PTBuffers &B = c->pt;

    const PTScene sc = make_ptscene(*s, st);
    const rg_camera &cam = s->camera;

    PTState ps;
    ps.ray = (rg_ray *)B.ray;
    ps.hit = (rg_hit *)B.hit;
    ps.beta = (float4 *)B.beta;
    ps.tp = (float4 *)B.tp;
    ps.L = (float4 *)B.L;
    ps.meta = (int4 *)B.meta;
    ps.sray = (rg_ray *)B.sray;
    ps.slot_stride = (uint32_t)B.cap;
    ps.ssrc = (uint32_t *)B.ssrc;
    ps.socc = (uint8_t *)B.socc;
    ps.scontrib = (float4 *)B.scontrib;
    ps.nshadow = (uint32_t *)B.nshadow;
    ps.q[0] = (uint32_t *)B.q0;
    ps.q[1] = (uint32_t *)B.q1;
    ps.count = (uint32_t *)B.count;
    ps.stats = (unsigned long long *)B.stats;
    ps.sample_base = 0;

    float *d_sum = (float *)B.sum, *d_sumsq = (float *)B.sumsq;
    uint32_t *d_nspp = (uint32_t *)B.tile_nspp;
    uint8_t *d_mask = (uint8_t *)B.tile_mask;
    if (int rc = launch_fill_u32(d_sum, 0u, (uint64_t)npix * 3, stream)) return rc;
    if (int rc = launch_fill_u32(d_sumsq, 0u, (uint64_t)npix * 3, stream)) return rc;
    if (int rc = launch_fill_u32(d_nspp, 0u, ntiles, stream)) return rc;
    if (int rc = launch_fill_u32(B.stats, 0u, 6, stream)) return rc;

    // Progress: after queueing a generation the host waits for the one
    // before it (the GPU keeps the new one) and reports the share of the
    // sample budget finished so far.
    float reported = 0.0f;
    auto report = [&](float f) {
        if (c->progress && f > reported) {
            reported = f;
            c->progress(f, c->progress_user);
        }
    };
    if (c->progress)
        for (void *&e : c->progress_ev)
            if (!e)
                if (int rc = hip_ok(hipEventCreateWithFlags((hipEvent_t *)&e, hipEventDisableTiming), "event"))
                    return rc;
    int gen = 0;
    float queued_frac = 0.0f;  // budget share of the samples queued before the newest generation
    // Queues one generation of n samples from sample `from` (and its
    // accumulation when `accumulate`).
    auto queue_generation = [&](uint32_t n, uint32_t from, const uint8_t *mask, bool accumulate) -> int {
        if (c->cancel.load()) return set_error(RG_E_CANCELLED, "cancelled");
        if (int rc = launch_generation(*s, sc, ps, n, from, mask, d_sum, d_sumsq, st.firefly_clamp, stream,
                                       accumulate))
            return rc;
        if (c->progress) {
            hipEvent_t ev = (hipEvent_t)c->progress_ev[gen & 1];
            if (int rc = hip_ok(hipEventRecord(ev, (hipStream_t)stream), "event record")) return rc;
            if (gen > 0) {
                if (int rc = hip_ok(hipEventSynchronize((hipEvent_t)c->progress_ev[(gen - 1) & 1]), "progress wait"))
                    return rc;
                report(queued_frac);
            }
            queued_frac = (float)(from + n) / (float)spp;
            gen++;
        }
        return RG_OK;
    };
    auto run_samples = [&](uint32_t first, uint32_t count, const uint8_t *mask) -> int {
        for (uint32_t done = 0; done < count;) {
            const uint32_t n = std::min(S, count - done);
            if (int rc = queue_generation(n, first + done, mask, true)) return rc;
            done += n;
        }
        return RG_OK;
    };

    if (!st.adaptive) {
        if (int rc = run_samples(0, spp, nullptr)) return rc;
        if (int rc = launch_fill_u32(d_nspp, spp, ntiles, stream)) return rc;
    } else {
        const uint32_t base = std::max<uint32_t>(1, st.adaptive_base);
        const uint32_t first = std::min(base, spp);
        if (int rc = run_samples(0, first, nullptr)) return rc;
        if (int rc = launch_fill_u32(d_nspp, first, ntiles, stream)) return rc;
        uint32_t spent = first;
        uint32_t *d_nactive = (uint32_t *)B.n_active;
        uint32_t window = 1;  // grows 1, 2, 4, ... up to window_max passes per generation
        while (spent < spp) {
            if (int rc = launch_tile_noise(d_sum, d_sumsq, d_nspp, W, H, spp, st.adaptive_threshold, d_mask,
                                           d_nactive, stream))
                return rc;
            uint32_t nact = 0;
            if (int rc = dev_download(&nact, d_nactive, 4, stream)) return rc;
            if (nact == 0) break;
            const uint32_t passes_left = (spp - spent + base - 1) / base;
            const uint32_t w = std::min(window, passes_left);
            const uint32_t count = std::min(w * base, spp - spent);
            if (w == 1) {
                if (int rc = queue_generation(count, spent, d_mask, true)) return rc;
                if (int rc = launch_tiles_add(d_nspp, d_mask, ntiles, count, stream)) return rc;
            } else {
                // the window's paths are those of the tiles active now; pass
                // j's samples are added for the tiles the decision after pass
                // j - 1 keeps, on the device, in pass order
                if (int rc = queue_generation(count, spent, d_mask, false)) return rc;
                for (uint32_t off = 0; off < count;) {
                    const uint32_t pc = std::min(base, count - off);
                    if (off > 0)
                        if (int rc = launch_tile_noise(d_sum, d_sumsq, d_nspp, W, H, spp, st.adaptive_threshold,
                                                       d_mask, d_nactive, stream))
                            return rc;
                    if (int rc = launch_accumulate(sc, ps, off, pc, d_mask, d_sum, d_sumsq, st.firefly_clamp, stream))
                        return rc;
                    if (int rc = launch_tiles_add(d_nspp, d_mask, ntiles, pc, stream)) return rc;
                    off += pc;
                }
            }
            spent += count;
            window = std::min(window * 2, window_max);
        }
    }
    float *d_hdr = (float *)B.hdr;
    if (int rc = launch_resolve(d_sum, d_nspp, W, H, cam.exposure, d_hdr, stream)) return rc;
    if (st.denoise) {
        if (int rc = launch_aov(*s, sc, B.aov, stream)) return rc;
        if (int rc = launch_denoise(st, W, H, sc, *s, d_hdr, B.aov, B.dn_a, B.dn_b, stream)) return rc;
    }
    if (hdr_rgba)
        if (int rc = dev_download(hdr_rgba, d_hdr, (size_t)npix * 16, stream)) return rc;
    if (sum_rgb)
        if (int rc = dev_download(sum_rgb, d_sum, (size_t)npix * 12, stream)) return rc;
    if (sumsq_rgb)
        if (int rc = dev_download(sumsq_rgb, d_sumsq, (size_t)npix * 12, stream)) return rc;
    if (nspp)
        if (int rc = dev_download(nspp, d_nspp, (size_t)ntiles * 4, stream)) return rc;
    if (int rc = dev_download(c->last_stats, B.stats, sizeof c->last_stats, stream)) return rc;
    if (int rc = hip_ok(hipStreamSynchronize((hipStream_t)stream), "render sync")) return rc;
    report(1.0f);
    return RG_OK;
}

int rg_render_stats(rg_ctx *ctx, uint64_t out[3]) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !out) return set_error(RG_E_INVALID, "null argument");
    for (int k = 0; k < 3; k++) out[k] = c->last_stats[k];
    return RG_OK;
}

}  // extern "C"

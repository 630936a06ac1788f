// rgx — command-line front-end with the reference's render flags
// (src/main.cpp:162-490): `rgx render <scene.json> [options]`.
//
//   -o FILE            output; .hdr = linear HDR (RGBE, after bloom when
//                      enabled), .jpg/.png/.bmp/.ppm = the tonemapped image
//                      (rg_postprocess: Reinhard on luminance + 1/2.2 gamma,
//                      RayRenderer::applyTonemapGamma); default
//                      <scene dir>/<scene name>.jpg (main.cpp:402-410)
//   -enpp B            HDR bloom (enableRenderingPostProcess)
//   -blth F -blst F -blcv F -blrd F   bloom threshold / strength / curve /
//                      radius (--bloom-threshold, ...), main.cpp:225-262
//   -r W,H | WxH | N   resolution (N = N x N)
//   -s N               samples per pixel
//   -c N               host threads (accepted; the GPU path does not use it)
//   -d N               shading system; only 5 (BSDF) is on this path
//   -enaa B  -endn B  -enad B   antialias / denoise / adaptive (yes|true|1|on)
//   -dni F  --adaptive-base N  --adaptive-threshold F
//   -encs B            colour sampling (enableColorSampling, default on)
//   -cb B              cullBackFace (accepted; shading system 5 never reads it)
//   -bc R,G,B[,A]      backColor (alpha 1 when only R,G,B are given)
//   --depth N          trace depth (the reference's MAX_TRACE_DEPTH, 32)
//   --seed N           Philox key of the sample stream
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rg.h"

static bool parse_bool(const char *s) {
    return !strcasecmp(s, "yes") || !strcasecmp(s, "true") || !strcmp(s, "1") || !strcasecmp(s, "on");
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::printf("usage: rgx render <scene.json> [-o out.jpg|.png|.bmp|.ppm|.hdr] [-r WxH] [-s N] [-d 0|1|2|3|5] [-enaa B] ...\n");
        return 0;
    }
    rg_settings st;
    rg_settings_default(&st);
    rg_bloom bloom;
    rg_bloom_default(&bloom);
    bool post = false;
    std::string cmd, scene, out;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        auto next = [&]() -> const char * {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", a);
                std::exit(1);
            }
            return argv[++i];
        };
        if (a[0] == '-') {
            if (!strcmp(a, "-o")) out = next();
            else if (!strcmp(a, "-r") || !strcmp(a, "--resolution")) {
                const char *v = next();
                int w = (int)st.width, h = (int)st.height;
                int n = std::sscanf(v, "%d,%d", &w, &h);
                if (n < 2) n = std::sscanf(v, "%dx%d", &w, &h);
                if (n == 1) h = w;
                if (n >= 1) {
                    st.width = (uint32_t)w;
                    st.height = (uint32_t)h;
                }
            } else if (!strcmp(a, "-s") || !strcmp(a, "--samples")) st.spp = (uint32_t)std::atoi(next());
            else if (!strcmp(a, "-c") || !strcmp(a, "--threads") || !strcmp(a, "--cores")) (void)next();
            else if (!strcmp(a, "-d") || !strcmp(a, "--shader")) st.shader = std::atoi(next());
            else if (!strcmp(a, "-enaa") || !strcmp(a, "--enable-antialias")) st.antialias = parse_bool(next());
            else if (!strcmp(a, "-endn") || !strcmp(a, "--enable-denoise")) st.denoise = parse_bool(next());
            else if (!strcmp(a, "-dni") || !strcmp(a, "--denoise-intensity")) st.denoise_intensity = (float)std::atof(next());
            else if (!strcmp(a, "-enad") || !strcmp(a, "--enable-adaptive")) st.adaptive = parse_bool(next());
            else if (!strcmp(a, "--adaptive-base")) st.adaptive_base = (uint32_t)std::atoi(next());
            else if (!strcmp(a, "--adaptive-threshold")) st.adaptive_threshold = (float)std::atof(next());
            else if (!strcmp(a, "-enpp") || !strcmp(a, "--enable-postprocess")) post = parse_bool(next());
            else if (!strcmp(a, "-blth") || !strcmp(a, "--bloom-threshold")) bloom.threshold = (float)std::atof(next());
            else if (!strcmp(a, "-blst") || !strcmp(a, "--bloom-strength")) bloom.strength = (float)std::atof(next());
            else if (!strcmp(a, "-blcv") || !strcmp(a, "--bloom-curve")) bloom.curve = (float)std::atof(next());
            else if (!strcmp(a, "-blrd") || !strcmp(a, "--bloom-radius")) bloom.radius = (float)std::atof(next());
            else if (!strcmp(a, "-encs") || !strcmp(a, "--enable-color-sampling")) st.color_sampling = parse_bool(next());
            else if (!strcmp(a, "-cb") || !strcmp(a, "--cullback")) st.cull_back_face = parse_bool(next());
            else if (!strcmp(a, "-bc") || !strcmp(a, "--backcolor")) {
                float c[4] = {st.back_color[0], st.back_color[1], st.back_color[2], st.back_color[3]};
                const int n = std::sscanf(next(), "%f,%f,%f,%f", &c[0], &c[1], &c[2], &c[3]);
                if (n == 3) c[3] = 1.0f;
                std::memcpy(st.back_color, c, sizeof c);
            } else if (!strcmp(a, "-wc") || !strcmp(a, "--worldcolor")) {
                // alpha is read and dropped: worldColor is a colour3 (rayrenderer.h:151)
                float c[4] = {st.world_color[0], st.world_color[1], st.world_color[2], 1.0f};
                std::sscanf(next(), "%f,%f,%f,%f", &c[0], &c[1], &c[2], &c[3]);
                std::memcpy(st.world_color, c, sizeof st.world_color);
            }
            else if (!strcmp(a, "--depth")) st.max_depth = (uint32_t)std::atoi(next());
            else if (!strcmp(a, "--seed")) st.seed = (uint32_t)std::strtoul(next(), nullptr, 0);
            else if (!strcmp(a, "-h") || !strcmp(a, "--help")) {
                std::printf("rgx render <scene.json> -o out.hdr -r 800,600 -s 100 -d 5 -enaa true\n");
                return 0;
            } else {
                std::printf("unknown argument: %s\n", a);
                return 1;
            }
        } else if (cmd.empty()) {
            cmd = a;
        } else if (scene.empty()) {
            scene = a;
        } else {
            std::printf("unknown argument: %s\n", a);
            return 1;
        }
    }
    if (st.shader != 0 && st.shader != 1 && st.shader != 2 && st.shader != 3 && st.shader != 5) {
        std::fprintf(stderr, "unknown shading system %d (0, 1, 2, 3 or 5)\n", st.shader);
        return 1;
    }
    if (cmd != "render") {
        std::fprintf(stderr, "only the render command is on this path\n");
        return 1;
    }
    if (out.empty()) {  // <scene dir>/<scene base name>.jpg
        const size_t slash = scene.find_last_of('/');
        const size_t dot = scene.find_last_of('.');
        const size_t base = slash == std::string::npos ? 0 : slash + 1;
        out = scene.substr(0, (dot != std::string::npos && dot > base) ? dot : scene.size()) + ".jpg";
    }
    rg_ctx *ctx = nullptr;
    if (rg_create(0, &ctx) != RG_OK) {
        std::fprintf(stderr, "rg_create: %s\n", rg_last_error());
        return 1;
    }
    auto t0 = std::chrono::steady_clock::now();
    rg_scene *sc = nullptr;
    if (rg_scene_load_json(ctx, scene.c_str(), &sc) != RG_OK) {
        std::fprintf(stderr, "load %s: %s\n", scene.c_str(), rg_last_error());
        return 1;
    }
    char warn[4096];
    if (rg_scene_warnings(sc, warn, sizeof warn) > 0) std::fprintf(stderr, "%s", warn);
    if (rg_bvh_build(ctx, sc, nullptr) != RG_OK) {
        std::fprintf(stderr, "bvh: %s\n", rg_last_error());
        return 1;
    }
    std::printf("resolution: %u x %u, samples: %u, shader: %d, antialias: %s, denoise: %s, adaptive: %s\n", st.width,
                st.height, st.spp, st.shader, st.antialias ? "yes" : "no", st.denoise ? "yes" : "no",
                st.adaptive ? "yes" : "no");
    std::vector<float> hdr((size_t)st.width * st.height * 4);
    // progress line as main.cpp:492-506 prints it
    rg_set_progress(
        ctx,
        [](float p, void *) {
            std::printf("\rrendering... %d%%   ", (int)(p * 100.0f));
            std::fflush(stdout);
        },
        nullptr);
    if (rg_render(ctx, sc, &st, hdr.data(), nullptr, nullptr, nullptr) != RG_OK) {
        std::fprintf(stderr, "\nrender: %s\n", rg_last_error());
        return 1;
    }
    std::printf("\n");
    // render()'s output stage: bloom (when enabled) and the tonemap
    const bool is_hdr = out.size() >= 4 && (out.compare(out.size() - 4, 4, ".hdr") == 0);
    std::vector<float> ldr(is_hdr ? 0 : hdr.size());
    if (rg_postprocess(ctx, st.width, st.height, hdr.data(), post ? &bloom : nullptr,
                       is_hdr ? nullptr : ldr.data()) != RG_OK) {
        std::fprintf(stderr, "post-process: %s\n", rg_last_error());
        return 1;
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rg_image_write(out.c_str(), st.width, st.height, is_hdr ? hdr.data() : ldr.data(), 0) != RG_OK) {
        std::fprintf(stderr, "%s\n", rg_last_error());
        return 1;
    }
    std::printf("done. (%.3f s) -> %s\n", secs, out.c_str());
    rg_scene_destroy(sc);
    rg_destroy(ctx);
    return 0;
}

/*
 * rg_oracle_post.c — CPU restatement of the reference's output stage: the
 * energy-based HDR bloom (RayRenderer::applyPostProcess,
 * src/raygen/rayrenderer.cpp:598-697, with downsampleArea :386-415 and
 * separableGaussianBlur :425-477) and the luminance Reinhard + 1/2.2 gamma
 * tonemap (RayRenderer::applyTonemapGamma, :2614-2647).
 *
 * TEST INFRASTRUCTURE ONLY (the checker; see rg_oracle.h).
 *
 * Parity pins (tests/test_oracle_pins.py):
 *   or_tonemap   closed forms: grey c -> (c/(1+c))^(1/2.2), black -> 0,
 *                a saturated colour whose Reinhard peak exceeds 1 -> the
 *                colour normalised by its peak, alpha copied.
 *   or_bloom     strength 0 and all-below-threshold images are unchanged;
 *                with size_aspect 1 and radius 0 it adds exactly
 *                strength * c * (L - threshold) / L; an integer-ratio
 *                downsample is the box mean; the blur of a single bright
 *                pixel is the outer product of the twice-convolved normalised
 *                Gaussian (computed independently in float64).
 * The upsample is the one step the reference delegates to an external
 * library (ugm Image::resize, not vendored): taken here as bilinear with
 * aligned pixel centres and clamped edges — parity unpinned for that step;
 * the tests pin the convention itself (constant and linear images).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "rg_oracle.h"

static float luma(const float *c) { return 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2]; }

/* downsampleArea (rayrenderer.cpp:386-415) */
static void downsample_area(const float *src, int sw, int sh, float *dst, int dw, int dh) {
    for (int y = 0; y < dh; ++y) {
        const int y0 = (int)((int64_t)y * sh / dh);
        int y1 = (int)((int64_t)(y + 1) * sh / dh);
        if (y1 < y0 + 1) y1 = y0 + 1;
        for (int x = 0; x < dw; ++x) {
            const int x0 = (int)((int64_t)x * sw / dw);
            int x1 = (int)((int64_t)(x + 1) * sw / dw);
            if (x1 < x0 + 1) x1 = x0 + 1;
            float ar = 0.0f, ag = 0.0f, ab = 0.0f, aa = 0.0f;
            int count = 0;
            for (int sy = y0; sy < y1 && sy < sh; ++sy)
                for (int sx = x0; sx < x1 && sx < sw; ++sx) {
                    const float *c = src + 4 * ((size_t)sy * sw + sx);
                    ar += c[0];
                    ag += c[1];
                    ab += c[2];
                    aa += c[3];
                    count++;
                }
            float *o = dst + 4 * ((size_t)y * dw + x);
            if (count > 0) {
                const float inv = 1.0f / (float)count;
                o[0] = ar * inv;
                o[1] = ag * inv;
                o[2] = ab * inv;
                o[3] = aa * inv;
            }
        }
    }
}

/* separableGaussianBlur (rayrenderer.cpp:425-477) */
static void gaussian_blur(float *img, int w, int h, float sigma) {
    if (w <= 0 || h <= 0 || sigma < 0.25f) return;
    int r = (int)ceilf(sigma * 3.0f);
    if (r < 1) r = 1;
    const int n = 2 * r + 1;
    float *k = (float *)malloc(sizeof(float) * (size_t)n);
    float sum = 0.0f;
    const float two_s2 = 2.0f * sigma * sigma;
    for (int i = 0; i < n; ++i) {
        const float x = (float)(i - r);
        const float v = expf(-(x * x) / two_s2);
        k[i] = v;
        sum += v;
    }
    const float inv = 1.0f / sum;
    for (int i = 0; i < n; ++i) k[i] *= inv;
    float *tmp = (float *)calloc((size_t)w * h * 4, sizeof(float));
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            float a[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int j = -r; j <= r; ++j) {
                int sx = x + j;
                if (sx < 0) sx = 0;
                else if (sx >= w) sx = w - 1;
                const float *c = img + 4 * ((size_t)y * w + sx);
                const float q = k[j + r];
                for (int ch = 0; ch < 4; ch++) a[ch] += c[ch] * q;
            }
            memcpy(tmp + 4 * ((size_t)y * w + x), a, sizeof a);
        }
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            float a[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int j = -r; j <= r; ++j) {
                int sy = y + j;
                if (sy < 0) sy = 0;
                else if (sy >= h) sy = h - 1;
                const float *c = tmp + 4 * ((size_t)sy * w + x);
                const float q = k[j + r];
                for (int ch = 0; ch < 4; ch++) a[ch] += c[ch] * q;
            }
            memcpy(img + 4 * ((size_t)y * w + x), a, sizeof a);
        }
    free(tmp);
    free(k);
}

/* Bilinear resize with aligned pixel centres, edges clamped (stands in for
 * the external Image::resize; see the header note). */
static void upsample_bilinear(const float *g, int gw, int gh, float *out, int w, int h) {
    const float rx = (float)gw / (float)w, ry = (float)gh / (float)h;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            float fx = ((float)x + 0.5f) * rx - 0.5f;
            float fy = ((float)y + 0.5f) * ry - 0.5f;
            if (fx < 0.0f) fx = 0.0f;
            if (fy < 0.0f) fy = 0.0f;
            int x0 = (int)fx, y0 = (int)fy;
            if (x0 > gw - 1) x0 = gw - 1;
            if (y0 > gh - 1) y0 = gh - 1;
            const int x1 = x0 + 1 < gw ? x0 + 1 : gw - 1;
            const int y1 = y0 + 1 < gh ? y0 + 1 : gh - 1;
            const float tx = fx - (float)x0, ty = fy - (float)y0;
            const float ux = 1.0f - tx, uy = 1.0f - ty;
            const float *a = g + 4 * ((size_t)y0 * gw + x0), *b = g + 4 * ((size_t)y0 * gw + x1);
            const float *c = g + 4 * ((size_t)y1 * gw + x0), *d = g + 4 * ((size_t)y1 * gw + x1);
            float *o = out + 4 * ((size_t)y * w + x);
            for (int ch = 0; ch < 3; ch++) {
                const float top = a[ch] * ux + b[ch] * tx;
                const float bot = c[ch] * ux + d[ch] * tx;
                o[ch] = top * uy + bot * ty;
            }
            o[3] = 0.0f;
        }
}

void or_bloom(float *hdr, int w, int h, float threshold_in, float strength, float radius, float size_aspect,
              float curve_in) {
    if (w <= 0 || h <= 0) return;
    const size_t npix = (size_t)w * h;
    float *glow = (float *)calloc(npix * 4, sizeof(float));
    const float threshold = fmaxf(threshold_in, 0.0f);
    const float curve = fmaxf(curve_in, 1e-3f);
    for (size_t i = 0; i < npix; i++) {
        const float *c = hdr + 4 * i;
        float *g = glow + 4 * i;
        const float L = luma(c);
        g[3] = c[3];
        if (L > threshold && L > 1e-6f) {
            float scale = (L - threshold) / L;
            if (curve != 1.0f) scale = powf(scale, curve);
            g[0] = c[0] * scale;
            g[1] = c[1] * scale;
            g[2] = c[2] * scale;
        }
    }
    const float aspect = fmaxf(0.02f, size_aspect);
    int gw = (int)((float)w * aspect), gh = (int)((float)h * aspect);
    if (gw < 1) gw = 1;
    if (gh < 1) gh = 1;
    float *small = (float *)calloc((size_t)gw * gh * 4, sizeof(float));
    downsample_area(glow, w, h, small, gw, gh);
    const float sigma_down = fmaxf(0.0f, radius) * (float)w * aspect;
    if (sigma_down >= 0.5f) {
        gaussian_blur(small, gw, gh, sigma_down * 0.70710678f);
        gaussian_blur(small, gw, gh, sigma_down * 0.70710678f);
    }
    upsample_bilinear(small, gw, gh, glow, w, h);
    for (size_t i = 0; i < npix; i++) {
        float *a = hdr + 4 * i;
        const float *g = glow + 4 * i;
        a[0] = a[0] + g[0] * strength;
        a[1] = a[1] + g[1] * strength;
        a[2] = a[2] + g[2] * strength;
    }
    free(small);
    free(glow);
}

void or_tonemap(const float *hdr, int w, int h, float *ldr) {
    const float inv_gamma = 1.0f / 2.2f;
    for (size_t i = 0; i < (size_t)w * h; i++) {
        const float *c = hdr + 4 * i;
        const float L = luma(c);
        float mr = 0.0f, mg = 0.0f, mb = 0.0f;
        if (L > 1e-6f) {
            const float Lm = L / (L + 1.0f);
            const float scale = Lm / L;
            mr = c[0] * scale;
            mg = c[1] * scale;
            mb = c[2] * scale;
            const float peak = fmaxf(fmaxf(mr, mg), mb);
            if (peak > 1.0f) {
                const float inv = 1.0f / peak;
                mr *= inv;
                mg *= inv;
                mb *= inv;
            }
        }
        float *o = ldr + 4 * i;
        o[0] = powf(fmaxf(mr, 0.0f), inv_gamma);
        o[1] = powf(fmaxf(mg, 0.0f), inv_gamma);
        o[2] = powf(fmaxf(mb, 0.0f), inv_gamma);
        o[3] = c[3];
        for (int ch = 0; ch < 3; ch++)
            if (o[ch] > 1.0f) o[ch] = 1.0f;
    }
}

// rg_post.hip — the output stage after the path tracer, for gfx950: the
// energy-based HDR bloom (RayRenderer::applyPostProcess,
// rayrenderer.cpp:598-697) and the luminance Reinhard + 1/2.2 gamma tonemap
// (RayRenderer::applyTonemapGamma, rayrenderer.cpp:2614-2647).
//
// Bloom extracts the radiance above a luma threshold, area-averages it down
// to a glow buffer of size_aspect x the image, blurs that with two separable
// Gaussian passes, upsamples it bilinearly and adds it to the HDR image.
// The launches:
//   k_glow_down    extract + area downsample, fused: one thread per glow
//                  pixel reads its source box, so the image is read once
//   k_blur<axis>   one 1-D Gaussian pass over the small buffer (four: two
//                  sigma/sqrt2 passes, each horizontal then vertical)
//   k_compose      bilinear upsample + HDR add + tonemap, fused: reads each
//                  HDR pixel once and writes the bloomed HDR and the LDR pixel
// The Gaussian weights are built on the host the way the reference builds
// them (binary32 expf, normalised by their float sum).  Every sum runs in the
// reference's order without contraction, so the bloomed HDR is the
// restatement's binary32 result; only powf (curve != 1, gamma) may differ in
// the last place from the host library.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "rg_internal.h"

namespace rg {

namespace {

constexpr int kBlk = 256;

__device__ __forceinline__ float luma(float r, float g, float b) { return 0.2126f * r + 0.7152f * g + 0.0722f * b; }

// Glow pixel (x, y) of a gw x gh buffer: the mean over its source box of the
// radiance above `threshold` (applyPostProcess :611-632 + downsampleArea
// :386-415).
__global__ __launch_bounds__(kBlk) void k_glow_down(const float4 *__restrict__ hdr, int sw, int sh, int dw, int dh,
                                                     float threshold, float curve, float4 *__restrict__ glow) {
    const int i = blockIdx.x * kBlk + threadIdx.x;
    if (i >= dw * dh) return;
    const int x = i % dw, y = i / dw;
    const int y0 = (int)((int64_t)y * sh / dh);
    const int y1 = max(y0 + 1, (int)((int64_t)(y + 1) * sh / dh));
    const int x0 = (int)((int64_t)x * sw / dw);
    const int x1 = max(x0 + 1, (int)((int64_t)(x + 1) * sw / dw));
    float ar = 0.0f, ag = 0.0f, ab = 0.0f, aa = 0.0f;
    int count = 0;
    for (int sy = y0; sy < y1 && sy < sh; ++sy) {
        for (int sx = x0; sx < x1 && sx < sw; ++sx) {
            const float4 c = hdr[(size_t)sy * sw + sx];
            const float L = luma(c.x, c.y, c.z);
            float gr = 0.0f, gg = 0.0f, gb = 0.0f;
            if (L > threshold && L > 1e-6f) {
                float scale = (L - threshold) / L;
                if (curve != 1.0f) scale = powf(scale, curve);
                gr = c.x * scale;
                gg = c.y * scale;
                gb = c.z * scale;
            }
            ar += gr;
            ag += gg;
            ab += gb;
            aa += c.w;
            count++;
        }
    }
    float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (count > 0) {
        const float inv = 1.0f / (float)count;
        o = make_float4(ar * inv, ag * inv, ab * inv, aa * inv);
    }
    glow[i] = o;
}

// Gaussian taps handed to the blur by value (kernel arguments), for the
// entry point that keeps no device buffer of its own.
constexpr int kArgTapR = 127;
struct TapArg {
    float w[2 * kArgTapR + 1];
    __device__ __forceinline__ float operator[](int i) const { return w[i]; }
};

// One pass of separableGaussianBlur (rayrenderer.cpp:425-477): taps
// -r..r in order, clamp to the edge.  Wt: a device pointer to the taps or
// a TapArg.
template <int kAxis, class Wt>
__global__ __launch_bounds__(kBlk) void k_blur(const float4 *__restrict__ src, float4 *__restrict__ dst, int w, int h,
                                                Wt wt, int r) {
    const int i = blockIdx.x * kBlk + threadIdx.x;
    if (i >= w * h) return;
    const int x = i % w, y = i / w;
    float ar = 0.0f, ag = 0.0f, ab = 0.0f, aa = 0.0f;
    for (int k = -r; k <= r; ++k) {
        int sx = x, sy = y;
        if (kAxis == 0) {
            sx = x + k;
            sx = sx < 0 ? 0 : (sx >= w ? w - 1 : sx);
        } else {
            sy = y + k;
            sy = sy < 0 ? 0 : (sy >= h ? h - 1 : sy);
        }
        const float4 c = src[(size_t)sy * w + sx];
        const float q = wt[k + r];
        ar += c.x * q;
        ag += c.y * q;
        ab += c.z * q;
        aa += c.w * q;
    }
    dst[i] = make_float4(ar, ag, ab, aa);
}

// Bilinear sample of the gw x gh glow buffer at full-resolution pixel (x, y):
// pixel centres aligned, clamped at the edges (the reference's
// Image::resize is an external library's; this is the convention taken).
__device__ __forceinline__ float4 glow_at(const float4 *__restrict__ g, int gw, int gh, float rx, float ry, int x,
                                          int y) {
    float fx = fmaxf(((float)x + 0.5f) * rx - 0.5f, 0.0f);
    float fy = fmaxf(((float)y + 0.5f) * ry - 0.5f, 0.0f);
    int x0 = min((int)fx, gw - 1), y0 = min((int)fy, gh - 1);
    const int x1 = min(x0 + 1, gw - 1), y1 = min(y0 + 1, gh - 1);
    const float tx = fx - (float)x0, ty = fy - (float)y0;
    const float4 a = g[(size_t)y0 * gw + x0], b = g[(size_t)y0 * gw + x1];
    const float4 c = g[(size_t)y1 * gw + x0], d = g[(size_t)y1 * gw + x1];
    const float ux = 1.0f - tx, uy = 1.0f - ty;
    const float tr = a.x * ux + b.x * tx, tg = a.y * ux + b.y * tx, tb = a.z * ux + b.z * tx;
    const float br = c.x * ux + d.x * tx, bg = c.y * ux + d.y * tx, bb = c.z * ux + d.z * tx;
    return make_float4(tr * uy + br * ty, tg * uy + bg * ty, tb * uy + bb * ty, 0.0f);
}

// Upsample + unclamped HDR add (applyPostProcess :684-696), then the
// tonemap (applyTonemapGamma :2614-2647) of the result.
__global__ __launch_bounds__(kBlk) void k_compose(float4 *__restrict__ hdr, int w, int h,
                                                   const float4 *__restrict__ glow, int gw, int gh, float rx,
                                                   float ry, float strength, int bloom, float4 *__restrict__ ldr) {
    const int i = blockIdx.x * kBlk + threadIdx.x;
    if (i >= w * h) return;
    float4 c = hdr[i];
    if (bloom) {
        const float4 g = glow_at(glow, gw, gh, rx, ry, i % w, i / w);
        c = make_float4(c.x + g.x * strength, c.y + g.y * strength, c.z + g.z * strength, c.w);
        hdr[i] = c;
    }
    if (!ldr) return;
    const float L = luma(c.x, c.y, c.z);
    float mr = 0.0f, mg = 0.0f, mb = 0.0f;
    if (L > 1e-6f) {
        const float Lm = L / (L + 1.0f);
        const float scale = Lm / L;
        mr = c.x * scale;
        mg = c.y * scale;
        mb = c.z * scale;
        const float peak = fmaxf(fmaxf(mr, mg), mb);
        if (peak > 1.0f) {
            const float inv = 1.0f / peak;
            mr *= inv;
            mg *= inv;
            mb *= inv;
        }
    }
    const float ig = 1.0f / 2.2f;
    float4 o = make_float4(powf(fmaxf(mr, 0.0f), ig), powf(fmaxf(mg, 0.0f), ig), powf(fmaxf(mb, 0.0f), ig), c.w);
    if (o.x > 1.0f) o.x = 1.0f;
    if (o.y > 1.0f) o.y = 1.0f;
    if (o.z > 1.0f) o.z = 1.0f;
    ldr[i] = o;
}

int hip_chk(hipError_t e, const char *what) {
    if (e == hipSuccess) return RG_OK;
    return set_error(RG_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Grows a device buffer to at least `bytes`.
int grow(void *&p, uint64_t &cap, uint64_t bytes) {
    if (bytes <= cap) return RG_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(RG_E_NOMEM, "post-process buffer allocation failed");
    }
    cap = bytes;
    return RG_OK;
}

// Taps of separableGaussianBlur (rayrenderer.cpp:429-441): r = max(1,
// ceil(3 sigma)), w_i = exp(-x^2 / (2 sigma^2)) normalised by its sum.
int gauss_radius(float sigma) { return std::max(1, (int)ceilf(sigma * 3.0f)); }
void gauss_fill(float sigma, int r, float *w) {
    float sum = 0.0f;
    const float two_s2 = 2.0f * sigma * sigma;
    for (int i = 0; i < 2 * r + 1; ++i) {
        const float x = (float)(i - r);
        const float v = expf(-(x * x) / two_s2);
        w[i] = v;
        sum += v;
    }
    const float inv = 1.0f / sum;
    for (int i = 0; i < 2 * r + 1; ++i) w[i] *= inv;
}
int gauss_taps(float sigma, std::vector<float> &w) {
    const int r = gauss_radius(sigma);
    w.assign((size_t)(2 * r + 1), 0.0f);
    gauss_fill(sigma, r, w.data());
    return r;
}

// Scratch of rg_postprocess_device_noalloc: its two glow buffers, at most
// kScratchGlowPix glow pixels each (2 x 16 MiB), one copy per GPU, held
// until the process ends.  Host code uses only the per-device addresses
// that post_scratch_init takes from hipGetSymbolAddress.
constexpr uint32_t kScratchGlowPix = 1u << 20;
constexpr int kMaxScratchDevices = 64;
__device__ __align__(256) float4 scratch[2 * (size_t)kScratchGlowPix];
std::mutex g_scratch_mu;
std::atomic<float4 *> g_scratch_ptr[kMaxScratchDevices];
std::string g_scratch_err[kMaxScratchDevices];

}  // namespace

// Once per device, with that device current (rg_create): the device
// address of the no-allocation entry point's scratch array.
int post_scratch_init(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return RG_OK;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device].load()) return RG_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return set_error(RG_E_HIP, g_scratch_err[device]);
    }
    g_scratch_ptr[device].store((float4 *)p);
    return RG_OK;
}

void release_post(Ctx *c) {
    PostBuffers &p = c->post;
    for (void *q : {p.glow_a, p.glow_b, p.taps, p.img_hdr, p.img_ldr})
        if (q) (void)hipFree(q);
    p = PostBuffers{};
}

// The output stage on device images: bloom in place (when `bloom`), then the
// tonemap into d_ldr (when non-null).
int post_device(Ctx *c, uint32_t W, uint32_t H, float4 *d_hdr, const rg_bloom *bloom, float4 *d_ldr) {
    hipStream_t s = (hipStream_t)c->stream;
    PostBuffers &P = c->post;
    const int w = (int)W, h = (int)H;
    const uint32_t npix = W * H;
    int gw = 1, gh = 1;
    float rx = 1.0f, ry = 1.0f;
    const float4 *glow = nullptr;
    if (bloom) {
        const float threshold = fmaxf(bloom->threshold, 0.0f);
        const float curve = fmaxf(bloom->curve, 1e-3f);
        const float aspect = fmaxf(0.02f, bloom->size_aspect);
        gw = std::max(1, (int)((float)w * aspect));
        gh = std::max(1, (int)((float)h * aspect));
        if ((uint64_t)gw * gh > (1ull << 28)) return set_error(RG_E_INVALID, "bloom buffer too large");
        const uint64_t gbytes = (uint64_t)gw * gh * 16;
        if (int rc = grow(P.glow_a, P.glow_a_cap, gbytes)) return rc;
        if (int rc = grow(P.glow_b, P.glow_b_cap, gbytes)) return rc;
        float4 *ga = (float4 *)P.glow_a, *gb = (float4 *)P.glow_b;
        const uint32_t gblocks = (uint32_t)(((uint64_t)gw * gh + kBlk - 1) / kBlk);
        hipLaunchKernelGGL(k_glow_down, dim3(gblocks), dim3(kBlk), 0, s, d_hdr, w, h, gw, gh, threshold, curve, ga);
        const float sigma_down = fmaxf(0.0f, bloom->radius) * (float)w * aspect;
        const float sigma = sigma_down * 0.70710678f;
        if (sigma_down >= 0.5f && sigma >= 0.25f) {
            if (sigma != P.taps_sigma) {
                // rebuilt and uploaded only when sigma changes; a pending
                // upload from the old host copy must finish first
                if (P.taps_sigma >= 0.0f)
                    if (int rc = hip_chk(hipStreamSynchronize(s), "bloom taps sync")) return rc;
                P.taps_sigma = -1.0f;
                std::vector<float> taps;
                const int r = gauss_taps(sigma, taps);
                if (r > (1 << 20)) return set_error(RG_E_INVALID, "bloom radius too large");
                if (int rc = grow(P.taps, P.taps_cap, taps.size() * 4)) return rc;
                P.taps_host.swap(taps);
                if (int rc = hip_chk(hipMemcpyAsync(P.taps, P.taps_host.data(), P.taps_host.size() * 4,
                                                    hipMemcpyHostToDevice, s),
                                     "bloom taps upload"))
                    return rc;
                P.taps_sigma = sigma;
                P.taps_r = r;
            }
            const int r = P.taps_r;
            const float *wt = (const float *)P.taps;
            for (int pass = 0; pass < 2; pass++) {
                hipLaunchKernelGGL((k_blur<0, const float *>), dim3(gblocks), dim3(kBlk), 0, s, ga, gb, gw, gh, wt, r);
                hipLaunchKernelGGL((k_blur<1, const float *>), dim3(gblocks), dim3(kBlk), 0, s, gb, ga, gw, gh, wt, r);
            }
        }
        glow = ga;
        rx = (float)gw / (float)w;
        ry = (float)gh / (float)h;
    }
    if (!bloom && !d_ldr) return RG_OK;
    hipLaunchKernelGGL(k_compose, dim3((npix + kBlk - 1) / kBlk), dim3(kBlk), 0, s, d_hdr, w, h, glow, gw, gh, rx, ry,
                       bloom ? bloom->strength : 0.0f, bloom ? 1 : 0, d_ldr);
    return hip_chk(hipGetLastError(), "post-process launch");
}

// The output stage with no allocation: the glow buffers live in the
// per-device scratch array, the Gaussian taps travel as kernel arguments.
int post_device_noalloc(Ctx *c, uint32_t W, uint32_t H, float4 *d_hdr, const rg_bloom *bloom, float4 *d_ldr) {
    hipStream_t s = (hipStream_t)c->stream;
    const int w = (int)W, h = (int)H;
    const uint32_t npix = W * H;
    int gw = 1, gh = 1;
    float rx = 1.0f, ry = 1.0f;
    const float4 *glow = nullptr;
    if (bloom) {
        float4 *base = (c->device >= 0 && c->device < kMaxScratchDevices) ? g_scratch_ptr[c->device].load() : nullptr;
        if (!base)
            return set_error(RG_E_HIP, "no-allocation scratch unavailable on this device" +
                                           (c->device >= 0 && c->device < kMaxScratchDevices
                                                ? ": " + g_scratch_err[c->device] : std::string()));
        const float threshold = fmaxf(bloom->threshold, 0.0f);
        const float curve = fmaxf(bloom->curve, 1e-3f);
        const float aspect = fmaxf(0.02f, bloom->size_aspect);
        gw = std::max(1, (int)((float)w * aspect));
        gh = std::max(1, (int)((float)h * aspect));
        if ((uint64_t)gw * gh > kScratchGlowPix)
            return set_error(RG_E_BUDGET, "bloom glow buffer of " + std::to_string(gw) + " x " + std::to_string(gh) +
                                              " pixels exceeds the no-allocation scratch (1,048,576 pixels)");
        const float sigma_down = fmaxf(0.0f, bloom->radius) * (float)w * aspect;
        const float sigma = sigma_down * 0.70710678f;
        const bool blur = sigma_down >= 0.5f && sigma >= 0.25f;
        TapArg taps;
        int r = 0;
        if (blur) {
            r = gauss_radius(sigma);
            if (r > kArgTapR)
                return set_error(RG_E_BUDGET, "bloom blur radius " + std::to_string(r) +
                                                  " exceeds the no-allocation limit of 127 taps per side");
            gauss_fill(sigma, r, taps.w);
        }
        float4 *ga = base, *gb = base + kScratchGlowPix;
        const uint32_t gblocks = (uint32_t)(((uint64_t)gw * gh + kBlk - 1) / kBlk);
        // every glow pixel is written before it is read: nothing to clear
        hipLaunchKernelGGL(k_glow_down, dim3(gblocks), dim3(kBlk), 0, s, d_hdr, w, h, gw, gh, threshold, curve, ga);
        if (blur) {
            for (int pass = 0; pass < 2; pass++) {
                hipLaunchKernelGGL((k_blur<0, TapArg>), dim3(gblocks), dim3(kBlk), 0, s, ga, gb, gw, gh, taps, r);
                hipLaunchKernelGGL((k_blur<1, TapArg>), dim3(gblocks), dim3(kBlk), 0, s, gb, ga, gw, gh, taps, r);
            }
        }
        glow = ga;
        rx = (float)gw / (float)w;
        ry = (float)gh / (float)h;
    }
    if (!bloom && !d_ldr) return RG_OK;
    hipLaunchKernelGGL(k_compose, dim3((npix + kBlk - 1) / kBlk), dim3(kBlk), 0, s, d_hdr, w, h, glow, gw, gh, rx, ry,
                       bloom ? bloom->strength : 0.0f, bloom ? 1 : 0, d_ldr);
    return hip_chk(hipGetLastError(), "post-process launch");
}

}  // namespace rg

using namespace rg;

extern "C" {

void rg_bloom_default(rg_bloom *b) {
    if (!b) return;
    b->threshold = 1.0f;
    b->strength = 1.0f;
    b->radius = 0.03f;
    b->size_aspect = 0.15f;
    b->curve = 1.0f;
    b->pad = 0.0f;
}

static int post_args(rg_ctx *ctx, uint32_t width, uint32_t height, const void *hdr) {
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    if (!c || !hdr) return set_error(RG_E_INVALID, "null argument");
    if (c->device < 0) return set_error(RG_E_NODEVICE, "host-only context has no device");
    if (width == 0 || height == 0 || width > 16384 || height > 16384)
        return set_error(RG_E_INVALID, "resolution must be 1..16384 per axis");
    return hip_chk(hipSetDevice(c->device), "hipSetDevice");
}

int rg_postprocess_device(rg_ctx *ctx, uint32_t width, uint32_t height, float *d_hdr_rgba, const rg_bloom *bloom,
                          float *d_ldr_rgba) {
    if (int rc = post_args(ctx, width, height, d_hdr_rgba)) return rc;
    return post_device(reinterpret_cast<Ctx *>(ctx), width, height, (float4 *)d_hdr_rgba, bloom,
                       (float4 *)d_ldr_rgba);
}

int rg_postprocess_device_noalloc(rg_ctx *ctx, uint32_t width, uint32_t height, float *d_hdr_rgba,
                                  const rg_bloom *bloom, float *d_ldr_rgba) {
    if (int rc = post_args(ctx, width, height, d_hdr_rgba)) return rc;
    return post_device_noalloc(reinterpret_cast<Ctx *>(ctx), width, height, (float4 *)d_hdr_rgba, bloom,
                               (float4 *)d_ldr_rgba);
}

int rg_postprocess(rg_ctx *ctx, uint32_t width, uint32_t height, float *hdr_rgba, const rg_bloom *bloom,
                   float *ldr_rgba) {
    if (int rc = post_args(ctx, width, height, hdr_rgba)) return rc;
    Ctx *c = reinterpret_cast<Ctx *>(ctx);
    PostBuffers &P = c->post;
    hipStream_t s = (hipStream_t)c->stream;
    const uint64_t bytes = (uint64_t)width * height * 16;
    if (int rc = grow(P.img_hdr, P.img_hdr_cap, bytes)) return rc;
    if (ldr_rgba)
        if (int rc = grow(P.img_ldr, P.img_ldr_cap, bytes)) return rc;
    if (int rc = hip_chk(hipMemcpyAsync(P.img_hdr, hdr_rgba, bytes, hipMemcpyHostToDevice, s), "hdr upload"))
        return rc;
    if (int rc = post_device(c, width, height, (float4 *)P.img_hdr, bloom, ldr_rgba ? (float4 *)P.img_ldr : nullptr))
        return rc;
    if (bloom)
        if (int rc = hip_chk(hipMemcpyAsync(hdr_rgba, P.img_hdr, bytes, hipMemcpyDeviceToHost, s), "hdr download"))
            return rc;
    if (ldr_rgba)
        if (int rc = hip_chk(hipMemcpyAsync(ldr_rgba, P.img_ldr, bytes, hipMemcpyDeviceToHost, s), "ldr download"))
            return rc;
    return hip_chk(hipStreamSynchronize(s), "post-process sync");
}

}  // extern "C"

// rg_denoise.hip — primary-hit guide buffers and the edge-avoiding A-Trous
// denoiser (RayRenderer::denoiseImage / atrousPass, rayrenderer.cpp:2536-2772)
// for gfx950.
//
// Each A-Trous level is a 5x5 B3-spline stencil with taps `step` pixels
// apart, weighted by luminance, normal and depth edge-stopping functions.
// Levels with step <= 8 stage their (16 + 4*step)^2 input window — colour,
// decoded normal and depth, 32 B per pixel — in LDS once per 16x16 output
// tile, so each input pixel is fetched from HBM once per tile instead of 25
// times; the widest level (step 16, a 80x80 window) reads through L2.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rg_device.h"
#include "rg_internal.h"
#include "rg_pt.h"

namespace rg {

namespace {

constexpr int kTile = 16;

__device__ __forceinline__ float lum3(float r, float g, float b) { return 0.2126f * r + 0.7152f * g + 0.0722f * b; }

// Guide buffers from the pixel-centre camera ray (accumulatePixelSamples
// :819-844): normal*0.5+0.5 / material colour / 1 - clamp(sqrt(dist/far)),
// alpha 1 on a hit; (0.5, backColor, 0) with alpha 0 on a miss.
template <int kStack, int kW>
__global__ __launch_bounds__(256) void k_aov(DevBvh bv, PTScene sc, float view_far, float4 *__restrict__ nrm,
                                              float4 *__restrict__ alb, float *__restrict__ dep) {
    RG_LDS_COLUMN(lc, 256);
    const uint32_t pix = blockIdx.x * 256 + threadIdx.x;
    const uint32_t npix = sc.width * sc.height;
    if (pix >= npix) return;
    const uint32_t x = pix % sc.width, y = pix / sc.width;
    const float dx = ((float)x + 0.5f - sc.half_w) * sc.view_scale;
    const float dy = -((float)y + 0.5f - sc.half_h) * sc.view_scale;
    float l = sqrtf(dx * dx + dy * dy + 1.0f);
    const float il = 1.0f / l;
    const float ddx = dx * il, ddy = dy * il, ddz = -1.0f * il;
    float t = 1.0e6f, u = 0.0f, v = 0.0f;
    uint32_t prim = 0xFFFFFFFFu;
    trace_closest_dev<kStack, 256, kW>(bv, 0.0f, 0.0f, 0.0f, ddx, ddy, ddz, 0.0f, t, prim, u, v, lc);
    if (prim != 0xFFFFFFFFu) {
        const DevTriShade &s = sc.shade[prim];
        const float w = 1.0f - u - v;
        float nx = s.n[0] * w + s.n[3] * u + s.n[6] * v;
        float ny = s.n[1] * w + s.n[4] * u + s.n[7] * v;
        float nz = s.n[2] * w + s.n[5] * u + s.n[8] * v;
        const float inv = 1.0f / sqrtf(nx * nx + ny * ny + nz * nz);
        nx = nx * inv;
        ny = ny * inv;
        nz = nz * inv;
        nrm[pix] = make_float4(nx * 0.5f + 0.5f, ny * 0.5f + 0.5f, nz * 0.5f + 0.5f, 1.0f);
        const DevObject &m = sc.objs[s.obj];
        alb[pix] = make_float4(m.color[0], m.color[1], m.color[2], 1.0f);
        const float hx = 0.0f + ddx * t, hy = 0.0f + ddy * t, hz = 0.0f + ddz * t;
        float d = sqrtf(hx * hx + hy * hy + hz * hz) / view_far;
        d = sqrtf(d);
        dep[pix] = 1.0f - fminf(fmaxf(d, 0.0f), 1.0f);
    } else {
        nrm[pix] = make_float4(0.5f, 0.5f, 0.5f, 0.0f);
        alb[pix] = make_float4(sc.back[0], sc.back[1], sc.back[2], sc.back[3]);
        dep[pix] = 0.0f;
    }
}

// Firefly clamp against the 3x3 neighbourhood and albedo demodulation
// (denoiseImage :2685-2722).
__global__ __launch_bounds__(256) void k_prepass(const float4 *__restrict__ noisy, const float4 *__restrict__ alb,
                                                  float4 *__restrict__ out, int w, int h) {
    const int pix = blockIdx.x * 256 + threadIdx.x;
    if (pix >= w * h) return;
    const int x = pix % w, y = pix / w;
    float4 c = noisy[pix];
    float mx = 0.0f;
    for (int dy = -1; dy <= 1; dy++) {
        const int ny = y + dy;
        if (ny < 0 || ny >= h) continue;
        for (int dx = -1; dx <= 1; dx++) {
            if (dx == 0 && dy == 0) continue;
            const int nx = x + dx;
            if (nx < 0 || nx >= w) continue;
            const float4 n = noisy[ny * w + nx];
            const float l = lum3(n.x, n.y, n.z);
            if (l > mx) mx = l;
        }
    }
    const float cl = lum3(c.x, c.y, c.z);
    const float cap = 1.5f * mx + 0.01f;
    if (cl > cap) {
        const float s = cap / cl;
        c.x *= s;
        c.y *= s;
        c.z *= s;
    }
    const float4 a = alb[pix];
    if (a.w > 0.5f) {
        out[pix] = make_float4(fminf(c.x / fmaxf(a.x, 0.3f), 3.0f), fminf(c.y / fmaxf(a.y, 0.3f), 3.0f),
                               fminf(c.z / fmaxf(a.z, 0.3f), 3.0f), 1.0f);
    } else {
        out[pix] = c;
    }
}

struct Tap {
    float r, g, b, l;   // colour and its luminance
    float nx, ny, nz, d; // decoded normal, depth
};

__device__ __forceinline__ Tap load_tap(const float4 *src, const float4 *nrm, const float *dep, int i) {
    const float4 c = src[i];
    const float4 n = nrm[i];
    Tap t;
    t.r = c.x;
    t.g = c.y;
    t.b = c.z;
    t.l = lum3(c.x, c.y, c.z);
    t.nx = n.x * 2.0f - 1.0f;
    t.ny = n.y * 2.0f - 1.0f;
    t.nz = n.z * 2.0f - 1.0f;
    t.d = dep[i];
    return t;
}

// One output pixel of atrousPass (:2560-2606); `fetch(x, y)` returns a tap.
template <class F>
__device__ __forceinline__ float4 atrous_pixel(int x, int y, int w, int h, int step, const Tap &c, float alpha,
                                               float inv_c2, float sigma_n, float dscale, F fetch) {
    const float k[5] = {1.0f / 16.0f, 1.0f / 4.0f, 3.0f / 8.0f, 1.0f / 4.0f, 1.0f / 16.0f};
    float sx = 0.0f, sy = 0.0f, sz = 0.0f, ws = 0.0f;
    for (int ky = 0; ky < 5; ky++) {
        const int ny = y + (ky - 2) * step;
        if (ny < 0 || ny >= h) continue;
        for (int kx = 0; kx < 5; kx++) {
            const int nx = x + (kx - 2) * step;
            if (nx < 0 || nx >= w) continue;
            const Tap s = fetch(nx, ny);
            const float ld = s.l - c.l;
            const float wc = expf(-(ld * ld) * inv_c2);
            float nd = s.nx * c.nx + s.ny * c.ny + s.nz * c.nz;
            if (nd < 0.0f) nd = 0.0f;
            const float wn = powf(nd, sigma_n);
            const float wd = expf(-fabsf(s.d - c.d) * dscale);
            const float wt = k[kx] * k[ky] * wc * wn * wd;
            sx += s.r * wt;
            sy += s.g * wt;
            sz += s.b * wt;
            ws += wt;
        }
    }
    if (ws > 1e-8f) return make_float4(sx / ws, sy / ws, sz / ws, alpha);
    return make_float4(c.r, c.g, c.b, alpha);
}

// LDS-tiled level: the 16x16 tile's (16 + 4*step)^2 window is staged once.
template <int kStep>
__global__ __launch_bounds__(256) void k_atrous_lds(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                     const float4 *__restrict__ nrm, const float *__restrict__ dep,
                                                     int w, int h, float inv_c2, float sigma_n, float dscale) {
    constexpr int kWin = kTile + 4 * kStep;
    __shared__ Tap win[kWin * kWin];
    const int x0 = blockIdx.x * kTile - 2 * kStep, y0 = blockIdx.y * kTile - 2 * kStep;
    for (int k = threadIdx.x; k < kWin * kWin; k += 256) {
        const int gx = x0 + k % kWin, gy = y0 + k / kWin;
        if (gx >= 0 && gx < w && gy >= 0 && gy < h) win[k] = load_tap(src, nrm, dep, gy * w + gx);
    }
    __syncthreads();
    const int lx = threadIdx.x % kTile, ly = threadIdx.x / kTile;
    const int x = blockIdx.x * kTile + lx, y = blockIdx.y * kTile + ly;
    if (x >= w || y >= h) return;
    const int ci = (ly + 2 * kStep) * kWin + (lx + 2 * kStep);
    const Tap c = win[ci];
    const float alpha = src[y * w + x].w;
    dst[y * w + x] = atrous_pixel(x, y, w, h, kStep, c, alpha, inv_c2, sigma_n, dscale,
                                  [&](int nx, int ny) { return win[(ny - y0) * kWin + (nx - x0)]; });
}

__global__ __launch_bounds__(256) void k_atrous_global(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                        const float4 *__restrict__ nrm, const float *__restrict__ dep,
                                                        int w, int h, int step, float inv_c2, float sigma_n,
                                                        float dscale) {
    const int lx = threadIdx.x % kTile, ly = threadIdx.x / kTile;
    const int x = blockIdx.x * kTile + lx, y = blockIdx.y * kTile + ly;
    if (x >= w || y >= h) return;
    const Tap c = load_tap(src, nrm, dep, y * w + x);
    const float alpha = src[y * w + x].w;
    dst[y * w + x] = atrous_pixel(x, y, w, h, step, c, alpha, inv_c2, sigma_n, dscale,
                                  [&](int nx, int ny) { return load_tap(src, nrm, dep, ny * w + nx); });
}

// Remodulation and blend with the noisy input (denoiseImage :2752-2771).
__global__ __launch_bounds__(256) void k_remod(const float4 *__restrict__ filt, const float4 *__restrict__ alb,
                                                float4 *__restrict__ hdr, int n, float t) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 f = filt[i], nz = hdr[i], a = alb[i];
    float rx = f.x, ry = f.y, rz = f.z;
    if (a.w > 0.5f) {
        rx = f.x * a.x;
        ry = f.y * a.y;
        rz = f.z * a.z;
    }
    const float s = 1.0f - t;
    hdr[i] = make_float4(nz.x * s + rx * t, nz.y * s + ry * t, nz.z * s + rz * t, 1.0f);
}

int chk(hipError_t e, const char *what) {
    if (e == hipSuccess) return RG_OK;
    return set_error(RG_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace

DevBvh dev_bvh(const Scene &s);

int launch_aov(const Scene &scn, const PTScene &sc, void *d_aov, void *stream) {
    const uint32_t npix = sc.width * sc.height;
    float4 *nrm = (float4 *)d_aov;
    float4 *alb = nrm + npix;
    float *dep = (float *)(alb + npix);
    const DevBvh bv = dev_bvh(scn);
    const dim3 grid((npix + 255) / 256);
    return with_traversal(scn.stack_need, scn.width, [&](auto tag) {
        using T = decltype(tag);
        hipLaunchKernelGGL((k_aov<T::kStack, T::kW>), grid, dim3(256), 0, (hipStream_t)stream, bv, sc,
                           scn.camera.view_far, nrm, alb, dep);
        return chk(hipGetLastError(), "aov launch");
    });
}

int launch_denoise(const rg_settings &st, uint32_t width, uint32_t height, const PTScene &, const Scene &,
                   float *d_hdr, void *d_aov, void *d_a, void *d_b, void *stream) {
    const int w = (int)width, h = (int)height, n = w * h;
    float4 *nrm = (float4 *)d_aov;
    float4 *alb = nrm + n;
    float *dep = (float *)(alb + n);
    float4 *hdr = (float4 *)d_hdr;
    float4 *a = (float4 *)d_a, *b = (float4 *)d_b;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_prepass, dim3((n + 255) / 256), dim3(256), 0, s, hdr, alb, a, w, h);
    const float sc = st.sigma_color, sn = st.sigma_normal, sd = st.sigma_depth;
    const float inv_c2 = 1.0f / (2.0f * sc * sc + 1e-8f);
    const int levels = st.denoise_levels < 1 ? 1 : (int)st.denoise_levels;
    const dim3 grid((w + kTile - 1) / kTile, (h + kTile - 1) / kTile);
    float4 *src = a, *dst = b;
    for (int l = 0; l < levels; l++) {
        const int step = 1 << l;
        const float dscale = 1.0f / (sd * (float)step + 1e-8f);
        switch (step) {
            case 1: hipLaunchKernelGGL(k_atrous_lds<1>, grid, dim3(256), 0, s, src, dst, nrm, dep, w, h, inv_c2, sn, dscale); break;
            case 2: hipLaunchKernelGGL(k_atrous_lds<2>, grid, dim3(256), 0, s, src, dst, nrm, dep, w, h, inv_c2, sn, dscale); break;
            case 4: hipLaunchKernelGGL(k_atrous_lds<4>, grid, dim3(256), 0, s, src, dst, nrm, dep, w, h, inv_c2, sn, dscale); break;
            case 8: hipLaunchKernelGGL(k_atrous_lds<8>, grid, dim3(256), 0, s, src, dst, nrm, dep, w, h, inv_c2, sn, dscale); break;
            default:
                hipLaunchKernelGGL(k_atrous_global, grid, dim3(256), 0, s, src, dst, nrm, dep, w, h, step, inv_c2, sn,
                                   dscale);
        }
        float4 *t = src;
        src = dst;
        dst = t;
    }
    const float t = fminf(fmaxf(st.denoise_intensity, 0.0f), 1.0f);
    hipLaunchKernelGGL(k_remod, dim3((n + 255) / 256), dim3(256), 0, s, src, alb, hdr, n, t);
    return chk(hipGetLastError(), "denoise launch");
}

}  // namespace rg

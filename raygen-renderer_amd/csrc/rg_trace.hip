// rg_trace.hip — closest-hit / any-hit BVH traversal and ray generation
// kernels for gfx950.
//
// Arithmetic contract: built with -ffp-contract=off and correctly rounded
// fp32 division, so every box and triangle test performs exactly the binary32
// operations of the reference (src/raygen/bvh.cpp:233-285, bvh.h:63-101,
// raycommon.cpp:151-235) and the hits are bit-identical to the CPU
// restatement.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rg_device.h"
#include "rg_internal.h"

namespace rg {

// ---------------------------------------------------------------------------
// HIP plumbing used by the host side
// ---------------------------------------------------------------------------
static int hip_check(hipError_t e, const char *what) {
    if (e == hipSuccess) return RG_OK;
    return set_error(RG_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

int dev_set_device(int device) { return hip_check(hipSetDevice(device), "hipSetDevice"); }

// Drops a pending error of a call whose failure was already handled, so a
// later launch check does not report it.
void dev_clear_error() { (void)hipGetLastError(); }

int dev_malloc(void **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return RG_OK;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_error(RG_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    return RG_OK;
}

void dev_free(void *p) {
    if (p) (void)hipFree(p);
}

int dev_upload(void *dst, const void *src, size_t bytes, void *stream) {
    if (bytes == 0) return RG_OK;
    int rc = hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream),
                       "hipMemcpyAsync(H2D)");
    if (rc) return rc;
    return hip_check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

int dev_download(void *dst, const void *src, size_t bytes, void *stream) {
    if (bytes == 0) return RG_OK;
    int rc = hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream),
                       "hipMemcpyAsync(D2H)");
    if (rc) return rc;
    return hip_check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

int dev_stream_create(void **s) {
    hipStream_t st = nullptr;
    int rc = hip_check(hipStreamCreate(&st), "hipStreamCreate");
    *s = rc ? nullptr : (void *)st;
    return rc;
}

void dev_stream_destroy(void *s) {
    if (s) (void)hipStreamDestroy((hipStream_t)s);
}

int dev_sync(void *stream) {
    return hip_check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

constexpr int kBlock = 128;

// One thread per ray.
template <int kStack, int kW>
__global__ __launch_bounds__(kBlock) void k_closest(DevBvh bv, const rg_ray *__restrict__ rays,
                                                    rg_hit *__restrict__ hits, uint64_t n) {
    RG_LDS_COLUMN(lc, kBlock);
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 r0 = reinterpret_cast<const float4 *>(rays)[2 * i];
    const float4 r1 = reinterpret_cast<const float4 *>(rays)[2 * i + 1];
    float t = r1.w, bu = 0.0f, bvv = 0.0f;
    uint32_t prim = 0xFFFFFFFFu;
    trace_closest_dev<kStack, kBlock, kW>(bv, r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, r0.w, t, prim, bu, bvv, lc);
    reinterpret_cast<float4 *>(hits)[i] = make_float4(__uint_as_float(prim), t, bu, bvv);
}

template <int kStack, int kW>
__global__ __launch_bounds__(kBlock) void k_any(DevBvh bv, const rg_ray *__restrict__ rays,
                                                uint8_t *__restrict__ occ, uint64_t n, uint32_t pred_bit) {
    RG_LDS_COLUMN(lc, kBlock);
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 r0 = reinterpret_cast<const float4 *>(rays)[2 * i];
    const float4 r1 = reinterpret_cast<const float4 *>(rays)[2 * i + 1];
    occ[i] = trace_any_dev<kStack, kBlock, kW>(bv, r0.x, r0.y, r0.z, r1.x, r1.y, r1.z, r1.w, pred_bit, lc) ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_gen_rays(float cx, float cy, float cz, float ex, float ey,
                                                  float ez, uint32_t key, uint64_t first,
                                                  rg_ray *__restrict__ rays, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint64_t idx = first + i;
    uint32_t a[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), 0u, 0u};
    uint32_t b[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), 1u, 0u};
    philox(a, key, 0u);
    philox(b, key, 0u);
    const float ox = cx + (u01(a[0]) - 0.5f) * ex;
    const float oy = cy + (u01(a[1]) - 0.5f) * ey;
    const float oz = cz + (u01(a[2]) - 0.5f) * ez;
    const float z = 1.0f - 2.0f * u01(a[3]);
    const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    const float phi = 6.28318530717958647692f * u01(b[0]);
    float sp, cp;
    sincosf(phi, &sp, &cp);
    float4 *out = reinterpret_cast<float4 *>(rays) + 2 * i;
    out[0] = make_float4(ox, oy, oz, 0.0f);
    out[1] = make_float4(r * cp, r * sp, z, __int_as_float(0x7F800000));
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
DevBvh dev_bvh(const Scene &s) {
    DevBvh b;
    b.nodes = s.d_nodes4;
    b.width = s.width;
    b.tris = (const float4 *)s.d_tris;
    if (!s.nodes2.empty()) {
        for (int k = 0; k < 3; k++) { b.root.lo[k] = s.nodes2[0].bmin[k]; b.root.hi[k] = s.nodes2[0].bmax[k]; }
    } else {
        for (int k = 0; k < 3; k++) { b.root.lo[k] = 0.0f; b.root.hi[k] = 0.0f; }
    }
    b.has_tree = (s.nodes4.empty() && s.nodes8.empty()) ? 0 : 1;
    return b;
}

int launch_closest(const Scene &s, const rg_ray *rays, rg_hit *hits, uint64_t n, void *stream) {
    if (n == 0) return RG_OK;
    const uint64_t blocks = (n + kBlock - 1) / kBlock;
    const DevBvh bv = dev_bvh(s);
    return with_traversal(s.stack_need, s.width, [&](auto tag) {
        using T = decltype(tag);
        hipLaunchKernelGGL((k_closest<T::kStack, T::kW>), dim3((uint32_t)blocks), dim3(kBlock), 0,
                           (hipStream_t)stream, bv, rays, hits, n);
        return hip_check(hipGetLastError(), "k_closest launch");
    });
}

int launch_any(const Scene &s, const rg_ray *rays, uint8_t *occ, uint64_t n, int pred, void *stream) {
    if (n == 0) return RG_OK;
    const uint64_t blocks = (n + kBlock - 1) / kBlock;
    const DevBvh bv = dev_bvh(s);
    const uint32_t bit = 1u << pred;
    return with_traversal(s.stack_need, s.width, [&](auto tag) {
        using T = decltype(tag);
        hipLaunchKernelGGL((k_any<T::kStack, T::kW>), dim3((uint32_t)blocks), dim3(kBlock), 0,
                           (hipStream_t)stream, bv, rays, occ, n, bit);
        return hip_check(hipGetLastError(), "k_any launch");
    });
}

int launch_gen_rays(const float box[6], float scale, uint32_t key, uint64_t first, rg_ray *rays,
                    uint64_t n, void *stream) {
    if (n == 0) return RG_OK;
    const float cx = (box[0] + box[3]) * 0.5f, cy = (box[1] + box[4]) * 0.5f,
                cz = (box[2] + box[5]) * 0.5f;
    const float ex = (box[3] - box[0]) * scale, ey = (box[4] - box[1]) * scale,
                ez = (box[5] - box[2]) * scale;
    const uint64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_gen_rays, dim3((uint32_t)blocks), dim3(256), 0, (hipStream_t)stream, cx, cy,
                       cz, ex, ey, ez, key, first, rays, n);
    return hip_check(hipGetLastError(), "k_gen_rays launch");
}

}  // namespace rg

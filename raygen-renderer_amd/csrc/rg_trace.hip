// rg_trace.hip — closest-hit / any-hit BVH traversal and ray generation
// kernels for gfx950.
//
// Arithmetic contract: built with -ffp-contract=off and correctly rounded
// fp32 division, so every box and triangle test performs exactly the binary32
// operations of the reference (src/raygen/bvh.cpp:233-285, bvh.h:63-101,
// raycommon.cpp:151-235) and the hits are bit-identical to the CPU
// restatement.
#include <hip/hip_runtime.h>

#include <algorithm>

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

}  // namespace rg

extern "C" {

int rg_host_alloc(uint64_t bytes, void **out) {
    if (!out) return rg::set_error(RG_E_INVALID, "null argument");
    *out = nullptr;
    if (bytes == 0) return RG_OK;
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess || !p) {
        (void)hipGetLastError();
        return rg::set_error(RG_E_NOMEM, "hipHostMalloc failed");
    }
    *out = p;
    return RG_OK;
}

void rg_host_free(void *p) {
    if (p) (void)hipHostFree(p);
}

}  // extern "C"

namespace rg {

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

// Closest / any hit of a ray array (rg_trace_closest / rg_trace_any).
struct RayArraySrc {
    const rg_ray *rays;
    rg_hit *hits;
    uint8_t *occ;
    uint32_t pred_bit;
    unsigned long long *work;  // zeroed work-sharing counter, or null
    static constexpr uint32_t kStaticQuarters = 1, kGrab = 128;
    __device__ __forceinline__ bool load(uint64_t i, const DevBvh &, PsRay &r) const {
        const float4 r0 = reinterpret_cast<const float4 *>(rays)[2 * i];
        const float4 r1 = reinterpret_cast<const float4 *>(rays)[2 * i + 1];
        r.ox = r0.x; r.oy = r0.y; r.oz = r0.z; r.tmin = r0.w;
        r.dx = r1.x; r.dy = r1.y; r.dz = r1.z; r.tmax = r1.w;
        r.pred_bit = pred_bit;
        return true;
    }
    __device__ __forceinline__ void store_closest(uint64_t i, uint32_t prim, float t, float u, float v) const {
        reinterpret_cast<float4 *>(hits)[i] = make_float4(__uint_as_float(prim), t, u, v);
    }
    __device__ __forceinline__ void store_any(uint64_t i, bool o) const { occ[i] = o ? 1 : 0; }
};

// Register budget of 7 waves per SIMD (72 VGPRs, a few spills): measured
// faster than the compiler's 78-VGPR, 6-wave allocation and than 8 waves.
template <int kStack, int kW, bool kAny>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(7))) void k_trace_ps(
    DevBvh bv, RayArraySrc src, uint64_t n) {
    RG_LDS_COLUMN(lc, kBlock);
    trace_ps_loop<kStack, kW, kAny, kBlock>(bv, src, n, lc);
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
    for (int k = 0; k < 3; k++) { b.root.lo[k] = s.root_box[k]; b.root.hi[k] = s.root_box[3 + k]; }
    b.has_tree = s.n_wide > 0 ? 1 : 0;
    const size_t n_nodes = s.n_wide;
    const size_t node_bytes = s.width == 8 ? sizeof(Node8) : sizeof(Node4);
    b.top = (uint32_t)std::min<size_t>(n_nodes, kTreeletBytes / node_bytes);
    // a tree held whole in LDS has no node fetch for idle lanes to wait on:
    // there the wave waits for every lane (measured on the 36-triangle cube
    // room, 0.93 against 0.96 ms per frame with batches of 16)
    b.leaf_batch = n_nodes <= b.top ? 64 : kLeafBatch;
    return b;
}

// Persistent launch: as many blocks as are resident at once.  With enough
// rays per wave they are shared through the caller's counter (a quarter in
// equal chunks, the rest in grabs of 128), which the launch clears first on
// the same stream; otherwise they are split evenly.
__global__ void k_clear_u64(unsigned long long *w) { *w = 0ull; }

template <int kStack, int kW, bool kAny>
static int launch_ps(const DevBvh &bv, const rg_ray *rays, rg_hit *hits, uint8_t *occ, uint64_t n, uint32_t bit,
                     unsigned long long *work, hipStream_t stream) {
    const uint32_t resident = resident_blocks<k_trace_ps<kStack, kW, kAny>>(kBlock);
    const uint64_t blocks = std::min<uint64_t>((uint64_t)resident, (n + kBlock - 1) / kBlock);
    if (n < kShareMinPerWave * (blocks * kBlock / 64)) work = nullptr;
    if (work) hipLaunchKernelGGL(k_clear_u64, dim3(1), dim3(1), 0, stream, work);
    const RayArraySrc src{rays, hits, occ, bit, work};
    hipLaunchKernelGGL((k_trace_ps<kStack, kW, kAny>), dim3((uint32_t)blocks), dim3(kBlock), 0, stream, bv, src, n);
    return hip_check(hipGetLastError(), "traversal launch");
}

static int trace_ps(const Scene &s, const rg_ray *rays, rg_hit *hits, uint8_t *occ, uint64_t n, uint32_t bit,
                    uint64_t *work, void *stream) {
    const DevBvh bv = dev_bvh(s);
    return with_traversal(s.stack_need, s.width, [&](auto tag) {
        using T = decltype(tag);
        unsigned long long *w = reinterpret_cast<unsigned long long *>(work);
        if (occ) return launch_ps<T::kStack, T::kW, true>(bv, rays, hits, occ, n, bit, w, (hipStream_t)stream);
        return launch_ps<T::kStack, T::kW, false>(bv, rays, hits, occ, n, bit, w, (hipStream_t)stream);
    });
}

int launch_closest(const Scene &s, const rg_ray *rays, rg_hit *hits, uint64_t n, uint64_t *work, void *stream) {
    if (n == 0) return RG_OK;
    return trace_ps(s, rays, hits, nullptr, n, 0, work, stream);
}

int launch_any(const Scene &s, const rg_ray *rays, uint8_t *occ, uint64_t n, int pred, uint64_t *work,
               void *stream) {
    if (n == 0) return RG_OK;
    return trace_ps(s, rays, nullptr, occ, n, 1u << pred, work, stream);
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

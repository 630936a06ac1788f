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

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------
struct RootBox {
    float lo[3], hi[3];
};

__device__ __forceinline__ float inv_dir(float d) {
    return (fabsf(d) > 1e-20f) ? 1.0f / d : 1e30f;
}

// Slab test of bvh.cpp:259-267; returns the entry distance, and whether the
// time-independent part (tmax >= 0, tmin <= tmax) and tmin <= tcur hold.
__device__ __forceinline__ bool slab(float lx, float ly, float lz, float hx, float hy, float hz,
                                     float ox, float oy, float oz, float ix, float iy, float iz,
                                     float tcur, float &tmin_out) {
    const float t1x = (lx - ox) * ix;
    const float t2x = (hx - ox) * ix;
    const float t1y = (ly - oy) * iy;
    const float t2y = (hy - oy) * iy;
    const float t1z = (lz - oz) * iz;
    const float t2z = (hz - oz) * iz;
    const float tmin = fmaxf(fmaxf(fminf(t1x, t2x), fminf(t1y, t2y)), fminf(t1z, t2z));
    const float tmax = fminf(fminf(fmaxf(t1x, t2x), fmaxf(t1y, t2y)), fmaxf(t1z, t2z));
    tmin_out = tmin;
    return !(tmax < 0.0f || tmin > tmax || tmin > tcur);
}

// Möller–Trumbore, closest variant (raycommon.cpp:200-235) on a precomputed
// edge record; e1 = v2 - v1 and e2 = v3 - v1 are the same binary32 values the
// reference forms on the fly.
__device__ __forceinline__ bool mt_closest(const float4 a, const float4 b, const float4 c,
                                           float ox, float oy, float oz, float dx, float dy,
                                           float dz, float tmin, float &t, float &u, float &v) {
    // pvec = cross(d, e2)
    const float px = dy * c.z - dz * c.y;
    const float py = dz * c.x - dx * c.z;
    const float pz = dx * c.y - dy * c.x;
    const float det = b.x * px + b.y * py + b.z * pz;
    if (fabsf(det) < 1e-8f) return false;
    const float inv_det = 1.0f / det;
    const float tx = ox - a.x, ty = oy - a.y, tz = oz - a.z;
    const float uu = (tx * px + ty * py + tz * pz) * inv_det;
    if (uu < 0.0f || uu > 1.0f) return false;
    // qvec = cross(tvec, e1)
    const float qx = ty * b.z - tz * b.y;
    const float qy = tz * b.x - tx * b.z;
    const float qz = tx * b.y - ty * b.x;
    const float vv = (dx * qx + dy * qy + dz * qz) * inv_det;
    if (vv < 0.0f || uu + vv > 1.0f) return false;
    const float tt = (c.x * qx + c.y * qy + c.z * qz) * inv_det;
    if (tt < tmin) return false;
    if (tt >= t) return false;
    t = tt;
    u = uu;
    v = vv;
    return true;
}

// Möller–Trumbore, any-hit variant (raycommon.cpp:151-198).
__device__ __forceinline__ bool mt_any(const float4 a, const float4 b, const float4 c, float ox,
                                       float oy, float oz, float dx, float dy, float dz,
                                       float maxt) {
    const float hx = dy * c.z - dz * c.y;
    const float hy = dz * c.x - dx * c.z;
    const float hz = dx * c.y - dy * c.x;
    const float aa = b.x * hx + b.y * hy + b.z * hz;
    if (fabsf(aa) < 1e-6f) return false;
    const float f = 1.0f / aa;
    const float sx = ox - a.x, sy = oy - a.y, sz = oz - a.z;
    const float u = f * (sx * hx + sy * hy + sz * hz);
    if (u < 0.0f || u > 1.0f) return false;
    const float qx = sy * b.z - sz * b.y;
    const float qy = sz * b.x - sx * b.z;
    const float qz = sx * b.y - sy * b.x;
    const float v = f * (dx * qx + dy * qy + dz * qz);
    if (v < 0.0f || u + v > 1.0f) return false;
    const float t = f * (c.x * qx + c.y * qy + c.z * qz);
    return t > 1e-6f && t < maxt;
}

constexpr int kBlock = 128;

// One thread per ray.  Entries on the stack carry the child's entry distance
// so the "tmin > closest t" rejection of the reference is applied when the
// entry is popped, exactly when the binary traversal would have tested it.
template <int kStack>
__global__ __launch_bounds__(kBlock) void k_closest(const Node4 *__restrict__ nodes,
                                                    const float4 *__restrict__ tris,
                                                    const rg_ray *__restrict__ rays,
                                                    rg_hit *__restrict__ hits, uint64_t n,
                                                    RootBox root, int has_tree) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 r0 = reinterpret_cast<const float4 *>(rays)[2 * i];
    const float4 r1 = reinterpret_cast<const float4 *>(rays)[2 * i + 1];
    const float ox = r0.x, oy = r0.y, oz = r0.z, rtmin = r0.w;
    const float dx = r1.x, dy = r1.y, dz = r1.z;
    float t = r1.w;
    uint32_t prim = 0xFFFFFFFFu;
    float bu = 0.0f, bv = 0.0f;

    const float ix = inv_dir(dx), iy = inv_dir(dy), iz = inv_dir(dz);
    const int oct = (ix < 0.0f ? 1 : 0) | (iy < 0.0f ? 2 : 0) | (iz < 0.0f ? 4 : 0);

    float root_tmin;
    if (has_tree && slab(root.lo[0], root.lo[1], root.lo[2], root.hi[0], root.hi[1], root.hi[2],
                         ox, oy, oz, ix, iy, iz, t, root_tmin)) {
        uint32_t stk_ref[kStack];
        float stk_t[kStack];
        int sp = 0;
        uint32_t cur = 0;
        for (;;) {
            if (!(cur & kLeafBit)) {
                const Node4 *nd = nodes + cur;
                const float4 lx = *reinterpret_cast<const float4 *>(nd->lox);
                const float4 ly = *reinterpret_cast<const float4 *>(nd->loy);
                const float4 lz = *reinterpret_cast<const float4 *>(nd->loz);
                const float4 hx = *reinterpret_cast<const float4 *>(nd->hix);
                const float4 hy = *reinterpret_cast<const float4 *>(nd->hiy);
                const float4 hz = *reinterpret_cast<const float4 *>(nd->hiz);
                const uint4 ch = *reinterpret_cast<const uint4 *>(nd->child);
                const uint2 ordw = *reinterpret_cast<const uint2 *>(nd->order);
                const uint32_t ord =
                    (oct < 4 ? (ordw.x >> (8 * oct)) : (ordw.y >> (8 * (oct - 4)))) & 0xFFu;
                float tm[4];
                bool ok[4];
                ok[0] = ch.x != kEmpty &&
                        slab(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, ox, oy, oz, ix, iy, iz, t, tm[0]);
                ok[1] = ch.y != kEmpty &&
                        slab(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, ox, oy, oz, ix, iy, iz, t, tm[1]);
                ok[2] = ch.z != kEmpty &&
                        slab(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, ox, oy, oz, ix, iy, iz, t, tm[2]);
                ok[3] = ch.w != kEmpty &&
                        slab(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, ox, oy, oz, ix, iy, iz, t, tm[3]);
                const uint32_t cref[4] = {ch.x, ch.y, ch.z, ch.w};
                // push in reverse visit order; the first one becomes `cur`
                bool have = false;
                uint32_t next = 0;
                float next_t = 0.0f;
#pragma unroll
                for (int j = 3; j >= 0; j--) {
                    const int s = (ord >> (2 * j)) & 3;
                    if (!ok[s]) continue;
                    if (have) {
                        stk_ref[sp] = next;
                        stk_t[sp] = next_t;
                        sp++;
                    }
                    next = cref[s];
                    next_t = tm[s];
                    have = true;
                }
                if (have) { cur = next; continue; }
            } else {
                uint32_t k = cur & ~kLeafBit;
                for (;;) {
                    const float4 a = tris[3 * (size_t)k];
                    const float4 b = tris[3 * (size_t)k + 1];
                    const float4 c = tris[3 * (size_t)k + 2];
                    if (mt_closest(a, b, c, ox, oy, oz, dx, dy, dz, rtmin, t, bu, bv))
                        prim = __float_as_uint(a.w);
                    if (__float_as_uint(b.w) & 1u) break;
                    k++;
                }
            }
            bool found = false;
            while (sp > 0) {
                sp--;
                if (!(stk_t[sp] > t)) { cur = stk_ref[sp]; found = true; break; }
            }
            if (!found) break;
        }
    }
    reinterpret_cast<float4 *>(hits)[i] = make_float4(__uint_as_float(prim), t, bu, bv);
}

// Any-hit traversal (bvh.h:63-101).  The boolean result does not depend on
// visit order, so children are visited in the same octant order as above.
template <int kStack>
__global__ __launch_bounds__(kBlock) void k_any(const Node4 *__restrict__ nodes,
                                                const float4 *__restrict__ tris,
                                                const rg_ray *__restrict__ rays,
                                                uint8_t *__restrict__ occ, uint64_t n,
                                                RootBox root, int has_tree, uint32_t pred_bit) {
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 r0 = reinterpret_cast<const float4 *>(rays)[2 * i];
    const float4 r1 = reinterpret_cast<const float4 *>(rays)[2 * i + 1];
    const float ox = r0.x, oy = r0.y, oz = r0.z;
    const float dx = r1.x, dy = r1.y, dz = r1.z;
    const float maxt = r1.w;
    const float ix = inv_dir(dx), iy = inv_dir(dy), iz = inv_dir(dz);
    const int oct = (ix < 0.0f ? 1 : 0) | (iy < 0.0f ? 2 : 0) | (iz < 0.0f ? 4 : 0);
    uint8_t hit = 0;
    float tm_unused;
    if (has_tree && slab(root.lo[0], root.lo[1], root.lo[2], root.hi[0], root.hi[1], root.hi[2],
                         ox, oy, oz, ix, iy, iz, maxt, tm_unused)) {
        uint32_t stk[kStack];
        int sp = 0;
        uint32_t cur = 0;
        for (;;) {
            if (!(cur & kLeafBit)) {
                const Node4 *nd = nodes + cur;
                const float4 lx = *reinterpret_cast<const float4 *>(nd->lox);
                const float4 ly = *reinterpret_cast<const float4 *>(nd->loy);
                const float4 lz = *reinterpret_cast<const float4 *>(nd->loz);
                const float4 hx = *reinterpret_cast<const float4 *>(nd->hix);
                const float4 hy = *reinterpret_cast<const float4 *>(nd->hiy);
                const float4 hz = *reinterpret_cast<const float4 *>(nd->hiz);
                const uint4 ch = *reinterpret_cast<const uint4 *>(nd->child);
                const uint2 ordw = *reinterpret_cast<const uint2 *>(nd->order);
                const uint32_t ord =
                    (oct < 4 ? (ordw.x >> (8 * oct)) : (ordw.y >> (8 * (oct - 4)))) & 0xFFu;
                float tm[4];
                bool ok[4];
                ok[0] = ch.x != kEmpty &&
                        slab(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, ox, oy, oz, ix, iy, iz, maxt, tm[0]);
                ok[1] = ch.y != kEmpty &&
                        slab(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, ox, oy, oz, ix, iy, iz, maxt, tm[1]);
                ok[2] = ch.z != kEmpty &&
                        slab(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, ox, oy, oz, ix, iy, iz, maxt, tm[2]);
                ok[3] = ch.w != kEmpty &&
                        slab(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, ox, oy, oz, ix, iy, iz, maxt, tm[3]);
                const uint32_t cref[4] = {ch.x, ch.y, ch.z, ch.w};
                bool have = false;
                uint32_t next = 0;
#pragma unroll
                for (int j = 3; j >= 0; j--) {
                    const int s = (ord >> (2 * j)) & 3;
                    if (!ok[s]) continue;
                    if (have) stk[sp++] = next;
                    next = cref[s];
                    have = true;
                }
                if (have) { cur = next; continue; }
            } else {
                uint32_t k = cur & ~kLeafBit;
                for (;;) {
                    const float4 a = tris[3 * (size_t)k];
                    const float4 b = tris[3 * (size_t)k + 1];
                    const float4 c = tris[3 * (size_t)k + 2];
                    if ((__float_as_uint(c.w) & pred_bit) &&
                        mt_any(a, b, c, ox, oy, oz, dx, dy, dz, maxt)) {
                        hit = 1;
                        break;
                    }
                    if (__float_as_uint(b.w) & 1u) break;
                    k++;
                }
                if (hit) break;
            }
            if (sp == 0) break;
            cur = stk[--sp];
        }
    }
    occ[i] = hit;
}

// ---------------------------------------------------------------------------
// Ray generation for the traversal microbenchmark: Philox4x32-10
// (Salmon et al., SC'11), counter = (index lo, index hi, block, 0).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint32_t lo0 = 0xD2511F53u * c[0];
        const uint32_t hi0 = __umulhi(0xD2511F53u, c[0]);
        const uint32_t lo1 = 0xCD9E8D57u * c[2];
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]);
        const uint32_t n0 = hi1 ^ c[1] ^ k0;
        const uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ __forceinline__ float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

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
static RootBox root_of(const Scene &s) {
    RootBox r;
    if (!s.nodes2.empty()) {
        for (int k = 0; k < 3; k++) { r.lo[k] = s.nodes2[0].bmin[k]; r.hi[k] = s.nodes2[0].bmax[k]; }
    } else {
        for (int k = 0; k < 3; k++) { r.lo[k] = 0.0f; r.hi[k] = 0.0f; }
    }
    return r;
}

int launch_closest(const Scene &s, const rg_ray *rays, rg_hit *hits, uint64_t n, void *stream) {
    if (n == 0) return RG_OK;
    const uint64_t blocks = (n + kBlock - 1) / kBlock;
    const int has = s.nodes4.empty() ? 0 : 1;
    if (s.stack_need <= 64)
        hipLaunchKernelGGL(k_closest<64>, dim3((uint32_t)blocks), dim3(kBlock), 0, (hipStream_t)stream,
                           (const Node4 *)s.d_nodes4, (const float4 *)s.d_tris, rays, hits, n, root_of(s), has);
    else if (s.stack_need <= 256)
        hipLaunchKernelGGL(k_closest<256>, dim3((uint32_t)blocks), dim3(kBlock), 0, (hipStream_t)stream,
                           (const Node4 *)s.d_nodes4, (const float4 *)s.d_tris, rays, hits, n, root_of(s), has);
    else
        return set_error(RG_E_UNSUPPORTED, "BVH too deep for the traversal stack");
    return hip_check(hipGetLastError(), "k_closest launch");
}

int launch_any(const Scene &s, const rg_ray *rays, uint8_t *occ, uint64_t n, int pred, void *stream) {
    if (n == 0) return RG_OK;
    const uint64_t blocks = (n + kBlock - 1) / kBlock;
    const int has = s.nodes4.empty() ? 0 : 1;
    const uint32_t bit = 1u << pred;
    if (s.stack_need <= 64)
        hipLaunchKernelGGL(k_any<64>, dim3((uint32_t)blocks), dim3(kBlock), 0, (hipStream_t)stream,
                           (const Node4 *)s.d_nodes4, (const float4 *)s.d_tris, rays, occ, n, root_of(s), has, bit);
    else if (s.stack_need <= 256)
        hipLaunchKernelGGL(k_any<256>, dim3((uint32_t)blocks), dim3(kBlock), 0, (hipStream_t)stream,
                           (const Node4 *)s.d_nodes4, (const float4 *)s.d_tris, rays, occ, n, root_of(s), has, bit);
    else
        return set_error(RG_E_UNSUPPORTED, "BVH too deep for the traversal stack");
    return hip_check(hipGetLastError(), "k_any launch");
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

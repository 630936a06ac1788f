// rg_device.h — device-side building blocks shared by the traversal and
// path-tracing kernels (gfx950).  Built with -ffp-contract=off so the box and
// triangle tests are the exact binary32 operations of the reference
// (src/raygen/bvh.cpp:233-285, bvh.h:63-101, raycommon.cpp:151-235).
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>

#include <cstdint>
#include <type_traits>

#include "rg_internal.h"
#include "rg_pt.h"

namespace rg {

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------

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


// One wide node's child tests: for each slot whether the ray enters the
// child box before tmax (and the entry distance), the child refs and the
// visit order for the ray's octant (kOrdBits per slot, first in the low bits).
template <int W>
struct WideTest;

template <>
struct WideTest<4> {
    static constexpr int kOrdBits = 2;
    __device__ __forceinline__ static void run(const void *nodes, uint32_t cur, int oct, float ox, float oy,
                                               float oz, float ix, float iy, float iz, float tmax, bool *ok,
                                               float *tm, uint32_t *cref, uint32_t &ord, uint32_t &rk) {
        const Node4 *nd = reinterpret_cast<const Node4 *>(nodes) + cur;
        const float4 lx = *reinterpret_cast<const float4 *>(nd->lox);
        const float4 ly = *reinterpret_cast<const float4 *>(nd->loy);
        const float4 lz = *reinterpret_cast<const float4 *>(nd->loz);
        const float4 hx = *reinterpret_cast<const float4 *>(nd->hix);
        const float4 hy = *reinterpret_cast<const float4 *>(nd->hiy);
        const float4 hz = *reinterpret_cast<const float4 *>(nd->hiz);
        const uint4 ch = *reinterpret_cast<const uint4 *>(nd->child);
        // the octant's order byte by a computed address: no branch, and the
        // load issues with the box loads
        const uint32_t ow = reinterpret_cast<const uint32_t *>(nd->order)[oct >> 2];
        ord = (ow >> (8 * (oct & 3))) & 0xFFu;
        const uint32_t rw = reinterpret_cast<const uint32_t *>(nd->rank)[oct >> 2];
        rk = (rw >> (8 * (oct & 3))) & 0xFFu;
        cref[0] = ch.x; cref[1] = ch.y; cref[2] = ch.z; cref[3] = ch.w;
        // every slot's slab test runs (an empty slot's zero box is masked
        // off): straight-line code instead of a branch per child
        ok[0] = slab(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, ox, oy, oz, ix, iy, iz, tmax, tm[0]) & (ch.x != kEmpty);
        ok[1] = slab(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, ox, oy, oz, ix, iy, iz, tmax, tm[1]) & (ch.y != kEmpty);
        ok[2] = slab(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, ox, oy, oz, ix, iy, iz, tmax, tm[2]) & (ch.z != kEmpty);
        ok[3] = slab(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, ox, oy, oz, ix, iy, iz, tmax, tm[3]) & (ch.w != kEmpty);
    }
};

template <>
struct WideTest<8> {
    static constexpr int kOrdBits = 3;
    __device__ __forceinline__ static void run(const void *nodes, uint32_t cur, int oct, float ox, float oy,
                                               float oz, float ix, float iy, float iz, float tmax, bool *ok,
                                               float *tm, uint32_t *cref, uint32_t &ord, uint32_t &rk) {
        const Node8 *nd = reinterpret_cast<const Node8 *>(nodes) + cur;
        ord = nd->order[oct];
        rk = 0;  // each slot's visit position
#pragma unroll
        for (int j = 0; j < 8; j++) rk |= (uint32_t)j << (3 * ((ord >> (3 * j)) & 7));
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const float4 lx = *reinterpret_cast<const float4 *>(nd->lox + 4 * h);
            const float4 ly = *reinterpret_cast<const float4 *>(nd->loy + 4 * h);
            const float4 lz = *reinterpret_cast<const float4 *>(nd->loz + 4 * h);
            const float4 hx = *reinterpret_cast<const float4 *>(nd->hix + 4 * h);
            const float4 hy = *reinterpret_cast<const float4 *>(nd->hiy + 4 * h);
            const float4 hz = *reinterpret_cast<const float4 *>(nd->hiz + 4 * h);
            const uint4 ch = *reinterpret_cast<const uint4 *>(nd->child + 4 * h);
            const int b = 4 * h;
            cref[b] = ch.x; cref[b + 1] = ch.y; cref[b + 2] = ch.z; cref[b + 3] = ch.w;
            ok[b] = slab(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, ox, oy, oz, ix, iy, iz, tmax, tm[b]) & (ch.x != kEmpty);
            ok[b + 1] =
                slab(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, ox, oy, oz, ix, iy, iz, tmax, tm[b + 1]) & (ch.y != kEmpty);
            ok[b + 2] =
                slab(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, ox, oy, oz, ix, iy, iz, tmax, tm[b + 2]) & (ch.z != kEmpty);
            ok[b + 3] =
                slab(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, ox, oy, oz, ix, iy, iz, tmax, tm[b + 3]) & (ch.w != kEmpty);
        }
    }
};

// The top of each lane's traversal stack lives in LDS, column-major across
// the block (entry e of thread x at [e * kStride + x]: consecutive lanes hit
// consecutive banks); entries deeper than kLdsStack fall back to private
// (scratch) memory.  Keeping the hot part out of scratch removes most of the
// traversal's HBM writes.
#ifndef RG_LDS_STACK
#define RG_LDS_STACK 8
#endif
constexpr int kLdsStack = RG_LDS_STACK;

constexpr int kLdsWords = kLdsStack > 0 ? kLdsStack : 1;

template <int kStride>
struct LdsColumn {
    uint32_t *ref;  // this lane's entry 0; entry e at ref[e * kStride]
    float *t;       // entry distances (closest hit), or more refs (any hit)
};

// Declares the block's LDS stack area and this lane's column.
#define RG_LDS_COLUMN(name, kBlk)                                  \
    __shared__ uint32_t name##_ref[kLdsWords * (kBlk)];            \
    __shared__ float name##_t[kLdsWords * (kBlk)];                 \
    const LdsColumn<(kBlk)> name{name##_ref + threadIdx.x, name##_t + threadIdx.x}

// Closest-hit traversal of one ray (TriangleBVH::intersectClosest,
// bvh.cpp:233-285) over the W-wide tree.  Entries on the stack carry the
// child's entry distance so the reference's "tmin > closest t" rejection is
// applied when the entry is popped, exactly when the binary traversal would
// have tested that node.  On return t/prim/u/v hold the closest hit (prim =
// 0xFFFFFFFF and t unchanged on a miss).
template <int kStack, int kStride, int W>
__device__ __forceinline__ void trace_closest_dev(const DevBvh &bv, float ox, float oy, float oz, float dx,
                                                float dy, float dz, float rtmin, float &t, uint32_t &prim,
                                                float &bu, float &bv_, const LdsColumn<kStride> col) {
    const float ix = inv_dir(dx), iy = inv_dir(dy), iz = inv_dir(dz);
    const int oct = (ix < 0.0f ? 1 : 0) | (iy < 0.0f ? 2 : 0) | (iz < 0.0f ? 4 : 0);
    float root_tmin;
    if (!bv.has_tree || !slab(bv.root.lo[0], bv.root.lo[1], bv.root.lo[2], bv.root.hi[0], bv.root.hi[1],
                              bv.root.hi[2], ox, oy, oz, ix, iy, iz, t, root_tmin))
        return;
    constexpr int kSpill = kStack > kLdsStack ? kStack - kLdsStack : 1;
    constexpr int kOrdBits = WideTest<W>::kOrdBits;
    uint32_t stk_ref[kSpill];
    float stk_t[kSpill];
    int sp = 0;
    uint32_t cur = 0;
    for (;;) {
        if (!(cur & kLeafBit)) {
            bool ok[W];
            float tm[W];
            uint32_t cref[W], ord, rk;
            WideTest<W>::run(bv.nodes, cur, oct, ox, oy, oz, ix, iy, iz, t, ok, tm, cref, ord, rk);
            bool have = false;
            uint32_t next = 0;
            float next_t = 0.0f;
#pragma unroll
            for (int j = W - 1; j >= 0; j--) {
                const int s = (ord >> (kOrdBits * j)) & (W - 1);
                if (!ok[s]) continue;
                if (have) {
                    if (sp < kLdsStack) {
                        col.ref[sp * kStride] = next;
                        col.t[sp * kStride] = next_t;
                    } else {
                        stk_ref[sp - kLdsStack] = next;
                        stk_t[sp - kLdsStack] = next_t;
                    }
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
                const float4 a = bv.tris[3 * (size_t)k];
                const float4 b = bv.tris[3 * (size_t)k + 1];
                const float4 c = bv.tris[3 * (size_t)k + 2];
                if (mt_closest(a, b, c, ox, oy, oz, dx, dy, dz, rtmin, t, bu, bv_)) prim = __float_as_uint(a.w);
                if (__float_as_uint(b.w) & 1u) break;
                k++;
            }
        }
        bool found = false;
        while (sp > 0) {
            sp--;
            const bool in_lds = sp < kLdsStack;
            const float et = in_lds ? col.t[sp * kStride] : stk_t[sp - kLdsStack];
            if (!(et > t)) {
                cur = in_lds ? col.ref[sp * kStride] : stk_ref[sp - kLdsStack];
                found = true;
                break;
            }
        }
        if (!found) break;
    }
}

// Any-hit traversal (TriangleBVH::intersectAny, bvh.h:63-101) with maxT =
// tmax; a triangle occludes when its occluder bits include pred_bit.  The
// boolean result does not depend on visit order.
template <int kStack, int kStride, int W>
__device__ __forceinline__ bool trace_any_dev(const DevBvh &bv, float ox, float oy, float oz, float dx, float dy,
                                            float dz, float maxt, uint32_t pred_bit,
                                            const LdsColumn<kStride> col) {
    // entries are refs only: the column's distance words hold more of them
    constexpr int kLds = 2 * kLdsStack;
    constexpr int kOrdBits = WideTest<W>::kOrdBits;
    uint32_t *const lt = reinterpret_cast<uint32_t *>(col.t);
    const float ix = inv_dir(dx), iy = inv_dir(dy), iz = inv_dir(dz);
    const int oct = (ix < 0.0f ? 1 : 0) | (iy < 0.0f ? 2 : 0) | (iz < 0.0f ? 4 : 0);
    float tm_unused;
    if (!bv.has_tree || !slab(bv.root.lo[0], bv.root.lo[1], bv.root.lo[2], bv.root.hi[0], bv.root.hi[1],
                              bv.root.hi[2], ox, oy, oz, ix, iy, iz, maxt, tm_unused))
        return false;
    constexpr int kSpill = kStack > kLds ? kStack - kLds : 1;
    uint32_t stk[kSpill];
    int sp = 0;
    uint32_t cur = 0;
    for (;;) {
        if (!(cur & kLeafBit)) {
            bool ok[W];
            float tm[W];
            uint32_t cref[W], ord, rk;
            WideTest<W>::run(bv.nodes, cur, oct, ox, oy, oz, ix, iy, iz, maxt, ok, tm, cref, ord, rk);
            bool have = false;
            uint32_t next = 0;
#pragma unroll
            for (int j = W - 1; j >= 0; j--) {
                const int s = (ord >> (kOrdBits * j)) & (W - 1);
                if (!ok[s]) continue;
                if (have) {
                    if (sp < kLdsStack) col.ref[sp * kStride] = next;
                    else if (sp < kLds) lt[(sp - kLdsStack) * kStride] = next;
                    else stk[sp - kLds] = next;
                    sp++;
                }
                next = cref[s];
                have = true;
            }
            if (have) { cur = next; continue; }
        } else {
            uint32_t k = cur & ~kLeafBit;
            for (;;) {
                const float4 a = bv.tris[3 * (size_t)k];
                const float4 b = bv.tris[3 * (size_t)k + 1];
                const float4 c = bv.tris[3 * (size_t)k + 2];
                if ((__float_as_uint(c.w) & pred_bit) && mt_any(a, b, c, ox, oy, oz, dx, dy, dz, maxt))
                    return true;
                if (__float_as_uint(b.w) & 1u) break;
                k++;
            }
        }
        if (sp == 0) break;
        sp--;
        cur = sp < kLdsStack ? col.ref[sp * kStride] : (sp < kLds ? lt[(sp - kLdsStack) * kStride] : stk[sp - kLds]);
    }
    return false;
}

// One ray of a persistent traversal lane.
struct PsRay {
    float ox, oy, oz, dx, dy, dz, tmin, tmax;
    uint32_t pred_bit;
};

constexpr uint32_t kNoNode = 0xFFFFFFFEu;

// Lanes of a wave that must hold a postponed leaf before the wave intersects
// leaves (unless every lane is ready sooner).  Waiting for all 64 left the
// lanes that had found a leaf idle behind the slowest one: measured on the
// 16M-ray closest-hit benchmark 3.37 ms at 64, 3.43 at 48, 3.18 at 40, 2.62 at
// 32, 2.41 at 24, 2.39 at 16, 2.52 at 8 (any hit 3.00 -> 2.11 ms at 16).
constexpr int kLeafBatch = 16;

// Persistent traversal with leaf postponement and lane refill (after Aila
// & Laine, HPG 2009), shared by every kernel that traces a batch of rays.
// Each wave owns a contiguous chunk of the n ray indices (and, with work
// sharing, grabs more); a lane whose ray is done takes the next index of the
// chunk, so a wave does not idle behind its
// longest ray.  Inner nodes are traversed until every lane of the wave holds
// a leaf to intersect, or kLeafBatch lanes do (the first leaf a lane reaches
// is postponed and the lane keeps descending with its current t), then the
// lanes holding one intersect their leaves together; the others carry on
// descending after that.
//
// The hits are those of trace_closest_dev, exact-t ties included, whatever
// rays share the wave: descending with the stale, larger t tests boxes
// against a t that the plain traversal would already have lowered, so after
// every leaf intersection the node the lane is about to enter is tested
// again against the new t (its entry distance is kept beside it).  Node
// boxes nest exactly and the slab test is monotone in the box bounds, so a
// node's entry distance is at least its ancestors': if it passes, every
// node on its path passes as well; the stack entries pushed meanwhile are
// re-tested when popped.  Relying instead on "a triangle inside a culled
// box cannot beat t" fails where the triangle test's t for a hit on a box
// face rounds below the slab's entry distance (a hit on an edge shared by
// triangles in two leaves).
//
// The tree's top levels (the first bv.top nodes, breadth-first) are copied
// into LDS when the block starts: every ray passes through them, and LDS
// answers in a fraction of an L2 hit's time.
//
// Src provides: bool load(i, bv, PsRay &) (false: nothing to trace at i),
// store_closest(i, prim, t, u, v), store_any(i, occluded).  Every thread of
// the block must call it (it synchronises the block once).
//
// Work sharing: a Src with a member `work` (pointer to a counter that is zero
// when the kernel starts, 32- or 64-bit) and the constants kStaticQuarters
// and kGrab hands each wave an equal chunk of the first kStaticQuarters / 4
// of the rays, then the rest in grabs of kGrab indices taken from the
// counter, so waves whose rays run short take over the long tail of waves
// whose rays run long.  The counter ends above n by at most kGrab per wave.
// Without `work`, or with fewer than kShareMinPerWave rays per wave, the n
// rays are split into equal chunks only: the counter is one address for the
// whole device, and a burst of read-modify-writes on it from every wave
// costs more than a short launch's imbalance.
constexpr uint64_t kShareMinPerWave = 256;
template <class S, class = void>
struct WorkShare {
    using T = unsigned long long;
    static constexpr uint32_t kStaticQuarters = 4, kGrab = 1;
    __device__ static T *get(const S &) { return nullptr; }
};
template <class S>
struct WorkShare<S, std::void_t<decltype(&S::work)>> {
    using T = std::remove_pointer_t<decltype(S::work)>;
    static constexpr uint32_t kStaticQuarters = S::kStaticQuarters, kGrab = S::kGrab;
    static_assert(kStaticQuarters <= 4 && kGrab >= 64, "static share at most all rays, grabs of a wave or more");
    __device__ static T *get(const S &s) { return s.work; }
};

template <int kStack, int kW, bool kAny, int kStride, class Src>
__device__ __forceinline__ void trace_ps_loop(const DevBvh &bv, const Src &src, uint64_t n,
                                              const LdsColumn<kStride> lc) {
    constexpr int kSpill = kStack > kLdsStack ? kStack - kLdsStack : 1;
    constexpr int kOrdBits = WideTest<kW>::kOrdBits;
    __shared__ uint4 treelet[kTreeletBytes / 16];
    {
        const uint32_t words = bv.top * (uint32_t)(kW == 8 ? sizeof(Node8) : sizeof(Node4)) / 16;
        const uint4 *src_words = reinterpret_cast<const uint4 *>(bv.nodes);
        for (uint32_t w = threadIdx.x; w < words; w += blockDim.x) treelet[w] = src_words[w];
        __syncthreads();
    }
    uint32_t stk_ref[kSpill];
    float stk_t[kSpill];
    const uint32_t lane = __lane_id();
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x >> 6);
    // the wave's index as a wave-uniform value, so that the chunk bounds
    // and the cursor stay in scalar registers
    const uint64_t wave = (uint64_t)blockIdx.x * (blockDim.x >> 6) +
                          (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    using Share = WorkShare<Src>;
    typename Share::T *const work = n >= kShareMinPerWave * waves ? Share::get(src) : nullptr;
    const uint64_t chunk = work ? n / 4 * Share::kStaticQuarters / waves : (n + waves - 1) / waves;
    const uint64_t static_end = work ? chunk * waves : n;  // <= n: grabs start here
    uint64_t cursor = wave * chunk;
    uint64_t chunk_end = cursor + chunk < n ? cursor + chunk : n;
    bool exhausted = work == nullptr;
    bool active = false;
    uint64_t ri = 0;
    PsRay r{};
    float ix = 0, iy = 0, iz = 0, t = 0, bu = 0, bvv = 0;
    int oct = 0, sp = 0;
    uint32_t prim = 0xFFFFFFFFu, cur = kNoNode, leaf = kNoNode;
    float cur_t = 0.0f;  // entry distance of cur (closest hit only)
    bool hit_any = false;

    // The LDS slot is read unconditionally (clamped to a valid one) and the
    // private one only past the LDS part.  Left alone, the compiler turns
    // the pair into one flat load from a selected address, and a flat load
    // waits for every load in flight, node fetches included; the empty asm
    // statements (no instruction, a register operand only) keep the LDS
    // reads as they are.
    auto pop = [&]() -> uint32_t {
        while (sp > 0) {
            sp--;
            const bool in_lds = sp < kLdsStack;
            const int li = in_lds ? sp : kLdsStack - 1;
            uint32_t ref = lc.ref[li * kStride];
            float et = kAny ? 0.0f : lc.t[li * kStride];
            asm volatile("" : "+v"(ref));
            if (!kAny) asm volatile("" : "+v"(et));
            if (!in_lds) {
                ref = stk_ref[sp - kLdsStack];
                if (!kAny) et = stk_t[sp - kLdsStack];
            }
            if (kAny) return ref;
            if (!(et > t)) {
                cur_t = et;
                return ref;
            }
        }
        return kNoNode;
    };
    auto finish = [&]() {
        if (kAny) src.store_any(ri, hit_any);
        else src.store_closest(ri, prim, t, bu, bvv);
        active = false;
    };

    for (;;) {
        // refill idle lanes from the wave's chunk or grab
        const uint64_t idle = __ballot(!active);
        if (idle && cursor >= chunk_end && !exhausted) {
            // the wave's chunk is used up: grab the next kGrab indices
            typename Share::T base = 0;
            if (lane == 0) base = atomicAdd(work, (typename Share::T)Share::kGrab);
            uint64_t got = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)base, 0);
            if (sizeof(base) == 8) got |= (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)base >> 32), 0) << 32;
            const uint64_t b = static_end + got;
            if (b < n) {
                cursor = b;
                chunk_end = b + Share::kGrab < n ? b + Share::kGrab : n;
            } else {
                exhausted = true;
            }
        }
        if (idle && cursor < chunk_end) {
            const uint64_t idx = cursor + (uint64_t)__popcll(idle & ((1ull << lane) - 1ull));
            if (!active && idx < chunk_end) {
                ri = idx;
                if (src.load(idx, bv, r)) {
                    t = r.tmax;
                    ix = inv_dir(r.dx); iy = inv_dir(r.dy); iz = inv_dir(r.dz);
                    oct = (ix < 0.0f ? 1 : 0) | (iy < 0.0f ? 2 : 0) | (iz < 0.0f ? 4 : 0);
                    prim = 0xFFFFFFFFu; bu = 0.0f; bvv = 0.0f; hit_any = false;
                    sp = 0; leaf = kNoNode;
                    float tm;
                    active = true;
                    if (bv.has_tree && slab(bv.root.lo[0], bv.root.lo[1], bv.root.lo[2], bv.root.hi[0],
                                            bv.root.hi[1], bv.root.hi[2], r.ox, r.oy, r.oz, ix, iy, iz, t, tm)) {
                        cur = 0;
                        cur_t = tm;
                    } else {
                        finish();
                    }
                }
            }
            cursor += (uint64_t)__popcll(idle);
        }
        if (__ballot(active) == 0) {
            if (cursor >= chunk_end && exhausted) break;
            continue;
        }
        // inner phase
        for (;;) {
            if (active && leaf == kNoNode && cur != kNoNode && (cur & kLeafBit)) {
                leaf = cur;
                cur = pop();
            }
            const bool want = active && cur != kNoNode && !(cur & kLeafBit);
            const bool ready = !active || leaf != kNoNode || cur == kNoNode;
            // to the leaf phase when no lane can descend, every lane is
            // ready, or bv.leaf_batch lanes (kLeafBatch, or 64 for a tree
            // held whole in LDS) hold a leaf (at least one, so every break
            // leaves the leaf phase work to do)
            const int holding = __popcll(__ballot(active && leaf != kNoNode));
            if (!__any(want) || __all(ready) || holding >= bv.leaf_batch) break;
            if (want) {
                bool ok[kW];
                float tm[kW];
                uint32_t cref[kW], ord, rk;
                // the top levels from LDS, the rest from HBM (a flat load
                // reaches either)
                const void *base = cur < bv.top ? (const void *)treelet : bv.nodes;
                WideTest<kW>::run(base, cur, oct, r.ox, r.oy, r.oz, ix, iy, iz, t, ok, tm, cref, ord, rk);
                // Each passing child's rank in the octant's visit order
                // fixes its stack position: the first is entered, the k-th
                // (k >= 1) of c goes to sp + c - 1 - k, so the second ends on
                // top, as pushing them one by one in reverse order leaves
                // them.  Slots that push nothing write to position sp + c - 1
                // (the new top, free), which keeps the writes unconditional.
                uint32_t pos[kW], m = 0;
#pragma unroll
                for (int i = 0; i < kW; i++) {
                    pos[i] = (rk >> (kOrdBits * i)) & (kW - 1);
                    m |= (uint32_t)ok[i] << pos[i];
                }
                if (m == 0) {
                    cur = pop();
                } else {
                    const int top = sp + __popc(m) - 1;  // stack size after the pushes
                    int e[kW];
                    uint32_t next = 0;
                    float next_t = 0.0f;
#pragma unroll
                    for (int i = 0; i < kW; i++) {
                        const int rank = __popc(m & ((1u << pos[i]) - 1u));
                        if (ok[i] && rank == 0) {
                            next = cref[i];
                            next_t = tm[i];
                        }
                        e[i] = ok[i] && rank > 0 ? top - rank : top;
                    }
                    // positions up to top stay within the stack: the tree's
                    // deepest push sequence sized it (stack_need)
                    if (top < kLdsStack) {
#pragma unroll
                        for (int i = 0; i < kW; i++) {
                            lc.ref[e[i] * kStride] = cref[i];
                            if (!kAny) lc.t[e[i] * kStride] = tm[i];
                        }
                    } else {
#pragma unroll
                        for (int i = 0; i < kW; i++) {
                            if (e[i] < kLdsStack) {
                                lc.ref[e[i] * kStride] = cref[i];
                                if (!kAny) lc.t[e[i] * kStride] = tm[i];
                            } else {
                                stk_ref[e[i] - kLdsStack] = cref[i];
                                if (!kAny) stk_t[e[i] - kLdsStack] = tm[i];
                            }
                        }
                    }
                    sp = top;
                    cur = next;
                    cur_t = next_t;
                }
            }
        }
        // leaf phase
        if (active && leaf != kNoNode) {
            uint32_t k = leaf & ~kLeafBit;
            for (;;) {
                float4 a = bv.tris[3 * (size_t)k];
                const float4 b = bv.tris[3 * (size_t)k + 1];
                float4 c = bv.tris[3 * (size_t)k + 2];
                // all 48 bytes as three aligned 16-byte loads: left to
                // itself the compiler reads only the words each variant
                // uses, as 12 + 16 + 16 bytes that can straddle cache lines
                asm volatile("" : "+v"(a.w));
                asm volatile("" : "+v"(c.w));
                if (kAny) {
                    if ((__float_as_uint(c.w) & r.pred_bit) &&
                        mt_any(a, b, c, r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, t)) {
                        hit_any = true;
                        break;
                    }
                } else if (mt_closest(a, b, c, r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, r.tmin, t, bu, bvv)) {
                    prim = __float_as_uint(a.w);
                }
                if (__float_as_uint(b.w) & 1u) break;
                k++;
            }
            leaf = kNoNode;
            if (kAny && hit_any) cur = kNoNode;
            // the node entered next was chosen against the t before this
            // leaf: test it against the new t, as the plain traversal does
            if (!kAny && cur != kNoNode && cur_t > t) cur = pop();
        }
        if (active && leaf == kNoNode && cur == kNoNode) finish();
    }
}

// Blocks of `kernel` the whole device holds at once (occupancy x compute
// units), looked up once per kernel: the grid of a persistent launch.
// Threads of the engine and of callers may ask at once: the cache is an
// atomic (two first callers both compute the same value).
template <auto kernel>
uint32_t resident_blocks(int block) {
    static std::atomic<int> cached{0};  // one per kernel instantiation
    int v = cached.load(std::memory_order_relaxed);
    if (!v) {
        int dev = 0, n_cu = 0, per_cu = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n_cu <= 0)
            n_cu = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu <= 0)
            per_cu = 1;
        v = n_cu * per_cu;
        cached.store(v, std::memory_order_relaxed);
    }
    return (uint32_t)v;
}

// Host-side dispatch to the traversal kernels instantiated for a stack size
// (64 or 256 entries, the smallest that covers the tree's stack need) and a
// node width (4 or 8): calls f(TravTag<stack, width>{}).
template <int S, int W>
struct TravTag {
    static constexpr int kStack = S;
    static constexpr int kW = W;
};

template <class F>
int with_traversal(uint32_t stack_need, int width, F &&f) {
    if (stack_need <= 64) return width == 8 ? f(TravTag<64, 8>{}) : f(TravTag<64, 4>{});
    if (stack_need <= 256) return width == 8 ? f(TravTag<256, 8>{}) : f(TravTag<256, 4>{});
    return set_error(RG_E_UNSUPPORTED, "BVH too deep for the traversal stack");
}

// Philox4x32-10 (Salmon et al., SC'11).
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


}  // namespace rg

// rg_build.hip — BVH construction on the GPU (gfx950).
//
// gpu_build_bvh makes, on the device, exactly what the host path makes with
// build_bvh2 + collapse_bvh4/8 + make_tris: the reference builder's binary
// tree (TriangleBVH::build/buildRecursive, bvh.cpp:76-231: the same node
// bounds, split decisions, node numbering and primitive permutation), its
// W-wide collapse with the per-octant visit orders and the top levels moved
// to the front, and the triangle records.
//
// Binary tree.  Ranges of more than kSmall primitives are split level by
// level over the whole array:
//   * bounds and bin boxes are min/max reductions.  The reference folds them
//     with fminf/fmaxf in primitive order, which on a tie (+0 against -0)
//     keeps the later operand; keying every value by (value, position) makes
//     that fold an order-free atomicMin/atomicMax on 64-bit keys, so the
//     reductions give the sequential bits whatever order the threads run in;
//   * the SAH sweep and the decision run per range in one thread with the
//     reference's arithmetic (-ffp-contract=off);
//   * the in-place partition (bvh.cpp:190-206) swaps the j-th left element
//     into position first + j.  Left elements therefore keep their order, a
//     right element at p >= first + nl stays, and one at p < first + nl moves
//     along the chain p -> l(p) -> l(l(p)) ... (l(j) = position of the j-th
//     left element) until it leaves [first, first + nl).  The chains are
//     resolved by pointer jumping, after a prefix sum gives every left
//     element its rank.
// Ranges of at most kSmall primitives are built by one wave each in LDS with
// the same steps, numbering their nodes in the recursion's order.  The host
// then numbers the level-built top nodes as the recursion allocates them
// (children of a node at the allocation counter when it splits, depth first,
// left first) and a kernel places every range's nodes.
//
// Wide tree.  Each wide node corresponds to a binary node whose greedy
// expansion (largest-area internal child first) fills its slots; the wide
// nodes are found level by level in breadth-first order, their depth-first
// numbers (the order the host collapse allocates them) follow from subtree
// sizes, and the first kTreeletBytes of nodes in breadth-first order go to
// the front as hoist_top places them.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "rg_internal.h"

namespace rg {

namespace {

typedef unsigned long long u64;

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr uint32_t kMaxBins = 32;    // bins the device builder supports
constexpr uint32_t kSmall = 256;     // ranges up to this size are built by one wave
constexpr uint32_t kBlk = 256;       // threads per block of the level kernels
constexpr uint32_t kPerBlk = 256;    // elements per block (<= kSmall: a block meets at most two ranges)
constexpr uint32_t kPerThread = kPerBlk / kBlk;
constexpr u64 kMinId = ~0ull;        // identity of the min keys
constexpr u64 kMaxId = 0ull;         // identity of the max keys

static_assert(kPerBlk <= kSmall, "a level block must not hold three ranges");

int hip_ok(hipError_t e, const char *what) {
    if (e == hipSuccess) return RG_OK;
    return set_error(RG_E_HIP, std::string("GPU BVH build, ") + what + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// Keys: a float's total order with -0 taken as +0, then the position, then
// the sign bit (which only tells +0 from -0 apart).  The smallest min key is
// the smallest value at its LAST position, the largest max key the largest
// value at its last position: what fminf/fmaxf(acc, x) folds in position
// order return, signed zeros included.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t ord_bits(float v) {
    uint32_t b = __float_as_uint(v);
    if (b == 0x80000000u) b = 0u;
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
// A NaN takes no part (the identity key), as fminf/fmaxf skip a NaN operand.
__device__ __forceinline__ u64 key_min(float v, uint32_t pos) {
    if (v != v) return ~0ull;
    return ((u64)ord_bits(v) << 32) | ((u64)(0x7FFFFFFFu - pos) << 1) | (u64)(__float_as_uint(v) >> 31);
}
__device__ __forceinline__ u64 key_max(float v, uint32_t pos) {
    if (v != v) return 0ull;
    return ((u64)ord_bits(v) << 32) | ((u64)pos << 1) | (u64)(__float_as_uint(v) >> 31);
}
// A key's value; the identities are the empty box's bounds (no real key
// equals either).
__device__ __forceinline__ float key_val(u64 k) {
    if (k == ~0ull) return FLT_MAX;
    if (k == 0ull) return -FLT_MAX;
    const uint32_t o = (uint32_t)(k >> 32);
    const uint32_t b = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
    return (b == 0u && (k & 1ull)) ? -0.0f : __uint_as_float(b);
}
// Key reductions over a wave (every lane gets the result).
__device__ __forceinline__ u64 wave_min(u64 v) {
    for (int o = 32; o > 0; o >>= 1) {
        const u64 w = __shfl_xor(v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}
__device__ __forceinline__ u64 wave_max(u64 v) {
    for (int o = 32; o > 0; o >>= 1) {
        const u64 w = __shfl_xor(v, o, 64);
        v = w > v ? w : v;
    }
    return v;
}

// the reference's fold steps, glibc's tie rule (the later operand)
__device__ __forceinline__ float min_acc(float acc, float x) { return (acc < x || x != x) ? acc : x; }
__device__ __forceinline__ float max_acc(float acc, float x) { return (acc > x || x != x) ? acc : x; }

__device__ __forceinline__ float box_area(const float lo[3], const float hi[3]) {
    const float dx = hi[0] - lo[0];
    const float dy = hi[1] - lo[1];
    const float dz = hi[2] - lo[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

// ---------------------------------------------------------------------------
// Device data
// ---------------------------------------------------------------------------
// Per-primitive arrays in the current order: box lo xyz, hi xyz, centroid
// xyz, original index, range id at the current level (kNone: not split here).
struct Prims {
    float *v[9];
    uint32_t *perm;
    uint32_t *seg;
};

// Key accumulators of one range at one level.
struct SegAcc {
    u64 box[12];  // pbox min xyz, pbox max xyz, centroid min xyz, centroid max xyz
    u64 bin[kMaxBins][6];
    uint32_t cnt[kMaxBins];
};

struct SegInfo {
    uint32_t first, count, node;
    int32_t axis;       // -1: leaf
    float cmin, scale;
    int32_t best;       // -1: leaf; else the split bin
    uint32_t nl;
    uint32_t child[2];  // range id at the next level, or kNone
};

// A node above the wave-built ranges.
struct TopNode {
    Node2 nd;           // bounds, axis; leaf: first and count
    uint32_t kind;      // 0 internal, 1 leaf, 2 root of a wave-built range
    uint32_t child[2];  // internal: top node ids
    uint32_t task;      // kind 2: the range's task id
};

struct Task {
    uint32_t first, count, node;
};

// Counters the host reads after each level.
struct LevelOut {
    uint32_t nseg_next, ntask, ntop, max_nl, error;
};

struct Build {
    uint32_t n, bins, leaf;
    uint32_t seg_cap, task_cap, top_cap;
    uint32_t nseg;               // ranges at the current level
    Prims a, b;                  // current and next order
    uint8_t *binv;
    uint32_t *flag, *scan, *jump;
    SegInfo *seg[2];             // this level's and the next level's ranges
    SegAcc *acc[2];
    TopNode *top;
    Task *task;
    uint32_t *task_nodes_n;      // nodes each wave-built range made
    Node2 *task_nodes;           // their nodes, range t at 2 * first
    LevelOut *out;
    uint32_t *final_perm;
};

// ---------------------------------------------------------------------------
// Binary build: level kernels
// ---------------------------------------------------------------------------
__global__ void k_prep(const float *__restrict__ v9, uint32_t n, Prims p, uint32_t root_seg) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *t = v9 + 9 * (size_t)i;
    float lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
        // BoundingBox::fromTriangle as the oracle folds it: empty box grown by v1, v2, v3
        lo[k] = min_acc(min_acc(min_acc(FLT_MAX, t[k]), t[3 + k]), t[6 + k]);
        hi[k] = max_acc(max_acc(max_acc(-FLT_MAX, t[k]), t[3 + k]), t[6 + k]);
    }
    for (int k = 0; k < 3; k++) {
        p.v[k][i] = lo[k];
        p.v[3 + k][i] = hi[k];
        p.v[6 + k][i] = (lo[k] + hi[k]) * 0.5f;
    }
    p.perm[i] = i;
    p.seg[i] = root_seg;
}

__device__ __forceinline__ void clear_box_keys(SegAcc &A) {
    for (int k = 0; k < 12; k++) A.box[k] = ((k / 3) & 1) ? kMaxId : kMinId;
}

__global__ void k_init_root(Build B, uint32_t n) {
    // level 0: one range (when n > kSmall) or one wave-built task
    TopNode &t = B.top[0];
    memset(&t, 0, sizeof t);
    if (n > kSmall) {
        SegInfo &s = B.seg[0][0];
        s.first = 0; s.count = n; s.node = 0;
        s.axis = -1; s.best = -1; s.nl = 0; s.child[0] = s.child[1] = kNone;
        clear_box_keys(B.acc[0][0]);
        t.kind = 0;
    } else {
        B.task[0] = Task{0, n, 0};
        t.kind = 2;
        t.task = 0;
    }
}

__global__ __launch_bounds__(kBlk) void k_bounds(Prims p, uint32_t n, uint32_t nseg, SegAcc *acc) {
    __shared__ u64 sk[2][12];
    __shared__ uint32_t s_lo, s_hi;
    const uint32_t tid = threadIdx.x;
    if (tid < 24) sk[tid / 12][tid % 12] = (((tid % 12) / 3) & 1) ? kMaxId : kMinId;
    if (tid == 0) { s_lo = kNone; s_hi = 0; }
    __syncthreads();
    const uint32_t base = blockIdx.x * kPerBlk;
    uint32_t sg[kPerThread];
    for (uint32_t e = 0; e < kPerThread; e++) {
        const uint32_t i = base + e * kBlk + tid;
        sg[e] = i < n ? p.seg[i] : kNone;
        if (sg[e] >= nseg) sg[e] = kNone;
        if (sg[e] != kNone) {
            atomicMin(&s_lo, sg[e]);
            atomicMax(&s_hi, sg[e]);
        }
    }
    __syncthreads();
    for (uint32_t e = 0; e < kPerThread; e++) {
        const uint32_t i = base + e * kBlk + tid;
        const uint32_t my = sg[e];
        const u64 act = __ballot(my != kNone);
        if (!act) continue;
        // a wave whose elements all belong to one range (the usual case)
        // reduces its keys first and updates LDS from one lane
        const uint32_t ws = __shfl(my, __ffsll((long long)act) - 1, 64);
        if (__ballot(my != kNone && my != ws) == 0) {
            u64 km[12];
            for (int k = 0; k < 3; k++) {
                const bool a = my != kNone;
                km[k] = a ? key_min(p.v[k][i], i) : kMinId;
                km[3 + k] = a ? key_max(p.v[3 + k][i], i) : kMaxId;
                km[6 + k] = a ? key_min(p.v[6 + k][i], i) : kMinId;
                km[9 + k] = a ? key_max(p.v[6 + k][i], i) : kMaxId;
            }
            for (int k = 0; k < 12; k++) km[k] = ((k / 3) & 1) ? wave_max(km[k]) : wave_min(km[k]);
            if ((threadIdx.x & 63u) == 0) {
                const int slot = ws == s_lo ? 0 : 1;
                for (int k = 0; k < 12; k++) {
                    if (((k / 3) & 1) == 0) atomicMin(&sk[slot][k], km[k]);
                    else atomicMax(&sk[slot][k], km[k]);
                }
            }
            continue;
        }
        if (my == kNone) continue;
        const int slot = my == s_lo ? 0 : 1;
        for (int k = 0; k < 3; k++) {
            atomicMin(&sk[slot][k], key_min(p.v[k][i], i));
            atomicMax(&sk[slot][3 + k], key_max(p.v[3 + k][i], i));
            atomicMin(&sk[slot][6 + k], key_min(p.v[6 + k][i], i));
            atomicMax(&sk[slot][9 + k], key_max(p.v[6 + k][i], i));
        }
    }
    __syncthreads();
    if (tid < 24 && s_lo != kNone) {
        const int slot = tid / 12, k = tid % 12;
        if (slot == 1 && s_hi == s_lo) return;
        const uint32_t sid = slot ? s_hi : s_lo;
        const u64 v = sk[slot][k];
        if (((k / 3) & 1) == 0) {
            if (v != kMinId) atomicMin(&acc[sid].box[k], v);
        } else {
            if (v != kMaxId) atomicMax(&acc[sid].box[k], v);
        }
    }
}

// Node bounds, leaf tests and the split axis of every range (bvh.cpp:102-133).
__global__ void k_axis(Build B, uint32_t nseg, const SegAcc *__restrict__ accin, SegInfo *segs, SegAcc *acc) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    SegInfo &S = segs[s];
    const SegAcc &A = accin[s];
    float plo[3], phi[3], clo[3], chi[3];
    for (int k = 0; k < 3; k++) {
        plo[k] = key_val(A.box[k]);
        phi[k] = key_val(A.box[3 + k]);
        clo[k] = key_val(A.box[6 + k]);
        chi[k] = key_val(A.box[9 + k]);
    }
    Node2 &nd = B.top[S.node].nd;
    for (int k = 0; k < 3; k++) { nd.bmin[k] = plo[k]; nd.bmax[k] = phi[k]; }
    S.axis = -1;
    S.best = -1;
    S.cmin = 0.0f;
    S.scale = 0.0f;
    if (S.count <= B.leaf) return;
    const float ext[3] = {chi[0] - clo[0], chi[1] - clo[1], chi[2] - clo[2]};
    int axis = 0;
    if (ext[1] > ext[0]) axis = 1;
    if (ext[2] > ((axis == 0) ? ext[0] : ext[1])) axis = 2;
    const float cext = ext[axis];
    if (cext < 1e-12f) return;
    S.axis = axis;
    S.cmin = clo[axis];
    S.scale = (float)B.bins / cext;
    SegAcc &W = acc[s];
    for (uint32_t b = 0; b < B.bins; b++) {
        for (int k = 0; k < 3; k++) { W.bin[b][k] = kMinId; W.bin[b][3 + k] = kMaxId; }
        W.cnt[b] = 0;
    }
}

__device__ __forceinline__ int bin_of(float c, float cmin, float scale, int last_bin) {
    int b = (int)((c - cmin) * scale);
    if (b < 0) b = 0;
    if (b > last_bin) b = last_bin;
    return b;
}

__global__ __launch_bounds__(kBlk) void k_bin(Prims p, uint32_t n, uint32_t nseg, uint32_t bins,
                                              const SegInfo *__restrict__ segs, SegAcc *acc, uint8_t *binv) {
    __shared__ u64 sk[2][kMaxBins][6];
    __shared__ uint32_t sc[2][kMaxBins];
    __shared__ uint32_t s_lo, s_hi;
    const uint32_t tid = threadIdx.x;
    for (uint32_t j = tid; j < 2 * kMaxBins * 6; j += kBlk) {
        const uint32_t slot = j / (kMaxBins * 6), r = j % (kMaxBins * 6);
        sk[slot][r / 6][r % 6] = (r % 6) < 3 ? kMinId : kMaxId;
    }
    if (tid < 2 * kMaxBins) sc[tid / kMaxBins][tid % kMaxBins] = 0;
    if (tid == 0) { s_lo = kNone; s_hi = 0; }
    __syncthreads();
    const uint32_t base = blockIdx.x * kPerBlk;
    uint32_t sg[kPerThread];
    for (uint32_t e = 0; e < kPerThread; e++) {
        const uint32_t i = base + e * kBlk + tid;
        sg[e] = i < n ? p.seg[i] : kNone;
        if (sg[e] >= nseg || segs[sg[e]].axis < 0) sg[e] = kNone;
        if (sg[e] != kNone) {
            atomicMin(&s_lo, sg[e]);
            atomicMax(&s_hi, sg[e]);
        }
    }
    __syncthreads();
    for (uint32_t e = 0; e < kPerThread; e++) {
        if (sg[e] == kNone) continue;
        const uint32_t i = base + e * kBlk + tid;
        const SegInfo &S = segs[sg[e]];
        const int b = bin_of(p.v[6 + S.axis][i], S.cmin, S.scale, (int)bins - 1);
        binv[i] = (uint8_t)b;
        const int slot = sg[e] == s_lo ? 0 : 1;
        for (int k = 0; k < 3; k++) {
            atomicMin(&sk[slot][b][k], key_min(p.v[k][i], i));
            atomicMax(&sk[slot][b][3 + k], key_max(p.v[3 + k][i], i));
        }
        atomicAdd(&sc[slot][b], 1u);
    }
    __syncthreads();
    if (s_lo == kNone) return;
    for (uint32_t j = tid; j < 2 * bins * 7; j += kBlk) {
        const uint32_t slot = j / (bins * 7), r = j % (bins * 7), b = r / 7, k = r % 7;
        if (slot == 1 && s_hi == s_lo) continue;
        const uint32_t sid = slot ? s_hi : s_lo;
        if (k == 6) {
            if (sc[slot][b]) atomicAdd(&acc[sid].cnt[b], sc[slot][b]);
        } else if (k < 3) {
            if (sk[slot][b][k] != kMinId) atomicMin(&acc[sid].bin[b][k], sk[slot][b][k]);
        } else {
            if (sk[slot][b][k] != kMaxId) atomicMax(&acc[sid].bin[b][k], sk[slot][b][k]);
        }
    }
}


// The SAH sweep and split decision of one range (bvh.cpp:145-186) from its
// bin keys, one lane per bin: the left boxes and counts are prefix
// min/max/sums over the lanes, the right ones suffix ones, the cost is the
// reference's 1 + 1.5 / A_p (A_L N_L + A_R N_R) in its operation order, and
// the first minimum cost below the leaf cost 1.5 N wins, as the reference's
// ascending strict-< scan picks it.  Returns the best bin (-1: leaf) and
// the left count.  The prefix boxes equal the
// sequential fold's up to the sign of a zero bound, which changes no area
// comparison.
__device__ int wave_sah(const u64 (*bk)[6], const uint32_t *bc, uint32_t bins, uint32_t count, const float plo[3],
                        const float phi[3], uint32_t lane, uint32_t &nl) {
    const float kTrav = 1.0f, kIsect = 1.5f;
    uint32_t c = 0;
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    if (lane < bins && bc[lane] > 0) {
        c = bc[lane];
        for (int k = 0; k < 3; k++) {
            lo[k] = key_val(bk[lane][k]);
            hi[k] = key_val(bk[lane][3 + k]);
        }
    }
    uint32_t lc = c, rc = c;
    float llo[3], lhi[3], rlo[3], rhi[3];
    for (int k = 0; k < 3; k++) { llo[k] = rlo[k] = lo[k]; lhi[k] = rhi[k] = hi[k]; }
    for (int o = 1; o < (int)kMaxBins; o <<= 1) {
        const uint32_t xl = __shfl_up(lc, o, 64), xr = __shfl_down(rc, o, 64);
        float al[3], ah[3], bl[3], bh[3];
        for (int k = 0; k < 3; k++) {
            al[k] = __shfl_up(llo[k], o, 64);
            ah[k] = __shfl_up(lhi[k], o, 64);
            bl[k] = __shfl_down(rlo[k], o, 64);
            bh[k] = __shfl_down(rhi[k], o, 64);
        }
        if (lane >= (uint32_t)o) {
            lc += xl;
            for (int k = 0; k < 3; k++) { llo[k] = fminf(llo[k], al[k]); lhi[k] = fmaxf(lhi[k], ah[k]); }
        }
        if (lane + o < bins) {
            rc += xr;
            for (int k = 0; k < 3; k++) { rlo[k] = fminf(rlo[k], bl[k]); rhi[k] = fmaxf(rhi[k], bh[k]); }
        }
    }
    // split i: left = prefix through bin i (this lane), right = suffix from bin i + 1
    const uint32_t rcn = __shfl_down(rc, 1, 64);
    float nlo[3], nhi[3];
    for (int k = 0; k < 3; k++) {
        nlo[k] = __shfl_down(rlo[k], 1, 64);
        nhi[k] = __shfl_down(rhi[k], 1, 64);
    }
    const float parea = box_area(plo, phi);
    const float inv_parea = (parea > 0.0f) ? 1.0f / parea : 0.0f;
    const float leaf_cost = (float)count * kIsect;
    float cost = FLT_MAX;
    bool valid = false;
    if (lane + 1 < bins && lc > 0 && rcn > 0) {
        const float la = box_area(llo, lhi), ra = box_area(nlo, nhi);
        cost = kTrav + kIsect * inv_parea * (la * (float)lc + ra * (float)rcn);
        valid = cost < leaf_cost;
    }
    // first lane holding the smallest valid cost
    float m = valid ? cost : FLT_MAX;
    for (int o = 32; o > 0; o >>= 1) m = fminf(m, __shfl_xor(m, o, 64));
    const u64 at = __ballot(valid && cost == m);
    if (!at) {
        nl = 0;
        return -1;
    }
    const int best = __ffsll((long long)at) - 1;
    nl = (uint32_t)__shfl(lc, best, 64);
    return best;
}

// The split decision of every range of the level, one wave per range (the
// bins' keys read by the lanes together): best bin and left count.
__global__ __launch_bounds__(256) void k_decide(Build B, uint32_t nseg, SegInfo *segs,
                                                const SegAcc *__restrict__ acc) {
    const uint32_t s = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63u;
    if (s >= nseg) return;
    SegInfo &S = segs[s];
    int best = -1;
    uint32_t nl = 0;
    if (S.axis >= 0) {
        float plo[3], phi[3];
        for (int k = 0; k < 3; k++) {
            plo[k] = key_val(acc[s].box[k]);
            phi[k] = key_val(acc[s].box[3 + k]);
        }
        best = wave_sah(acc[s].bin, acc[s].cnt, B.bins, S.count, plo, phi, lane, nl);
        if (best >= 0 && (nl == 0 || nl >= S.count)) best = -1;
    }
    if (lane == 0) {
        S.best = best;
        S.nl = best >= 0 ? nl : 0;
    }
}

// One block: allocates the children of every range k_decide split (top
// nodes in range order, left then right; next-level ranges and wave tasks in
// the same order), marks the others leaves and writes the counters.
constexpr uint32_t kSplitBlk = 1024;

__global__ __launch_bounds__(kSplitBlk) void k_split(Build B, uint32_t nseg, uint32_t ntop0, uint32_t ntask0,
                                                     SegInfo *segs, SegInfo *next, SegAcc *next_acc) {
    typedef hipcub::BlockScan<uint32_t, kSplitBlk> Scan;
    __shared__ typename Scan::TempStorage tmp;
    __shared__ uint32_t run_top, run_seg, run_task, max_nl;
    const uint32_t tid = threadIdx.x;
    if (tid == 0) { run_top = ntop0; run_seg = 0; run_task = ntask0; max_nl = 0; }
    __syncthreads();
    for (uint32_t base = 0; base < nseg; base += kSplitBlk) {
        const uint32_t s = base + tid;
        uint32_t nsplit = 0, nbig = 0, nsmall = 0;
        uint32_t nl = 0;
        int best = -1;
        if (s < nseg) {
            const SegInfo &S = segs[s];
            best = S.best;
            nl = S.nl;
            if (best >= 0) {
                nsplit = 1;
                nbig = (nl > kSmall) + (S.count - nl > kSmall);
                nsmall = 2 - nbig;
                atomicMax(&max_nl, nl);
            }
        }
        uint32_t r_top, r_big, r_small;
        Scan(tmp).ExclusiveSum(nsplit, r_top);
        __syncthreads();
        Scan(tmp).ExclusiveSum(nbig, r_big);
        __syncthreads();
        Scan(tmp).ExclusiveSum(nsmall, r_small);
        __syncthreads();
        if (s < nseg) {
            SegInfo &S = segs[s];
            TopNode &T = B.top[S.node];
            const uint32_t ids = run_top + 2 * r_top;
            if (best >= 0 && (ids + 2 > B.top_cap || run_seg + r_big + nbig > B.seg_cap ||
                              run_task + r_small + nsmall > B.task_cap)) {
                B.out->error = 4;
                best = -1;
                S.best = -1;
            }
            if (best < 0) {
                T.kind = 1;
                T.nd.first_or_left = S.first;
                T.nd.count = (uint16_t)min(S.count, 0xFFFFu);
                T.nd.axis = 0;
                S.child[0] = S.child[1] = kNone;
            } else {
                T.kind = 0;
                T.nd.count = 0;
                T.nd.axis = (uint16_t)S.axis;
                T.child[0] = ids;
                T.child[1] = ids + 1;
                uint32_t big = run_seg + r_big, small = run_task + r_small;
                for (int c = 0; c < 2; c++) {
                    const uint32_t f = c == 0 ? S.first : S.first + nl;
                    const uint32_t m = c == 0 ? nl : S.count - nl;
                    TopNode &C = B.top[ids + c];
                    memset(&C, 0, sizeof C);
                    if (m > kSmall) {
                        SegInfo &X = next[big];
                        X.first = f; X.count = m; X.node = ids + c;
                        X.axis = -1; X.best = -1; X.nl = 0; X.child[0] = X.child[1] = kNone;
                        clear_box_keys(next_acc[big]);
                        S.child[c] = big++;
                        C.kind = 0;
                    } else {
                        B.task[small] = Task{f, m, ids + c};
                        C.kind = 2;
                        C.task = small++;
                        S.child[c] = kNone;
                    }
                }
            }
        }
        __syncthreads();
        if (tid == kSplitBlk - 1) {
            run_top += 2 * (r_top + nsplit);
            run_seg += r_big + nbig;
            run_task += r_small + nsmall;
        }
        __syncthreads();
    }
    if (tid == 0) {
        B.out->nseg_next = run_seg;
        B.out->ntask = run_task;
        B.out->ntop = run_top;
        B.out->max_nl = max_nl;
    }
}

__device__ __forceinline__ bool goes_left(const SegInfo &S, uint8_t b) { return (int)b <= S.best; }

__global__ void k_flags(Build B, const SegInfo *__restrict__ segs) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n) return;
    const uint32_t s = B.a.seg[i];
    uint32_t f = 0;
    if (s < B.nseg) {
        const SegInfo &S = segs[s];
        f = (S.best >= 0 && goes_left(S, B.binv[i])) ? 1u : 0u;
    }
    B.flag[i] = f;
}

// The j-th left element of a range sits at l(j): jump[first + j] = l(j).
__global__ void k_left(Build B, const SegInfo *__restrict__ segs) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n || !B.flag[i]) return;
    const SegInfo &S = segs[B.a.seg[i]];  // flagged elements have a range of this level
    const uint32_t j = B.scan[i] - B.scan[S.first];
    if (j >= S.nl) { B.out->error = 1; return; }
    B.jump[S.first + j] = i;
}

// One round of pointer jumping over [first, first + nl) of every split range.
__global__ void k_jump(Build B, const SegInfo *__restrict__ segs) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= B.n) return;
    const uint32_t s = B.a.seg[q];
    if (s >= B.nseg) return;
    const SegInfo &S = segs[s];
    if (S.best < 0 || q - S.first >= S.nl) return;
    const uint32_t x = B.jump[q];
    if (x - S.first < S.nl) B.jump[q] = B.jump[x];
}

__global__ void k_scatter(Build B, const SegInfo *__restrict__ segs) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B.n) return;
    const uint32_t s = B.a.seg[i];
    uint32_t dest = i, nseg = kNone;
    if (s < B.nseg) {
        const SegInfo &S = segs[s];
        if (S.best >= 0) {
            if (B.flag[i]) {
                dest = S.first + (B.scan[i] - B.scan[S.first]);
                nseg = S.child[0];
            } else {
                dest = (i - S.first < S.nl) ? B.jump[i] : i;
                nseg = S.child[1];
                if (dest - S.first - S.nl >= S.count - S.nl) { B.out->error = 2; dest = i; }
            }
        }
    }
    if (dest >= B.n) { B.out->error = 3; return; }
    for (int k = 0; k < 9; k++) B.b.v[k][dest] = B.a.v[k][i];
    B.b.perm[dest] = B.a.perm[i];
    B.b.seg[dest] = nseg;
}

// ---------------------------------------------------------------------------
// Binary build: one wave per range of at most kSmall primitives
// ---------------------------------------------------------------------------
struct WaveLds {
    float v[6][kSmall];  // box lo xyz, hi xyz (the centroid is (lo + hi) * 0.5 again)
    uint32_t perm[kSmall];
    uint16_t ord[kSmall], ord2[kSmall], dst[kSmall], jump[kSmall];
    uint8_t binv[kSmall];
    u64 bk[kMaxBins][6];
    uint32_t bc[kMaxBins];
    uint16_t stack[kSmall][3];
};

__global__ __launch_bounds__(64) void k_wave_build(Build B, uint32_t ntask) {
    __shared__ WaveLds L;
    const uint32_t t = blockIdx.x;
    if (t >= ntask) return;
    const uint32_t lane = threadIdx.x;
    const Task T = B.task[t];
    const uint32_t first = T.first, m = T.count;
    if (m > kSmall || first + m > B.n) {
        if (lane == 0) B.out->error = 5;
        return;
    }
    Node2 *out = B.task_nodes + 2 * (size_t)first;
    for (uint32_t p = lane; p < m; p += 64) {
        for (int k = 0; k < 6; k++) L.v[k][p] = B.a.v[k][first + p];
        L.perm[p] = B.a.perm[first + p];
        L.ord[p] = (uint16_t)p;
    }
    uint32_t nn = 1, sp = 0;
    if (lane == 0) { L.stack[0][0] = 0; L.stack[0][1] = 0; L.stack[0][2] = (uint16_t)m; }
    sp = 1;
    __syncthreads();
    const int last_bin = (int)B.bins - 1;
    while (sp > 0) {
        sp--;
        const uint32_t k = L.stack[sp][0], a = L.stack[sp][1], c = L.stack[sp][2];
        __syncthreads();
        // bounds over [a, a + c) in the current order
        u64 km[12];
        for (int j = 0; j < 12; j++) km[j] = ((j / 3) & 1) ? kMaxId : kMinId;
        for (uint32_t p = a + lane; p < a + c; p += 64) {
            const uint32_t e = L.ord[p];
            for (int j = 0; j < 3; j++) {
                const float lo = L.v[j][e], hi = L.v[3 + j][e];
                const float ce = (lo + hi) * 0.5f;
                const u64 a0 = key_min(lo, p), a1 = key_max(hi, p);
                const u64 a2 = key_min(ce, p), a3 = key_max(ce, p);
                km[j] = a0 < km[j] ? a0 : km[j];
                km[3 + j] = a1 > km[3 + j] ? a1 : km[3 + j];
                km[6 + j] = a2 < km[6 + j] ? a2 : km[6 + j];
                km[9 + j] = a3 > km[9 + j] ? a3 : km[9 + j];
            }
        }
        for (int j = 0; j < 12; j++) km[j] = ((j / 3) & 1) ? wave_max(km[j]) : wave_min(km[j]);
        Node2 nd;
        float clo[3], chi[3];
        for (int j = 0; j < 3; j++) {
            nd.bmin[j] = key_val(km[j]);
            nd.bmax[j] = key_val(km[3 + j]);
            clo[j] = key_val(km[6 + j]);
            chi[j] = key_val(km[9 + j]);
        }
        nd.first_or_left = first + a;
        nd.count = (uint16_t)min(c, 0xFFFFu);
        nd.axis = 0;
        bool leaf = c <= B.leaf;
        int axis = 0, best = -1;
        uint32_t nl = 0;
        if (!leaf) {
            const float ext[3] = {chi[0] - clo[0], chi[1] - clo[1], chi[2] - clo[2]};
            if (ext[1] > ext[0]) axis = 1;
            if (ext[2] > ((axis == 0) ? ext[0] : ext[1])) axis = 2;
            if (ext[axis] < 1e-12f) leaf = true;
            const float cmin = clo[axis];
            const float scale = (float)B.bins / ext[axis];
            if (!leaf) {
                for (uint32_t j = lane; j < B.bins; j += 64) {
                    for (int q = 0; q < 3; q++) { L.bk[j][q] = kMinId; L.bk[j][3 + q] = kMaxId; }
                    L.bc[j] = 0;
                }
                __syncthreads();
                for (uint32_t p = a + lane; p < a + c; p += 64) {
                    const uint32_t e = L.ord[p];
                    const float ce = (L.v[axis][e] + L.v[3 + axis][e]) * 0.5f;
                    const int b = bin_of(ce, cmin, scale, last_bin);
                    L.binv[p] = (uint8_t)b;
                    for (int q = 0; q < 3; q++) {
                        atomicMin(&L.bk[b][q], key_min(L.v[q][e], p));
                        atomicMax(&L.bk[b][3 + q], key_max(L.v[3 + q][e], p));
                    }
                    atomicAdd(&L.bc[b], 1u);
                }
                __syncthreads();
                best = wave_sah(L.bk, L.bc, B.bins, c, nd.bmin, nd.bmax, lane, nl);
                if (best >= 0 && (nl == 0 || nl >= c)) best = -1;
                if (best < 0) leaf = true;
            }
        }
        if (!leaf && sp + 2 > kSmall) {
            if (lane == 0) B.out->error = 6;
            leaf = true;
        }
        if (leaf) {
            if (lane == 0) out[k] = nd;
            __syncthreads();
            continue;
        }
        // partition: left ranks by ballot, right elements along their chains
        uint32_t run = 0;
        for (uint32_t p0 = a; p0 < a + c; p0 += 64) {
            const uint32_t p = p0 + lane;
            const bool f = p < a + c && (int)L.binv[p] <= best;
            const u64 bal = __ballot(f);
            const uint32_t rank = run + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
            if (f) {
                L.jump[a + rank] = (uint16_t)p;
                L.dst[p] = (uint16_t)(a + rank);
            }
            run += (uint32_t)__popcll(bal);
        }
        __syncthreads();
        for (uint32_t span = 1; span < 2 * nl; span <<= 1) {
            for (uint32_t q = a + lane; q < a + nl; q += 64) {
                const uint32_t x = L.jump[q];
                if (x < a + nl) L.jump[q] = L.jump[x];
            }
            __syncthreads();
        }
        for (uint32_t p = a + lane; p < a + c; p += 64)
            if ((int)L.binv[p] > best) L.dst[p] = p < a + nl ? L.jump[p] : (uint16_t)p;
        __syncthreads();
        for (uint32_t p = a + lane; p < a + c; p += 64) L.ord2[L.dst[p]] = L.ord[p];
        __syncthreads();
        for (uint32_t p = a + lane; p < a + c; p += 64) L.ord[p] = L.ord2[p];
        const uint32_t left = nn;
        nn += 2;
        nd.first_or_left = left;
        nd.count = 0;
        nd.axis = (uint16_t)axis;
        if (lane == 0) {
            out[k] = nd;
            L.stack[sp][0] = (uint16_t)(left + 1); L.stack[sp][1] = (uint16_t)(a + nl); L.stack[sp][2] = (uint16_t)(c - nl);
            L.stack[sp + 1][0] = (uint16_t)left; L.stack[sp + 1][1] = (uint16_t)a; L.stack[sp + 1][2] = (uint16_t)nl;
        }
        sp += 2;
        __syncthreads();
    }
    for (uint32_t p = lane; p < m; p += 64) B.final_perm[first + p] = L.perm[L.ord[p]];
    if (lane == 0) B.task_nodes_n[t] = nn;
}

// Places the level-built nodes (host-numbered) and every wave-built range's
// nodes: local node 0 at the range root's number, local k >= 1 at base + k - 1.
__global__ void k_place_top(const Node2 *__restrict__ top, const uint32_t *__restrict__ idx, uint32_t ntop,
                            Node2 *nodes, uint32_t total) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < ntop && idx[j] < total) nodes[idx[j]] = top[j];
}

__global__ void k_place_tasks(Build B, const uint2 *__restrict__ where, uint32_t ntask, Node2 *nodes,
                              uint32_t total) {
    const uint32_t t = blockIdx.x;
    if (t >= ntask) return;
    const uint32_t nn = min(B.task_nodes_n[t], 2 * kSmall);
    const uint32_t root = where[t].x, base = where[t].y;
    const Node2 *src = B.task_nodes + 2 * (size_t)B.task[t].first;
    for (uint32_t k = threadIdx.x; k < nn; k += blockDim.x) {
        Node2 nd = src[k];
        if (nd.count == 0) nd.first_or_left = base + nd.first_or_left - 1;
        const uint32_t d = k == 0 ? root : base + k - 1;
        if (d < total) nodes[d] = nd;
    }
}

// ---------------------------------------------------------------------------
// Wide tree
// ---------------------------------------------------------------------------
template <int W>
struct WideRec {
    uint32_t slot[W];     // binary node per slot (kNone: empty)
    uint32_t exp[W];      // expanded binary nodes (exp[0] = the wide node's own)
    uint32_t child[W];    // breadth-first index of the slot's wide node (kNone: leaf or empty)
    uint32_t nslot, nexp, ninner;
};

__device__ __forceinline__ float half_area(const Node2 &n) {
    const float dx = n.bmax[0] - n.bmin[0];
    const float dy = n.bmax[1] - n.bmin[1];
    const float dz = n.bmax[2] - n.bmin[2];
    return dx * dy + dy * dz + dz * dx;
}

// Slots of the wide node of binary node k2 (Collapse::fill's greedy
// expansion, rg_bvh.cpp).
template <int W>
__global__ void k_wide_expand(const Node2 *__restrict__ n2, uint32_t nn2, const uint32_t *__restrict__ frontier,
                              uint32_t lo, uint32_t cnt, WideRec<W> *rec, uint32_t *ninner, uint32_t *error) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cnt) return;
    const uint32_t b = lo + j;
    const uint32_t k2 = frontier[b];
    WideRec<W> r;
    if (k2 >= nn2 || n2[k2].first_or_left + 1 >= nn2) {
        // not an internal node of this tree: an empty record
        *error = 1;
        for (int s = 0; s < W; s++) r.slot[s] = r.exp[s] = r.child[s] = kNone;
        r.nslot = r.nexp = r.ninner = 0;
        rec[b] = r;
        ninner[j] = 0;
        return;
    }
    r.slot[0] = n2[k2].first_or_left;
    r.slot[1] = r.slot[0] + 1;
    r.nslot = 2;
    r.exp[0] = k2;
    r.nexp = 1;
    while ((int)r.nslot < W) {
        int pick = -1;
        float best = -1.0f;
        for (int s = 0; s < (int)r.nslot; s++) {
            const Node2 &c = n2[r.slot[s]];
            if (c.count == 0 && half_area(c) > best) { best = half_area(c); pick = s; }
        }
        if (pick < 0) break;
        if (n2[r.slot[pick]].first_or_left + 1 >= nn2) {
            *error = 1;
            break;
        }
        const uint32_t k = r.slot[pick];
        r.exp[r.nexp++] = k;
        for (int s = (int)r.nslot; s > pick + 1; s--) r.slot[s] = r.slot[s - 1];
        r.slot[pick] = n2[k].first_or_left;
        r.slot[pick + 1] = n2[k].first_or_left + 1;
        r.nslot++;
    }
    uint32_t inner = 0;
    for (int s = 0; s < W; s++) {
        if (s >= (int)r.nslot) { r.slot[s] = kNone; r.child[s] = kNone; continue; }
        r.child[s] = kNone;
        if (n2[r.slot[s]].count == 0) inner++;
    }
    for (int s = (int)r.nexp; s < W; s++) r.exp[s] = kNone;
    r.ninner = inner;
    rec[b] = r;
    ninner[j] = inner;
}

template <int W>
__global__ void k_wide_link(const Node2 *__restrict__ n2, uint32_t *frontier, uint32_t lo, uint32_t cnt,
                            WideRec<W> *rec, const uint32_t *__restrict__ pos, uint32_t cap, uint32_t *error) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cnt) return;
    const uint32_t b = lo + j;
    WideRec<W> &r = rec[b];
    uint32_t at = lo + cnt + pos[j];
    for (int s = 0; s < W; s++) {
        if (r.slot[s] == kNone || n2[r.slot[s]].count != 0) continue;
        if (at >= cap) {
            *error = 1;
            r.child[s] = kNone;
            continue;
        }
        frontier[at] = r.slot[s];
        r.child[s] = at++;
    }
}

template <int W>
__global__ void k_wide_size(const WideRec<W> *__restrict__ rec, uint32_t lo, uint32_t cnt, uint32_t *size) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cnt) return;
    const uint32_t b = lo + j;
    uint32_t s = 1;
    for (int q = 0; q < W; q++)
        if (rec[b].child[q] != kNone) s += size[rec[b].child[q]];
    size[b] = s;
}

// Depth-first numbers (Collapse::fill allocates a node's inner children as
// one block, then fills them in slot order) and the traversal stack depth
// (stack_need_wide).
template <int W>
__global__ void k_wide_dfs(const WideRec<W> *__restrict__ rec, uint32_t lo, uint32_t cnt,
                           const uint32_t *__restrict__ size, uint32_t *dfs, uint32_t *cbase, uint32_t *acc,
                           uint32_t *stack_best) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cnt) return;
    const uint32_t b = lo + j;
    const WideRec<W> &r = rec[b];
    const uint32_t here = acc[b] + (r.nslot > 0 ? r.nslot - 1 : 0);
    atomicMax(stack_best, here);
    uint32_t at = cbase[b];
    uint32_t next_base = at + r.ninner;
    for (int s = 0; s < W; s++) {
        const uint32_t c = r.child[s];
        if (c == kNone) continue;
        dfs[c] = at++;
        cbase[c] = next_base;
        acc[c] = here;
        next_base += size[c] - 1;
    }
}

__device__ __forceinline__ uint32_t final_index(uint32_t b, const uint32_t *__restrict__ dfs, uint32_t top) {
    if (b < top) return b;
    const uint32_t d = dfs[b];
    uint32_t below = 0;
    for (uint32_t h = 0; h < top; h++) below += dfs[h] < d;
    return top + d - below;
}

__device__ __forceinline__ void set_order_dev(Node4 &nd, int oct, const int *ord) {
    nd.order[oct] = (uint8_t)(ord[0] | (ord[1] << 2) | (ord[2] << 4) | (ord[3] << 6));
    nd.rank[oct] = (uint8_t)(0 << (2 * ord[0]) | 1 << (2 * ord[1]) | 2 << (2 * ord[2]) | 3 << (2 * ord[3]));
}
__device__ __forceinline__ void set_order_dev(Node8 &nd, int oct, const int *ord) {
    uint32_t o = 0;
    for (int s = 0; s < 8; s++) o |= (uint32_t)ord[s] << (3 * s);
    nd.order[oct] = o;
}

template <class NodeT, int W>
__global__ void k_wide_write(const Node2 *__restrict__ n2, const WideRec<W> *__restrict__ rec, uint32_t nw,
                             const uint32_t *__restrict__ dfs, uint32_t top, NodeT *out) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nw) return;
    const WideRec<W> &r = rec[b];
    NodeT nd;
    memset(&nd, 0, sizeof nd);
    for (int s = 0; s < W; s++) {
        if (r.slot[s] == kNone) {
            nd.lox[s] = nd.loy[s] = nd.loz[s] = nd.hix[s] = nd.hiy[s] = nd.hiz[s] = 0.0f;
            nd.child[s] = kEmpty;
            continue;
        }
        const Node2 &c = n2[r.slot[s]];
        nd.lox[s] = c.bmin[0]; nd.loy[s] = c.bmin[1]; nd.loz[s] = c.bmin[2];
        nd.hix[s] = c.bmax[0]; nd.hiy[s] = c.bmax[1]; nd.hiz[s] = c.bmax[2];
        nd.child[s] = c.count > 0 ? (kLeafBit | c.first_or_left)
                                  : (r.child[s] < nw ? final_index(r.child[s], dfs, top) : kEmpty);
    }
    for (int oct = 0; oct < 8; oct++) {
        int ord[W], m = 0;
        uint32_t st[2 * W];
        int sp = 0;
        st[sp++] = r.exp[0];
        while (sp > 0) {
            const uint32_t k = st[--sp];
            bool expanded = false;
            for (int e = 0; e < (int)r.nexp; e++) expanded |= r.exp[e] == k;
            if (!expanded) {
                for (int s = 0; s < (int)r.nslot; s++)
                    if (r.slot[s] == k) { ord[m++] = s; break; }
                continue;
            }
            const Node2 &x = n2[k];
            const uint32_t l = x.first_or_left, rr = l + 1;
            // pushed in reverse of the visit order
            if ((oct >> x.axis) & 1) { st[sp++] = l; st[sp++] = rr; }
            else { st[sp++] = rr; st[sp++] = l; }
        }
        for (int s = 0; s < W; s++) {
            bool seen = false;
            for (int q = 0; q < m; q++) seen |= ord[q] == s;
            if (!seen) ord[m++] = s;
        }
        set_order_dev(nd, oct, ord);
    }
    const uint32_t at = final_index(b, dfs, top);
    if (at < nw) out[at] = nd;
}

// ---------------------------------------------------------------------------
// Triangle records (make_tris)
// ---------------------------------------------------------------------------
__global__ void k_tris(const float *__restrict__ v9, const uint32_t *__restrict__ perm,
                       const uint8_t *__restrict__ pred, uint32_t n, Tri48 *out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t o = perm[k];
    const float *t = v9 + 9 * (size_t)o;
    const uint32_t mask = pred ? (uint32_t)pred[o] : 0xFFu;
    float4 *r = reinterpret_cast<float4 *>(out + k);
    r[0] = make_float4(t[0], t[1], t[2], __uint_as_float(o));
    r[1] = make_float4(t[3] - t[0], t[4] - t[1], t[5] - t[2], 0.0f);
    r[2] = make_float4(t[6] - t[0], t[7] - t[1], t[8] - t[2], __uint_as_float(mask));
}

__global__ void k_leaf_last(const Node2 *__restrict__ n2, uint32_t nn, uint32_t n, Tri48 *out) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nn) return;
    const Node2 &nd = n2[j];
    if (nd.count == 0) return;
    const uint32_t last = nd.first_or_left + nd.count - 1;
    if (last < n) out[last].b[3] = __uint_as_float(1u);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
uint32_t blocks_for(uint64_t n, uint32_t per) { return (uint32_t)std::max<uint64_t>(1, (n + per - 1) / per); }

// One device allocation carved into aligned pieces.
struct Arena {
    char *base = nullptr;
    size_t used = 0, cap = 0;
    template <class T>
    T *take(size_t count) {
        used = (used + 255) & ~(size_t)255;
        T *p = reinterpret_cast<T *>(base + used);
        used += count * sizeof(T);
        return p;
    }
};

template <class T>
size_t piece(size_t count) { return ((count * sizeof(T)) + 255) & ~(size_t)255; }

// The context's workspace of at least `bytes` (kept between builds).
int workspace(Ctx &c, size_t bytes, char **out) {
    if (c.build_ws_cap < bytes) {
        dev_free(c.build_ws);
        c.build_ws = nullptr;
        c.build_ws_cap = 0;
        const size_t cap = bytes + bytes / 8;
        if (int rc = dev_malloc(&c.build_ws, cap)) return rc;
        c.build_ws_cap = cap;
    }
    *out = (char *)c.build_ws;
    return RG_OK;
}

struct HostTop {
    Node2 nd;
    uint32_t kind, child[2], task;
};
static_assert(sizeof(HostTop) == sizeof(TopNode), "top node layouts differ");

template <class NodeT, int W>
int build_wide(Ctx &c, const Node2 *d_n2, uint32_t nn2, hipStream_t st, void **d_wide_out, uint32_t &n_wide,
               uint32_t &stack_need) {
    // upper bound: one wide node per internal binary node
    const uint32_t cap = std::max<uint32_t>(1, nn2);
    size_t scan_bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (uint32_t *)nullptr, (uint32_t *)nullptr, cap, st);
    const size_t bytes = piece<WideRec<W>>(cap) + 6 * piece<uint32_t>(cap) + piece<uint32_t>(64) +
                         piece<char>(scan_bytes) + 1024;
    char *mem = nullptr;
    if (int rc = workspace(c, bytes, &mem)) return rc;
    Arena ar{mem, 0, bytes};
    WideRec<W> *rec = ar.take<WideRec<W>>(cap);
    uint32_t *frontier = ar.take<uint32_t>(cap);
    uint32_t *ninner = ar.take<uint32_t>(cap);
    uint32_t *pos = ar.take<uint32_t>(cap);
    uint32_t *size = ar.take<uint32_t>(cap);
    uint32_t *dfs = ar.take<uint32_t>(cap);
    uint32_t *cbase = ar.take<uint32_t>(cap);
    uint32_t *small = ar.take<uint32_t>(64);
    void *scan_tmp = ar.take<char>(scan_bytes);
    uint32_t *acc = pos;  // reused after the levels are known
    int rc = RG_OK;
    std::vector<uint32_t> lv_lo, lv_cnt;
    const uint32_t zero = 0, one = 1;
    const uint32_t zeros[2] = {0, 0};
    rc = hip_ok(hipMemcpyAsync(frontier, &zero, 4, hipMemcpyHostToDevice, st), "frontier");
    if (!rc) rc = hip_ok(hipMemcpyAsync(small, zeros, 8, hipMemcpyHostToDevice, st), "wide flags");
    uint32_t lo = 0, cnt = 1;
    while (!rc && cnt > 0) {
        if ((uint64_t)lo + cnt > cap) { rc = set_error(RG_E_INTERNAL, "GPU BVH build: wide node overflow"); break; }
        lv_lo.push_back(lo);
        lv_cnt.push_back(cnt);
        hipLaunchKernelGGL((k_wide_expand<W>), dim3(blocks_for(cnt, 256)), dim3(256), 0, st, d_n2, nn2, frontier, lo,
                           cnt, rec, ninner, small + 1);
        size_t tb = scan_bytes;
        rc = hip_ok(hipcub::DeviceScan::ExclusiveSum(scan_tmp, tb, ninner, pos, cnt, st), "wide scan");
        if (rc) break;
        hipLaunchKernelGGL((k_wide_link<W>), dim3(blocks_for(cnt, 256)), dim3(256), 0, st, d_n2, frontier, lo, cnt,
                           rec, pos, cap, small + 1);
        uint32_t tail[2];
        rc = hip_ok(hipMemcpyAsync(&tail[0], pos + cnt - 1, 4, hipMemcpyDeviceToHost, st), "wide count");
        if (!rc) rc = hip_ok(hipMemcpyAsync(&tail[1], ninner + cnt - 1, 4, hipMemcpyDeviceToHost, st), "wide count");
        if (!rc) rc = hip_ok(hipStreamSynchronize(st), "wide level");
        if (rc) break;
        lo += cnt;
        cnt = tail[0] + tail[1];
    }
    const uint32_t nw = lo;
    if (!rc) {
        for (size_t l = lv_lo.size(); l-- > 0;)
            hipLaunchKernelGGL((k_wide_size<W>), dim3(blocks_for(lv_cnt[l], 256)), dim3(256), 0, st, rec, lv_lo[l],
                               lv_cnt[l], size);
        rc = hip_ok(hipMemcpyAsync(dfs, &zero, 4, hipMemcpyHostToDevice, st), "wide root");
        if (!rc) rc = hip_ok(hipMemcpyAsync(cbase, &one, 4, hipMemcpyHostToDevice, st), "wide root");
        if (!rc) rc = hip_ok(hipMemcpyAsync(acc, &zero, 4, hipMemcpyHostToDevice, st), "wide root");
    }
    if (!rc)
        for (size_t l = 0; l < lv_lo.size(); l++)
            hipLaunchKernelGGL((k_wide_dfs<W>), dim3(blocks_for(lv_cnt[l], 256)), dim3(256), 0, st, rec, lv_lo[l],
                               lv_cnt[l], size, dfs, cbase, acc, small);
    NodeT *wide = nullptr;
    if (!rc) rc = dev_malloc((void **)&wide, (size_t)nw * sizeof(NodeT));
    const uint32_t top = std::min<uint32_t>(nw, kTreeletBytes / (uint32_t)sizeof(NodeT));
    if (!rc)
        hipLaunchKernelGGL((k_wide_write<NodeT, W>), dim3(blocks_for(nw, 128)), dim3(128), 0, st, d_n2, rec, nw, dfs,
                           top, wide);
    uint32_t flags[2] = {0, 0};
    if (!rc) rc = hip_ok(hipGetLastError(), "wide kernels");
    if (!rc) rc = hip_ok(hipMemcpyAsync(flags, small, 8, hipMemcpyDeviceToHost, st), "wide stack");
    if (!rc) rc = hip_ok(hipStreamSynchronize(st), "wide tree");
    if (!rc && flags[1]) rc = set_error(RG_E_INTERNAL, "GPU BVH build: wide tree inconsistent");
    const uint32_t best = flags[0];
    if (rc) {
        dev_free(wide);
        return rc;
    }
    *d_wide_out = wide;
    n_wide = nw;
    stack_need = best + 1;
    return RG_OK;
}

}  // namespace

bool gpu_build_supported(uint32_t bins, uint32_t leaf) { return bins >= 2 && bins <= kMaxBins && leaf >= 1; }

// The scene's triangles on the device: the copy kept from an earlier upload
// (they never change after rg_scene_create), else uploaded now and kept.
static int device_v9(Scene &s, hipStream_t st) {
    if (s.d_v9) return RG_OK;
    if (int rc = dev_malloc(&s.d_v9, (size_t)s.n_tris * 36)) return rc;
    return hip_ok(hipMemcpyAsync(s.d_v9, s.v9.data(), (size_t)s.n_tris * 36, hipMemcpyHostToDevice, st), "upload");
}

int gpu_make_tris(Scene &s, void *stream) {
    hipStream_t st = (hipStream_t)stream;
    const uint32_t n = s.n_tris;
    if (n == 0 || !s.d_tris || !s.d_perm || !s.d_nodes2) return RG_OK;
    uint8_t *pred = nullptr;
    int rc = device_v9(s, st);
    const float *v9 = (const float *)s.d_v9;
    if (!rc && !s.pred_mask.empty()) rc = dev_malloc((void **)&pred, n);
    if (!rc && pred) rc = hip_ok(hipMemcpyAsync(pred, s.pred_mask.data(), n, hipMemcpyHostToDevice, st), "upload");
    if (!rc) {
        hipLaunchKernelGGL(k_tris, dim3(blocks_for(n, 256)), dim3(256), 0, st, v9, (const uint32_t *)s.d_perm, pred,
                           n, (Tri48 *)s.d_tris);
        hipLaunchKernelGGL(k_leaf_last, dim3(blocks_for(s.n_nodes2, 256)), dim3(256), 0, st,
                           (const Node2 *)s.d_nodes2, s.n_nodes2, n, (Tri48 *)s.d_tris);
        rc = hip_ok(hipGetLastError(), "triangle records");
    }
    if (!rc) rc = hip_ok(hipStreamSynchronize(st), "triangle records");
    dev_free(pred);
    return rc;
}

void release_build(Ctx &c) {
    dev_free(c.build_ws);
    c.build_ws = nullptr;
    c.build_ws_cap = 0;
    if (c.build_host) (void)hipHostFree(c.build_host);
    c.build_host = nullptr;
}

int gpu_build_bvh(Ctx &c, Scene &s, uint32_t bins, uint32_t leaf, int width) {
    hipStream_t st = (hipStream_t)c.stream;
    const uint32_t n = s.n_tris;
    if (n == 0 || !gpu_build_supported(bins, leaf))
        return set_error(RG_E_INTERNAL, "GPU BVH build: unsupported input");
    if (n >= 0x7FFFFFFFu) return set_error(RG_E_UNSUPPORTED, "GPU BVH build: too many triangles");
    const uint32_t seg_cap = n / (kSmall + 1) + 2;
    const uint32_t task_cap = 4 * (n / kSmall) + 8;
    const uint32_t top_cap = 2 * seg_cap + task_cap + 8;
    size_t scan_bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (uint32_t *)nullptr, (uint32_t *)nullptr, n, st);
    const size_t bytes = 20 * piece<float>(n) + 4 * piece<uint32_t>(n) + piece<uint8_t>(n) + 3 * piece<uint32_t>(n) +
                         2 * piece<SegInfo>(seg_cap) + 2 * piece<SegAcc>(seg_cap) + piece<TopNode>(top_cap) +
                         piece<Task>(task_cap) + piece<uint32_t>(task_cap) + piece<Node2>(2 * (size_t)n) +
                         piece<LevelOut>(1) + piece<char>(scan_bytes) + 4096;
    char *mem = nullptr;
    if (int rc = workspace(c, bytes, &mem)) return rc;
    Arena ar{mem, 0, bytes};
    Build B;
    B.n = n;
    B.bins = bins;
    B.leaf = leaf;
    B.seg_cap = seg_cap;
    B.task_cap = task_cap;
    B.top_cap = top_cap;
    B.nseg = 0;
    for (Prims *p : {&B.a, &B.b}) {
        for (int k = 0; k < 9; k++) p->v[k] = ar.take<float>(n);
        p->perm = ar.take<uint32_t>(n);
        p->seg = ar.take<uint32_t>(n);
    }
    B.binv = ar.take<uint8_t>(n);
    B.flag = ar.take<uint32_t>(n);
    B.scan = ar.take<uint32_t>(n);
    B.jump = ar.take<uint32_t>(n);
    B.seg[0] = ar.take<SegInfo>(seg_cap);
    B.seg[1] = ar.take<SegInfo>(seg_cap);
    B.acc[0] = ar.take<SegAcc>(seg_cap);
    B.acc[1] = ar.take<SegAcc>(seg_cap);
    B.top = ar.take<TopNode>(top_cap);
    B.task = ar.take<Task>(task_cap);
    B.task_nodes_n = ar.take<uint32_t>(task_cap);
    B.task_nodes = ar.take<Node2>(2 * (size_t)n);
    B.out = ar.take<LevelOut>(1);
    void *scan_tmp = ar.take<char>(scan_bytes);

    // results kept by the scene
    Node2 *d_n2 = nullptr;
    uint32_t *d_perm = nullptr;
    Tri48 *d_tris = nullptr;
    int rc = dev_malloc((void **)&d_perm, (size_t)n * 4);
    if (!rc) rc = dev_malloc((void **)&d_tris, (size_t)n * sizeof(Tri48));
    B.final_perm = d_perm;

    if (!rc && !c.build_host) rc = hip_ok(hipHostMalloc(&c.build_host, 64), "hipHostMalloc");
    LevelOut *host_out = (LevelOut *)c.build_host;
    if (!rc) rc = device_v9(s, st);
    const float *v9 = (const float *)s.d_v9;
    if (!rc) rc = hip_ok(hipMemsetAsync(B.out, 0, sizeof(LevelOut), st), "clear");
    if (!rc) rc = hip_ok(hipMemsetAsync(B.jump, 0, (size_t)n * 4, st), "clear");
    if (!rc) {
        hipLaunchKernelGGL(k_prep, dim3(blocks_for(n, 256)), dim3(256), 0, st, v9, n, B.a, n > kSmall ? 0u : kNone);
        hipLaunchKernelGGL(k_init_root, dim3(1), dim3(1), 0, st, B, n);
        rc = hip_ok(hipGetLastError(), "prep");
    }
    uint32_t nseg = n > kSmall ? 1 : 0, ntop = 1, ntask = n > kSmall ? 0 : 1;
    int cur = 0;
    for (uint32_t level = 0; !rc && nseg > 0; level++) {
        if (level > n) {
            rc = set_error(RG_E_INTERNAL, "GPU BVH build: levels do not terminate");
            break;
        }
        SegInfo *segs = B.seg[cur], *next = B.seg[cur ^ 1];
        SegAcc *acc = B.acc[cur], *next_acc = B.acc[cur ^ 1];
        const uint32_t nb = blocks_for(n, kPerBlk);
        B.nseg = nseg;
        hipLaunchKernelGGL(k_bounds, dim3(nb), dim3(kBlk), 0, st, B.a, n, nseg, acc);
        hipLaunchKernelGGL(k_axis, dim3(blocks_for(nseg, 128)), dim3(128), 0, st, B, nseg, acc, segs, acc);
        hipLaunchKernelGGL(k_bin, dim3(nb), dim3(kBlk), 0, st, B.a, n, nseg, bins, segs, acc, B.binv);
        hipLaunchKernelGGL(k_decide, dim3(blocks_for((uint64_t)nseg * 64, 256)), dim3(256), 0, st, B, nseg, segs, acc);
        hipLaunchKernelGGL(k_split, dim3(1), dim3(kSplitBlk), 0, st, B, nseg, ntop, ntask, segs, next, next_acc);
        rc = hip_ok(hipGetLastError(), "level kernels");
        if (!rc) rc = hip_ok(hipMemcpyAsync(host_out, B.out, sizeof(LevelOut), hipMemcpyDeviceToHost, st), "level");
        if (!rc) rc = hip_ok(hipStreamSynchronize(st), "level");
        if (rc) break;
        const LevelOut lo = *host_out;
        if (lo.error || lo.ntop > top_cap || lo.ntask > task_cap || lo.nseg_next > seg_cap) {
            rc = set_error(RG_E_INTERNAL, "GPU BVH build: level bookkeeping out of range");
            break;
        }
        const uint32_t nb1 = blocks_for(n, 256);
        hipLaunchKernelGGL(k_flags, dim3(nb1), dim3(256), 0, st, B, segs);
        size_t tb = scan_bytes;
        rc = hip_ok(hipcub::DeviceScan::ExclusiveSum(scan_tmp, tb, B.flag, B.scan, n, st), "scan");
        if (rc) break;
        hipLaunchKernelGGL(k_left, dim3(nb1), dim3(256), 0, st, B, segs);
        for (uint64_t span = 1; span < 2ull * lo.max_nl; span <<= 1)
            hipLaunchKernelGGL(k_jump, dim3(nb1), dim3(256), 0, st, B, segs);
        hipLaunchKernelGGL(k_scatter, dim3(nb1), dim3(256), 0, st, B, segs);
        rc = hip_ok(hipGetLastError(), "partition kernels");
        std::swap(B.a, B.b);
        nseg = lo.nseg_next;
        ntop = lo.ntop;
        ntask = lo.ntask;
        cur ^= 1;
    }
    // ranges left as leaves by the levels keep their order: copy it to the
    // final permutation, then let the waves build and permute theirs
    if (!rc) rc = hip_ok(hipMemcpyAsync(d_perm, B.a.perm, (size_t)n * 4, hipMemcpyDeviceToDevice, st), "perm");
    if (!rc && ntask > 0) {
        hipLaunchKernelGGL(k_wave_build, dim3(ntask), dim3(64), 0, st, B, ntask);
        rc = hip_ok(hipGetLastError(), "wave build");
    }
    std::vector<HostTop> top(ntop);
    std::vector<Task> tasks(ntask);
    std::vector<uint32_t> tnn(ntask);
    if (!rc) rc = hip_ok(hipMemcpyAsync(top.data(), B.top, ntop * sizeof(TopNode), hipMemcpyDeviceToHost, st), "top");
    if (!rc && ntask)
        rc = hip_ok(hipMemcpyAsync(tasks.data(), B.task, ntask * sizeof(Task), hipMemcpyDeviceToHost, st), "tasks");
    if (!rc && ntask)
        rc = hip_ok(hipMemcpyAsync(tnn.data(), B.task_nodes_n, ntask * 4, hipMemcpyDeviceToHost, st), "tasks");
    if (!rc) rc = hip_ok(hipMemcpyAsync(host_out, B.out, sizeof(LevelOut), hipMemcpyDeviceToHost, st), "status");
    if (!rc) rc = hip_ok(hipStreamSynchronize(st), "wave build");
    if (!rc && host_out->error) rc = set_error(RG_E_INTERNAL, "GPU BVH build: partition out of range");

    // number the top nodes as the recursion allocates them
    uint32_t total = 0;
    std::vector<uint32_t> idx(ntop, kNone), base(ntop, 0), size(ntop, 0);
    std::vector<uint2> where(ntask);
    if (!rc) {
        for (uint32_t j = ntop; j-- > 0;) {
            const HostTop &t = top[j];
            if (t.kind == 2) size[j] = tnn[t.task];
            else if (t.kind == 1) size[j] = 1;
            else size[j] = 1 + size[t.child[0]] + size[t.child[1]];
        }
        total = size[0];
        idx[0] = 0;
        base[0] = 1;
        for (uint32_t j = 0; j < ntop; j++) {
            HostTop &t = top[j];
            if (t.kind == 0) {
                const uint32_t l = t.child[0], r = t.child[1];
                idx[l] = base[j];
                idx[r] = base[j] + 1;
                base[l] = base[j] + 2;
                base[r] = base[l] + size[l] - 1;
                t.nd.first_or_left = base[j];
            } else if (t.kind == 2) {
                where[t.task] = make_uint2(idx[j], base[j]);
                idx[j] = kNone;  // placed by its wave's nodes
            }
        }
    }
    Node2 *d_top = nullptr;
    uint32_t *d_idx = nullptr;
    uint2 *d_where = nullptr;
    std::vector<Node2> top_nodes(ntop);
    for (uint32_t j = 0; j < ntop; j++) top_nodes[j] = top[j].nd;
    if (!rc) rc = dev_malloc((void **)&d_n2, (size_t)total * sizeof(Node2));
    if (!rc) rc = dev_malloc((void **)&d_top, (size_t)ntop * sizeof(Node2));
    if (!rc) rc = dev_malloc((void **)&d_idx, (size_t)ntop * 4);
    if (!rc && ntask) rc = dev_malloc((void **)&d_where, (size_t)ntask * sizeof(uint2));
    if (!rc) rc = hip_ok(hipMemcpyAsync(d_top, top_nodes.data(), ntop * sizeof(Node2), hipMemcpyHostToDevice, st), "top");
    if (!rc) rc = hip_ok(hipMemcpyAsync(d_idx, idx.data(), ntop * 4, hipMemcpyHostToDevice, st), "top");
    if (!rc && ntask)
        rc = hip_ok(hipMemcpyAsync(d_where, where.data(), ntask * sizeof(uint2), hipMemcpyHostToDevice, st), "tasks");
    if (!rc) {
        hipLaunchKernelGGL(k_place_top, dim3(blocks_for(ntop, 256)), dim3(256), 0, st, d_top, d_idx, ntop, d_n2, total);
        if (ntask) hipLaunchKernelGGL(k_place_tasks, dim3(ntask), dim3(256), 0, st, B, d_where, ntask, d_n2, total);
        hipLaunchKernelGGL(k_tris, dim3(blocks_for(n, 256)), dim3(256), 0, st, v9, d_perm, (const uint8_t *)nullptr, n,
                           d_tris);
        hipLaunchKernelGGL(k_leaf_last, dim3(blocks_for(total, 256)), dim3(256), 0, st, d_n2, total, n, d_tris);
        rc = hip_ok(hipGetLastError(), "placement");
    }
    if (!rc && !s.pred_mask.empty()) {
        // occluder bits per triangle: rewrite the c.w words from the mask
        uint8_t *pred = reinterpret_cast<uint8_t *>(B.flag);  // n bytes of scratch
        rc = hip_ok(hipMemcpyAsync(pred, s.pred_mask.data(), n, hipMemcpyHostToDevice, st), "predicates");
        if (!rc) {
            hipLaunchKernelGGL(k_tris, dim3(blocks_for(n, 256)), dim3(256), 0, st, v9, d_perm, pred, n, d_tris);
            hipLaunchKernelGGL(k_leaf_last, dim3(blocks_for(total, 256)), dim3(256), 0, st, d_n2, total, n, d_tris);
        }
    }
    if (!rc) rc = hip_ok(hipStreamSynchronize(st), "binary tree");
    dev_free(d_top);
    dev_free(d_idx);
    dev_free(d_where);

    // the wide tree
    void *d_wide = nullptr;
    uint32_t n_wide = 0, stack_need = 0;
    Node2 root{};
    if (!rc) rc = hip_ok(hipMemcpyAsync(&root, d_n2, sizeof(Node2), hipMemcpyDeviceToHost, st), "root");
    if (!rc) rc = hip_ok(hipStreamSynchronize(st), "root");
    if (!rc) {
        if (root.count > 0) {
            // a leaf root: one wide node with one leaf child (collapse_wide)
            std::vector<uint8_t> raw(width == 8 ? sizeof(Node8) : sizeof(Node4), 0);
            if (width == 8) {
                Node8 nd;
                std::memset(&nd, 0, sizeof nd);
                nd.lox[0] = root.bmin[0]; nd.loy[0] = root.bmin[1]; nd.loz[0] = root.bmin[2];
                nd.hix[0] = root.bmax[0]; nd.hiy[0] = root.bmax[1]; nd.hiz[0] = root.bmax[2];
                nd.child[0] = kLeafBit | root.first_or_left;
                for (int q = 1; q < 8; q++) nd.child[q] = kEmpty;
                for (int oct = 0; oct < 8; oct++) {
                    uint32_t o = 0;
                    for (int q = 0; q < 8; q++) o |= (uint32_t)q << (3 * q);
                    nd.order[oct] = o;
                }
                std::memcpy(raw.data(), &nd, sizeof nd);
            } else {
                Node4 nd;
                std::memset(&nd, 0, sizeof nd);
                nd.lox[0] = root.bmin[0]; nd.loy[0] = root.bmin[1]; nd.loz[0] = root.bmin[2];
                nd.hix[0] = root.bmax[0]; nd.hiy[0] = root.bmax[1]; nd.hiz[0] = root.bmax[2];
                nd.child[0] = kLeafBit | root.first_or_left;
                for (int q = 1; q < 4; q++) nd.child[q] = kEmpty;
                for (int oct = 0; oct < 8; oct++) {
                    nd.order[oct] = (uint8_t)(0 | (1 << 2) | (2 << 4) | (3 << 6));
                    nd.rank[oct] = nd.order[oct];  // the identity is its own inverse
                }
                std::memcpy(raw.data(), &nd, sizeof nd);
            }
            rc = dev_malloc(&d_wide, raw.size());
            if (!rc) rc = dev_upload(d_wide, raw.data(), raw.size(), st);
            n_wide = 1;
            stack_need = 1;
        } else if (width == 8) {
            rc = build_wide<Node8, 8>(c, d_n2, total, st, &d_wide, n_wide, stack_need);
        } else {
            rc = build_wide<Node4, 4>(c, d_n2, total, st, &d_wide, n_wide, stack_need);
        }
    }
    if (rc) {
        dev_free(d_n2);
        dev_free(d_perm);
        dev_free(d_tris);
        dev_free(d_wide);
        return rc;
    }
    s.d_nodes2 = d_n2;
    s.d_perm = d_perm;
    s.d_tris = d_tris;
    s.d_nodes4 = d_wide;
    s.n_nodes2 = total;
    s.n_wide = n_wide;
    s.stack_need = stack_need;
    for (int k = 0; k < 3; k++) { s.root_box[k] = root.bmin[k]; s.root_box[3 + k] = root.bmax[k]; }
    return RG_OK;
}

int download_host_tree(Scene &s, void *stream) {
    if (s.host_tree || !s.d_nodes2) return RG_OK;
    std::vector<Node2> nodes(s.n_nodes2);
    std::vector<uint32_t> perm(s.n_tris);
    hipStream_t st = (hipStream_t)stream;
    int rc = hip_ok(hipMemcpyAsync(nodes.data(), s.d_nodes2, nodes.size() * sizeof(Node2), hipMemcpyDeviceToHost, st),
                    "download");
    if (!rc) rc = hip_ok(hipMemcpyAsync(perm.data(), s.d_perm, perm.size() * 4, hipMemcpyDeviceToHost, st), "download");
    if (!rc) rc = hip_ok(hipStreamSynchronize(st), "download");
    if (rc) return rc;
    s.nodes2.swap(nodes);
    s.perm.swap(perm);
    s.host_tree = true;
    return RG_OK;
}

}  // namespace rg

// rg_bvh.cpp — host BVH construction.
//
// build_bvh2 follows the reference builder TriangleBVH::build/buildRecursive
// (src/raygen/bvh.cpp:76-231) decision for decision: the same prim/centroid
// bounding boxes, widest-centroid-axis choice, 16-bin SAH with cost
// 1 + 1.5/A_p (A_L N_L + A_R N_R) against 1.5 N, the same forced-leaf rules and
// the same in-place partition, so it yields the same nodes in the same order
// and the same permutation.  It is written as an explicit work stack (DFS,
// left subtree first), which allocates sibling pairs in the same order as the
// reference's recursion.  Large inputs split their top levels on the calling
// thread and build the subtrees below on worker threads (each owns its
// primitive range, which the partition only permutes within), then number
// the nodes exactly as the sequential recursion would: every decision of a
// subtree depends only on its own range, so the tree is the same.
//
// collapse_bvh4 then folds the binary tree into 128-byte 4-wide nodes for the
// GPU.  Every leaf and every leaf's primitive order is kept, and each node
// records, per ray-direction octant, the order in which the binary tree's
// near-child-first traversal (bvh.cpp:270-282) would reach its children.  With
// the box test repeated against the current closest t when an entry is popped,
// the 4-wide traversal then visits leaves in exactly the binary traversal's
// order and prunes exactly the same leaves (child boxes are nested inside
// parent boxes and the slab arithmetic is monotone), so hits, including the
// winner of exact-t ties, are identical.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <thread>

#include "rg_internal.h"

namespace rg {
namespace {

// fminf(acc, x) / fmaxf(acc, x) as the reference's expandBBox calls them
// (bvh.cpp:36-51) with glibc's tie rule: of two equal operands (+0 and -0)
// the second is returned, and a NaN operand yields the other one.  Written
// out because clang's inline fminf keeps the first operand on ties.
inline float min_acc(float acc, float x) { return (acc < x || x != x) ? acc : x; }
inline float max_acc(float acc, float x) { return (acc > x || x != x) ? acc : x; }

struct Box {
    float lo[3], hi[3];
    void clear() {
        for (int k = 0; k < 3; k++) { lo[k] = FLT_MAX; hi[k] = -FLT_MAX; }
    }
    void grow(const Box &b) {
        for (int k = 0; k < 3; k++) {
            lo[k] = min_acc(lo[k], b.lo[k]);
            hi[k] = max_acc(hi[k], b.hi[k]);
        }
    }
    void grow(const float p[3]) {
        for (int k = 0; k < 3; k++) {
            lo[k] = min_acc(lo[k], p[k]);
            hi[k] = max_acc(hi[k], p[k]);
        }
    }
    float area() const {
        const float dx = hi[0] - lo[0];
        const float dy = hi[1] - lo[1];
        const float dz = hi[2] - lo[2];
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct Bins {
    std::vector<Box> bbox;
    std::vector<int> bcnt, lcnt, rcnt;
    std::vector<float> larea, rarea;
    explicit Bins(uint32_t n) : bbox(n), bcnt(n), lcnt(n), rcnt(n), larea(n), rarea(n) {}
};

// Per-primitive arrays in the current permutation order: the partition
// swaps every one of them with perm, so pbox[i] is the box of prim perm[i]
// and every pass over a node's range reads memory sequentially.
struct BuildData {
    uint32_t bins, leaf;
    std::vector<uint32_t> &perm;
    std::vector<Box> pbox;
    std::vector<float> cent[3];
    std::vector<uint8_t> bin;  // bin of each primitive in the node being split
    unsigned threads;
    uint32_t par_min;          // ranges at least this large bin on all threads
};

constexpr uint32_t kLeafNode = 0xFFFFFFFFu;
constexpr uint32_t kParallelRange = 1u << 18;  // smallest range that bins on all threads
constexpr uint32_t kChunk = 1u << 15;

template <class F>
void parallel_for(uint32_t n, unsigned threads, F &&f);

// Bounds of the primitive boxes and of the centroids over [first, end).
void range_bounds(const BuildData &D, uint32_t first, uint32_t end, Box &pb, Box &cb) {
    const Box *__restrict__ bx = D.pbox.data();
    const float *__restrict__ x = D.cent[0].data();
    const float *__restrict__ y = D.cent[1].data();
    const float *__restrict__ z = D.cent[2].data();
    pb.clear();
    cb.clear();
    for (uint32_t i = first; i < end; i++) {
        pb.grow(bx[i]);
        const float c[3] = {x[i], y[i], z[i]};
        cb.grow(c);
    }
}

// Bin boxes and counts over [a, e) along the split axis; keeps each
// primitive's bin for the partition.
void bin_range(BuildData &D, uint32_t a, uint32_t e, const float *__restrict__ ca, float cmin, float scale,
               int last_bin, Box *__restrict__ bb, int *__restrict__ bc) {
    const Box *__restrict__ bx = D.pbox.data();
    uint8_t *__restrict__ bn = D.bin.data();
    for (int b = 0; b <= last_bin; b++) { bb[b].clear(); bc[b] = 0; }
    for (uint32_t i = a; i < e; i++) {
        int b = (int)((ca[i] - cmin) * scale);
        if (b < 0) b = 0;
        if (b > last_bin) b = last_bin;
        bn[i] = (uint8_t)b;
        bb[b].grow(bx[i]);
        bc[b]++;
    }
}

// One node of buildRecursive (bvh.cpp:100-231) over [first, first + count):
// its bounds go to nd; returns the split point (the left child is
// [first, mid)) with nd.axis set, or kLeafNode when the node is a leaf
// (nd then holds the leaf's range and count).  Bounds and bin boxes are
// min/max reductions and bin counts integer sums, so splitting those passes
// over threads gives the sequential results exactly; the partition runs on
// one thread in the reference's swap order.
uint32_t split_node(BuildData &D, Bins &B, uint32_t first, uint32_t count, Node2 &nd) {
    const float kTrav = 1.0f, kIsect = 1.5f;
    const uint32_t end = first + count;
    const uint32_t bins = D.bins;
    const bool par = D.threads > 1 && count >= D.par_min;
    const uint32_t nchunks = par ? (count + kChunk - 1) / kChunk : 1;
    Box pb, cb;
    if (par) {
        std::vector<Box> ppb(nchunks), pcb(nchunks);
        parallel_for(nchunks, D.threads, [&](uint32_t k, unsigned) {
            range_bounds(D, first + k * kChunk, std::min(end, first + (k + 1) * kChunk), ppb[k], pcb[k]);
        });
        pb.clear();
        cb.clear();
        for (uint32_t k = 0; k < nchunks; k++) {
            pb.grow(ppb[k]);
            cb.grow(pcb[k]);
        }
    } else {
        range_bounds(D, first, end, pb, cb);
    }
    for (int k = 0; k < 3; k++) { nd.bmin[k] = pb.lo[k]; nd.bmax[k] = pb.hi[k]; }
    auto make_leaf = [&]() {
        nd.first_or_left = first;
        nd.count = (uint16_t)std::min<uint32_t>(count, 0xFFFFu);
        nd.axis = 0;
        return kLeafNode;
    };
    if (count <= D.leaf) return make_leaf();

    const float ext[3] = {cb.hi[0] - cb.lo[0], cb.hi[1] - cb.lo[1], cb.hi[2] - cb.lo[2]};
    int axis = 0;
    if (ext[1] > ext[0]) axis = 1;
    if (ext[2] > ((axis == 0) ? ext[0] : ext[1])) axis = 2;
    const float cmin = cb.lo[axis];
    const float cext = ext[axis];
    if (cext < 1e-12f) return make_leaf();
    const float scale = (float)bins / cext;
    const int last_bin = (int)bins - 1;
    const float *ca = D.cent[axis].data();
    if (par) {
        std::vector<Box> cbb((size_t)nchunks * bins);
        std::vector<int> cbc((size_t)nchunks * bins);
        parallel_for(nchunks, D.threads, [&](uint32_t k, unsigned) {
            bin_range(D, first + k * kChunk, std::min(end, first + (k + 1) * kChunk), ca, cmin, scale, last_bin,
                      &cbb[(size_t)k * bins], &cbc[(size_t)k * bins]);
        });
        for (uint32_t b = 0; b < bins; b++) { B.bbox[b].clear(); B.bcnt[b] = 0; }
        for (uint32_t k = 0; k < nchunks; k++)
            for (uint32_t b = 0; b < bins; b++) {
                B.bbox[b].grow(cbb[(size_t)k * bins + b]);
                B.bcnt[b] += cbc[(size_t)k * bins + b];
            }
    } else {
        bin_range(D, first, end, ca, cmin, scale, last_bin, B.bbox.data(), B.bcnt.data());
    }
    {
        Box acc; acc.clear();
        int c = 0;
        for (int i = 0; i < last_bin; i++) {
            if (B.bcnt[i] > 0) acc.grow(B.bbox[i]);
            c += B.bcnt[i];
            B.larea[i] = (c > 0) ? acc.area() : 0.0f;
            B.lcnt[i] = c;
        }
        acc.clear();
        c = 0;
        for (int i = last_bin; i > 0; i--) {
            if (B.bcnt[i] > 0) acc.grow(B.bbox[i]);
            c += B.bcnt[i];
            B.rarea[i - 1] = (c > 0) ? acc.area() : 0.0f;
            B.rcnt[i - 1] = c;
        }
    }
    const float parea = pb.area();
    const float inv_parea = (parea > 0.0f) ? 1.0f / parea : 0.0f;
    float best = (float)count * kIsect;
    int best_bin = -1;
    for (int i = 0; i < last_bin; i++) {
        if (B.lcnt[i] == 0 || B.rcnt[i] == 0) continue;
        const float cost = kTrav + kIsect * inv_parea *
                                       (B.larea[i] * (float)B.lcnt[i] + B.rarea[i] * (float)B.rcnt[i]);
        if (cost < best) { best = cost; best_bin = i; }
    }
    if (best_bin < 0) return make_leaf();

    // in-place partition in the reference's order (bvh.cpp:190-206); the
    // bins were kept from the binning pass
    uint32_t mid = first;
    {
        uint32_t *__restrict__ pm = D.perm.data();
        Box *__restrict__ bx = D.pbox.data();
        float *__restrict__ x = D.cent[0].data();
        float *__restrict__ y = D.cent[1].data();
        float *__restrict__ z = D.cent[2].data();
        uint8_t *__restrict__ bn = D.bin.data();
        const uint8_t bb = (uint8_t)best_bin;
        for (uint32_t i = first; i < end; i++) {
            if (bn[i] <= bb) {
                std::swap(pm[i], pm[mid]);
                std::swap(bx[i], bx[mid]);
                std::swap(x[i], x[mid]);
                std::swap(y[i], y[mid]);
                std::swap(z[i], z[mid]);
                std::swap(bn[i], bn[mid]);
                mid++;
            }
        }
    }
    const uint32_t nl = mid - first, nr = count - nl;
    if (nl == 0 || nr == 0) return make_leaf();
    nd.count = 0;
    nd.axis = (uint16_t)axis;
    return mid;
}

struct Task {
    uint32_t node, first, count;
};

// The subtree of [first, first + count) by the reference's recursion, as an
// explicit work stack (left subtree first, sibling pairs allocated when
// their parent splits): root at nodes[0], the rest in the order the
// recursion allocates them.
void build_local(BuildData &D, Bins &B, uint32_t first, uint32_t count, std::vector<Node2> &nodes) {
    nodes.clear();
    nodes.push_back(Node2{});
    std::vector<Task> work;
    work.push_back({0, first, count});
    while (!work.empty()) {
        const Task tk = work.back();
        work.pop_back();
        Node2 nd{};
        const uint32_t mid = split_node(D, B, tk.first, tk.count, nd);
        if (mid == kLeafNode) {
            nodes[tk.node] = nd;
            continue;
        }
        const uint32_t left = (uint32_t)nodes.size();
        nd.first_or_left = left;
        nodes[tk.node] = nd;
        nodes.push_back(Node2{});
        nodes.push_back(Node2{});
        work.push_back({left + 1, mid, tk.first + tk.count - mid});  // right after left, as the recursion does
        work.push_back({left, tk.first, mid - tk.first});
    }
}

// Top of the tree, split on the calling thread down to ranges small enough
// to build as independent subtrees on worker threads.
struct TopNode {
    Node2 nd{};
    int left = -1, right = -1;  // children (internal node)
    int sub = -1;               // subtree index (a range built by a worker)
    uint32_t first = 0, count = 0;
};

int expand_top(BuildData &D, Bins &B, uint32_t first, uint32_t count, uint32_t cutoff, std::vector<TopNode> &top,
               std::vector<int> &subs) {
    const int id = (int)top.size();
    top.emplace_back();
    top[id].first = first;
    top[id].count = count;
    if (count <= cutoff) {
        top[id].sub = (int)subs.size();
        subs.push_back(id);
        return id;
    }
    Node2 nd{};
    const uint32_t mid = split_node(D, B, first, count, nd);
    top[id].nd = nd;
    if (mid == kLeafNode) return id;
    const int l = expand_top(D, B, first, mid - first, cutoff, top, subs);
    const int r = expand_top(D, B, mid, first + count - mid, cutoff, top, subs);
    top[id].left = l;
    top[id].right = r;
    return id;
}

size_t count_top(const std::vector<TopNode> &top, const std::vector<std::vector<Node2>> &sub_nodes, int id) {
    const TopNode &t = top[id];
    if (t.sub >= 0) return sub_nodes[t.sub].size();
    if (t.left < 0) return 1;
    return 1 + count_top(top, sub_nodes, t.left) + count_top(top, sub_nodes, t.right);
}

// Writes the top node id at out[at] and its descendants from out[alloc] on,
// in the order the sequential recursion allocates them.
void emit_top(const std::vector<TopNode> &top, const std::vector<std::vector<Node2>> &sub_nodes, int id,
              uint32_t at, uint32_t &alloc, std::vector<Node2> &out) {
    const TopNode &t = top[id];
    if (t.sub >= 0) {
        const std::vector<Node2> &S = sub_nodes[t.sub];
        const uint32_t base = alloc;  // local node k >= 1 -> base + k - 1
        auto remap = [&](Node2 nd) {
            if (nd.count == 0) nd.first_or_left = base + nd.first_or_left - 1;
            return nd;
        };
        out[at] = remap(S[0]);
        for (size_t k = 1; k < S.size(); k++) out[base + k - 1] = remap(S[k]);
        alloc += (uint32_t)S.size() - 1;
        return;
    }
    Node2 nd = t.nd;
    if (t.left < 0) {
        out[at] = nd;
        return;
    }
    const uint32_t l = alloc;
    alloc += 2;
    nd.first_or_left = l;
    out[at] = nd;
    emit_top(top, sub_nodes, t.left, l, alloc, out);
    emit_top(top, sub_nodes, t.right, l + 1, alloc, out);
}

unsigned build_threads() {
    const unsigned hw = std::thread::hardware_concurrency();
    return std::max(1u, std::min(16u, hw ? hw : 1u));
}

template <class F>
void parallel_for(uint32_t n, unsigned threads, F &&f) {
    if (threads <= 1 || n < 2) {
        for (uint32_t i = 0; i < n; i++) f(i, 0u);
        return;
    }
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < threads; t++)
        pool.emplace_back([&, t]() {
            for (uint32_t i; (i = next.fetch_add(1)) < n;) f(i, t);
        });
    for (std::thread &th : pool) th.join();
}

}  // namespace

void build_bvh2(const float *v9, uint32_t n, uint32_t bins, uint32_t leaf,
                std::vector<Node2> &nodes, std::vector<uint32_t> &perm) {
    nodes.clear();
    perm.assign(n, 0);
    if (n == 0) return;
    const unsigned threads = n >= (1u << 16) ? build_threads() : 1u;
    BuildData D{bins, leaf, perm, std::vector<Box>(n), {std::vector<float>(n), std::vector<float>(n),
                std::vector<float>(n)}, std::vector<uint8_t>(n), threads, kParallelRange};
    const uint32_t blocks = (n + kChunk - 1) / kChunk;
    parallel_for(blocks, threads, [&](uint32_t blk, unsigned) {
        const uint32_t end = std::min(n, (blk + 1) * kChunk);
        for (uint32_t i = blk * kChunk; i < end; i++) {
            const float *t = v9 + 9 * (size_t)i;
            Box &b = D.pbox[i];
            b.clear();
            b.grow(t);
            b.grow(t + 3);
            b.grow(t + 6);
            D.cent[0][i] = (b.lo[0] + b.hi[0]) * 0.5f;
            D.cent[1][i] = (b.lo[1] + b.hi[1]) * 0.5f;
            D.cent[2][i] = (b.lo[2] + b.hi[2]) * 0.5f;
            perm[i] = i;
        }
    });
    Bins B(bins);
    if (threads <= 1) {
        build_local(D, B, 0, n, nodes);
    } else {
        // the calling thread splits the top (binning on all threads); the
        // subtrees of at most `cutoff` primitives are built by the workers
        // (each range is touched by one thread only) and stitched in the
        // sequential allocation order
        const uint32_t cutoff = std::max<uint32_t>(4096, n / (8 * threads));
        D.par_min = std::max(kParallelRange, cutoff + 1);  // the workers' subtrees bin on their own thread
        std::vector<TopNode> top;
        std::vector<int> subs;
        expand_top(D, B, 0, n, cutoff, top, subs);
        std::vector<std::vector<Node2>> sub_nodes(subs.size());
        std::vector<uint32_t> order(subs.size());
        for (uint32_t i = 0; i < subs.size(); i++) order[i] = i;
        std::sort(order.begin(), order.end(),
                  [&](uint32_t a, uint32_t b) { return top[subs[a]].count > top[subs[b]].count; });
        std::vector<Bins> tb(threads, Bins(bins));
        parallel_for((uint32_t)subs.size(), threads, [&](uint32_t i, unsigned th) {
            const uint32_t k = order[i];
            const TopNode &t = top[subs[k]];
            build_local(D, tb[th], t.first, t.count, sub_nodes[k]);
        });
        nodes.resize(count_top(top, sub_nodes, 0));
        uint32_t alloc = 1;
        emit_top(top, sub_nodes, 0, 0, alloc, nodes);
    }
}

namespace {

// Packs a visit order (slots ord[0..W-1], first visited first) into a node.
inline void set_order(Node4 &nd, int oct, const int *ord) {
    nd.order[oct] = (uint8_t)(ord[0] | (ord[1] << 2) | (ord[2] << 4) | (ord[3] << 6));
    nd.rank[oct] = (uint8_t)(0 << (2 * ord[0]) | 1 << (2 * ord[1]) | 2 << (2 * ord[2]) | 3 << (2 * ord[3]));
}
inline void set_order(Node8 &nd, int oct, const int *ord) {
    uint32_t o = 0;
    for (int s = 0; s < 8; s++) o |= (uint32_t)ord[s] << (3 * s);
    nd.order[oct] = o;
}

// Folds binary levels into W-wide nodes: starting from a binary node, the
// internal child with the largest surface area is expanded until W slots are
// filled or only leaves remain.  Per octant, the slots are ordered the way
// the binary tree's near-child-first rule would reach them.
template <class NodeT, int W>
struct Collapse {
    const std::vector<Node2> &n2;
    std::vector<NodeT> &out;

    static float area(const Node2 &n) {
        const float dx = n.bmax[0] - n.bmin[0];
        const float dy = n.bmax[1] - n.bmin[1];
        const float dz = n.bmax[2] - n.bmin[2];
        return dx * dy + dy * dz + dz * dx;
    }

    // Depth-first order of the slots below binary node k for octant oct.
    void order(uint32_t k, int oct, const std::vector<uint32_t> &slots,
               const std::vector<uint32_t> &expanded, int *ord, int &m) const {
        if (std::find(expanded.begin(), expanded.end(), k) == expanded.end()) {
            for (int s = 0; s < (int)slots.size(); s++)
                if (slots[s] == k) { ord[m++] = s; return; }
            return;
        }
        const Node2 &nd = n2[k];
        const uint32_t l = nd.first_or_left, r = l + 1;
        if ((oct >> nd.axis) & 1) {
            order(r, oct, slots, expanded, ord, m);
            order(l, oct, slots, expanded, ord, m);
        } else {
            order(l, oct, slots, expanded, ord, m);
            order(r, oct, slots, expanded, ord, m);
        }
    }

    void fill(uint32_t idxw, uint32_t k2) {
        std::vector<uint32_t> slots = {n2[k2].first_or_left, n2[k2].first_or_left + 1};
        std::vector<uint32_t> expanded = {k2};
        while ((int)slots.size() < W) {
            int pick = -1;
            float best = -1.0f;
            for (int s = 0; s < (int)slots.size(); s++) {
                const Node2 &c = n2[slots[s]];
                if (c.count == 0 && area(c) > best) { best = area(c); pick = s; }
            }
            if (pick < 0) break;
            const uint32_t k = slots[pick];
            expanded.push_back(k);
            slots[pick] = n2[k].first_or_left;
            slots.insert(slots.begin() + pick + 1, n2[k].first_or_left + 1);
        }
        NodeT nd;
        std::memset(&nd, 0, sizeof nd);
        std::vector<std::pair<int, uint32_t>> inner;  // (slot, binary node)
        for (int s = 0; s < W; s++) {
            if (s >= (int)slots.size()) { nd.child[s] = kEmpty; continue; }
            const Node2 &c = n2[slots[s]];
            nd.lox[s] = c.bmin[0]; nd.loy[s] = c.bmin[1]; nd.loz[s] = c.bmin[2];
            nd.hix[s] = c.bmax[0]; nd.hiy[s] = c.bmax[1]; nd.hiz[s] = c.bmax[2];
            if (c.count > 0) nd.child[s] = kLeafBit | c.first_or_left;
            else inner.push_back({s, slots[s]});
        }
        for (int oct = 0; oct < 8; oct++) {
            int ord[W], m = 0;
            order(k2, oct, slots, expanded, ord, m);
            for (int s = 0; s < W; s++)
                if (std::find(ord, ord + m, s) == ord + m) ord[m++] = s;
            set_order(nd, oct, ord);
        }
        const uint32_t base = (uint32_t)out.size();
        out.resize(out.size() + inner.size());
        for (size_t i = 0; i < inner.size(); i++) nd.child[inner[i].first] = base + (uint32_t)i;
        out[idxw] = nd;
        for (size_t i = 0; i < inner.size(); i++) fill(base + (uint32_t)i, inner[i].second);
    }
};

// Moves the first `top` nodes in breadth-first order to the front (root
// first), keeping the others in their depth-first order, and renumbers the
// child refs.  Siblings stay contiguous either way.
template <class NodeT, int W>
void hoist_top(std::vector<NodeT> &nodes, uint32_t top) {
    const uint32_t n = (uint32_t)nodes.size();
    if (n <= 1) return;
    top = std::min(top, n);
    std::vector<uint32_t> bfs;
    bfs.reserve(top);
    bfs.push_back(0);
    for (size_t h = 0; h < bfs.size() && bfs.size() < top; h++) {
        const NodeT &nd = nodes[bfs[h]];
        for (int s = 0; s < W && bfs.size() < top; s++)
            if (nd.child[s] != kEmpty && !(nd.child[s] & kLeafBit)) bfs.push_back(nd.child[s]);
    }
    std::vector<uint32_t> to(n, ~0u);
    uint32_t next = 0;
    for (uint32_t k : bfs) to[k] = next++;
    for (uint32_t k = 0; k < n; k++)
        if (to[k] == ~0u) to[k] = next++;
    std::vector<NodeT> out(n);
    for (uint32_t k = 0; k < n; k++) {
        NodeT nd = nodes[k];
        for (int s = 0; s < W; s++)
            if (nd.child[s] != kEmpty && !(nd.child[s] & kLeafBit)) nd.child[s] = to[nd.child[s]];
        out[to[k]] = nd;
    }
    nodes.swap(out);
}

template <class NodeT, int W>
void collapse_wide(const std::vector<Node2> &nodes2, std::vector<NodeT> &out) {
    out.clear();
    if (nodes2.empty()) return;
    out.resize(1);
    if (nodes2[0].count > 0) {
        // A leaf root: one leaf child whose box is the root box.
        NodeT nd;
        std::memset(&nd, 0, sizeof nd);
        const Node2 &r = nodes2[0];
        nd.lox[0] = r.bmin[0]; nd.loy[0] = r.bmin[1]; nd.loz[0] = r.bmin[2];
        nd.hix[0] = r.bmax[0]; nd.hiy[0] = r.bmax[1]; nd.hiz[0] = r.bmax[2];
        nd.child[0] = kLeafBit | r.first_or_left;
        for (int s = 1; s < W; s++) nd.child[s] = kEmpty;
        int ord[W];
        for (int s = 0; s < W; s++) ord[s] = s;
        for (int oct = 0; oct < 8; oct++) set_order(nd, oct, ord);
        out[0] = nd;
        return;
    }
    Collapse<NodeT, W> c{nodes2, out};
    c.fill(0, 0);
    hoist_top<NodeT, W>(out, kTreeletBytes / (uint32_t)sizeof(NodeT));
}

// Deepest sum of (children - 1) pushed along any root-to-node path, plus one.
template <class NodeT, int W>
uint32_t stack_need_wide(const std::vector<NodeT> &nodes) {
    if (nodes.empty()) return 0;
    std::vector<std::pair<uint32_t, uint32_t>> st = {{0u, 0u}};
    uint32_t best = 0;
    while (!st.empty()) {
        auto [idx, acc] = st.back();
        st.pop_back();
        const NodeT &nd = nodes[idx];
        uint32_t valid = 0;
        for (int s = 0; s < W; s++) valid += nd.child[s] != kEmpty;
        const uint32_t here = acc + (valid > 0 ? valid - 1 : 0);
        best = std::max(best, here);
        for (int s = 0; s < W; s++)
            if (nd.child[s] != kEmpty && !(nd.child[s] & kLeafBit)) st.push_back({nd.child[s], here});
    }
    return best + 1;
}

}  // namespace

void collapse_bvh4(const std::vector<Node2> &nodes2, std::vector<Node4> &out) {
    collapse_wide<Node4, 4>(nodes2, out);
}

void collapse_bvh8(const std::vector<Node2> &nodes2, std::vector<Node8> &out) {
    collapse_wide<Node8, 8>(nodes2, out);
}

void make_tris(const float *v9, const std::vector<Node2> &nodes2,
               const std::vector<uint32_t> &perm, const std::vector<uint8_t> &pred_mask,
               std::vector<Tri48> &out) {
    const size_t n = perm.size();
    out.resize(n);
    for (size_t k = 0; k < n; k++) {
        const float *t = v9 + 9 * (size_t)perm[k];
        Tri48 &r = out[k];
        r.a[0] = t[0]; r.a[1] = t[1]; r.a[2] = t[2];
        std::memcpy(&r.a[3], &perm[k], 4);
        r.b[0] = t[3] - t[0]; r.b[1] = t[4] - t[1]; r.b[2] = t[5] - t[2];
        r.b[3] = 0.0f;
        r.c[0] = t[6] - t[0]; r.c[1] = t[7] - t[1]; r.c[2] = t[8] - t[2];
        // occluder bits per RG_PRED_* (all set when no materials are known)
        const uint32_t mask = pred_mask.empty() ? 0xFFu : (uint32_t)pred_mask[perm[k]];
        std::memcpy(&r.c[3], &mask, 4);
    }
    for (const Node2 &nd : nodes2) {
        if (nd.count == 0) continue;
        const size_t last = (size_t)nd.first_or_left + nd.count - 1;
        uint32_t flags;
        std::memcpy(&flags, &out[last].b[3], 4);
        flags |= 1u;
        std::memcpy(&out[last].b[3], &flags, 4);
    }
}

uint32_t stack_need4(const std::vector<Node4> &nodes) { return stack_need_wide<Node4, 4>(nodes); }
uint32_t stack_need8(const std::vector<Node8> &nodes) { return stack_need_wide<Node8, 8>(nodes); }

}  // namespace rg

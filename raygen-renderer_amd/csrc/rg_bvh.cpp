// rg_bvh.cpp — host BVH construction.
//
// build_bvh2 follows the reference builder TriangleBVH::build/buildRecursive
// (src/raygen/bvh.cpp:76-231) decision for decision: the same prim/centroid
// bounding boxes, widest-centroid-axis choice, 16-bin SAH with cost
// 1 + 1.5/A_p (A_L N_L + A_R N_R) against 1.5 N, the same forced-leaf rules and
// the same in-place partition, so it yields the same nodes in the same order
// and the same permutation.  It is written as an explicit work stack (DFS,
// left subtree first), which allocates sibling pairs in the same order as the
// reference's recursion.
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
#include <cfloat>
#include <cmath>
#include <cstring>

#include "rg_internal.h"

namespace rg {
namespace {

struct Box {
    float lo[3], hi[3];
    void clear() {
        for (int k = 0; k < 3; k++) { lo[k] = FLT_MAX; hi[k] = -FLT_MAX; }
    }
    void grow(const Box &b) {
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], b.lo[k]);
            hi[k] = fmaxf(hi[k], b.hi[k]);
        }
    }
    void grow(const float p[3]) {
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], p[k]);
            hi[k] = fmaxf(hi[k], p[k]);
        }
    }
    float area() const {
        const float dx = hi[0] - lo[0];
        const float dy = hi[1] - lo[1];
        const float dz = hi[2] - lo[2];
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct Task {
    uint32_t node, first, count;
};

}  // namespace

void build_bvh2(const float *v9, uint32_t n, uint32_t bins, uint32_t leaf,
                std::vector<Node2> &nodes, std::vector<uint32_t> &perm) {
    nodes.clear();
    perm.assign(n, 0);
    if (n == 0) return;
    const float kTrav = 1.0f, kIsect = 1.5f;

    std::vector<Box> pbox(n);
    std::vector<float> cx(n), cy(n), cz(n);
    for (uint32_t i = 0; i < n; i++) {
        const float *t = v9 + 9 * (size_t)i;
        pbox[i].clear();
        pbox[i].grow(t);
        pbox[i].grow(t + 3);
        pbox[i].grow(t + 6);
        cx[i] = (pbox[i].lo[0] + pbox[i].hi[0]) * 0.5f;
        cy[i] = (pbox[i].lo[1] + pbox[i].hi[1]) * 0.5f;
        cz[i] = (pbox[i].lo[2] + pbox[i].hi[2]) * 0.5f;
        perm[i] = i;
    }
    float *cent[3] = {cx.data(), cy.data(), cz.data()};

    nodes.reserve((size_t)n * 2);
    nodes.push_back(Node2{});
    std::vector<Box> bbox(bins);
    std::vector<int> bcnt(bins);
    std::vector<float> larea(bins), rarea(bins);
    std::vector<int> lcnt(bins), rcnt(bins);

    auto make_leaf = [&](uint32_t idx, uint32_t first, uint32_t count) {
        nodes[idx].first_or_left = first;
        nodes[idx].count = (uint16_t)std::min<uint32_t>(count, 0xFFFFu);
        nodes[idx].axis = 0;
    };

    std::vector<Task> work;
    work.push_back({0, 0, n});
    while (!work.empty()) {
        const Task tk = work.back();
        work.pop_back();
        const uint32_t first = tk.first, count = tk.count, end = tk.first + tk.count;

        Box pb, cb;
        pb.clear();
        cb.clear();
        for (uint32_t i = first; i < end; i++) {
            pb.grow(pbox[perm[i]]);
            const float c[3] = {cx[i], cy[i], cz[i]};
            cb.grow(c);
        }
        Node2 &nd = nodes[tk.node];
        for (int k = 0; k < 3; k++) { nd.bmin[k] = pb.lo[k]; nd.bmax[k] = pb.hi[k]; }

        if (count <= leaf) { make_leaf(tk.node, first, count); continue; }

        const float ext[3] = {cb.hi[0] - cb.lo[0], cb.hi[1] - cb.lo[1], cb.hi[2] - cb.lo[2]};
        int axis = 0;
        if (ext[1] > ext[0]) axis = 1;
        if (ext[2] > ((axis == 0) ? ext[0] : ext[1])) axis = 2;
        const float cmin = cb.lo[axis];
        const float cext = ext[axis];
        if (cext < 1e-12f) { make_leaf(tk.node, first, count); continue; }
        const float scale = (float)bins / cext;
        const int last_bin = (int)bins - 1;
        const float *ca = cent[axis];
        auto bin_of = [&](uint32_t i) {
            int b = (int)((ca[i] - cmin) * scale);
            if (b < 0) b = 0;
            if (b > last_bin) b = last_bin;
            return b;
        };

        for (uint32_t b = 0; b < bins; b++) { bbox[b].clear(); bcnt[b] = 0; }
        for (uint32_t i = first; i < end; i++) {
            const int b = bin_of(i);
            bbox[b].grow(pbox[perm[i]]);
            bcnt[b]++;
        }
        {
            Box acc; acc.clear();
            int c = 0;
            for (int i = 0; i < last_bin; i++) {
                if (bcnt[i] > 0) acc.grow(bbox[i]);
                c += bcnt[i];
                larea[i] = (c > 0) ? acc.area() : 0.0f;
                lcnt[i] = c;
            }
            acc.clear();
            c = 0;
            for (int i = last_bin; i > 0; i--) {
                if (bcnt[i] > 0) acc.grow(bbox[i]);
                c += bcnt[i];
                rarea[i - 1] = (c > 0) ? acc.area() : 0.0f;
                rcnt[i - 1] = c;
            }
        }
        const float parea = pb.area();
        const float inv_parea = (parea > 0.0f) ? 1.0f / parea : 0.0f;
        float best = (float)count * kIsect;
        int best_bin = -1;
        for (int i = 0; i < last_bin; i++) {
            if (lcnt[i] == 0 || rcnt[i] == 0) continue;
            const float cost = kTrav + kIsect * inv_parea *
                                           (larea[i] * (float)lcnt[i] + rarea[i] * (float)rcnt[i]);
            if (cost < best) { best = cost; best_bin = i; }
        }
        if (best_bin < 0) { make_leaf(tk.node, first, count); continue; }

        uint32_t mid = first;
        for (uint32_t i = first; i < end; i++) {
            if (bin_of(i) <= best_bin) {
                std::swap(perm[i], perm[mid]);
                std::swap(cx[i], cx[mid]);
                std::swap(cy[i], cy[mid]);
                std::swap(cz[i], cz[mid]);
                mid++;
            }
        }
        const uint32_t nl = mid - first, nr = count - nl;
        if (nl == 0 || nr == 0) { make_leaf(tk.node, first, count); continue; }

        const uint32_t left = (uint32_t)nodes.size();
        nodes.push_back(Node2{});
        nodes.push_back(Node2{});
        Node2 &p = nodes[tk.node];
        p.first_or_left = left;
        p.count = 0;
        p.axis = (uint16_t)axis;
        work.push_back({left + 1, mid, nr});  // right after left, as the recursion does
        work.push_back({left, first, nl});
    }
}

namespace {

// Packs a visit order (slots ord[0..W-1], first visited first) into a node.
inline void set_order(Node4 &nd, int oct, const int *ord) {
    nd.order[oct] = (uint8_t)(ord[0] | (ord[1] << 2) | (ord[2] << 4) | (ord[3] << 6));
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

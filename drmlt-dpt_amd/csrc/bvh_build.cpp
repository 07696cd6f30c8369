// Binned-SAH BVH builder (host, once per scene) replacing the reference's
// Embree build (scene.cpp:42-54). A binary SAH tree is built first and then
// collapsed into a 4-wide tree, which is what the gfx950 traversal reads:
// 8 float4 (128 bytes, one cache line) per node -- the four child boxes as
// structure of arrays (lo.x[4], hi.x[4], lo.y[4], hi.y[4], lo.z[4], hi.z[4])
// followed by child[4] and count[4]: count > 0 is a leaf whose triangles are
// [child, child + count) in the reordered triangle arrays, count == 0 an
// internal node, count < 0 an empty slot.
#include "bvh_build.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#ifndef DRMLT_BVH_BINS
#define DRMLT_BVH_BINS 32
#endif
#ifndef DRMLT_BVH_AXES
#define DRMLT_BVH_AXES 3
#endif

namespace drmlt {

namespace {
struct Box {
    float lo[3], hi[3];
    Box() { for (int k = 0; k < 3; k++) { lo[k] = 3e38f; hi[k] = -3e38f; } }
    void grow(const Box &b) { for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], b.lo[k]); hi[k] = std::max(hi[k], b.hi[k]); } }
    void grow(const float *p) { for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); } }
    float area() const {
        float d[3];
        for (int k = 0; k < 3; k++) d[k] = std::max(hi[k] - lo[k], 0.0f);
        return 2.0f * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
    }
};

// binary tree of the first pass
struct BNode {
    Box box[2];
    int index[2], count[2];  // count > 0: leaf range; 0: internal node index
};

struct Builder {
    const float *triP;
    std::vector<int> idx;
    std::vector<Box> tb;
    std::vector<float> cen;
    std::vector<BNode> bnodes;
    int maxDepth;
    int maxLeaf = 4;
    int depthReached = 0;

    static void pad(Box &b) {
        for (int k = 0; k < 3; k++) {
            float p = 1e-5f * (std::fabs(b.lo[k]) + std::fabs(b.hi[k])) + 1e-7f;
            b.lo[k] -= p;
            b.hi[k] += p;
        }
    }
    Box rangeBox(int first, int n) const {
        Box b;
        for (int i = first; i < first + n; i++) b.grow(tb[idx[i]]);
        return b;
    }
    // Returns (index, count): count > 0 for a leaf, count == 0 for an internal node.
    void build(int first, int n, int depth, int &outIndex, int &outCount, Box &outBox) {
        depthReached = std::max(depthReached, depth);
        outBox = rangeBox(first, n);
        if (n <= maxLeaf || depth >= maxDepth) {
            outIndex = first;
            outCount = n;
            return;
        }
        Box cb;
        for (int i = first; i < first + n; i++) cb.grow(&cen[3 * idx[i]]);
        int axis = 0;
        float ext = -1;
        for (int k = 0; k < 3; k++)
            if (cb.hi[k] - cb.lo[k] > ext) { ext = cb.hi[k] - cb.lo[k]; axis = k; }
        // binned SAH over the centroid extent of every axis (the longest
        // axis alone with DRMLT_BVH_AXES=1)
        constexpr int NB = DRMLT_BVH_BINS;
        float best = 3e38f;
        int bestAxis = -1, bestSplit = -1;
        for (int k = 0; k < 3; k++) {
            if (DRMLT_BVH_AXES == 1 && k != axis) continue;
            const float ek = cb.hi[k] - cb.lo[k];
            if (!(ek > 0)) continue;
            Box bb[NB];
            int bc[NB] = {0};
            const float scale = NB / ek;
            for (int i = first; i < first + n; i++) {
                const int t = idx[i];
                const int b = std::min(std::max(int((cen[3 * t + k] - cb.lo[k]) * scale), 0), NB - 1);
                bc[b]++;
                bb[b].grow(tb[t]);
            }
            Box lacc[NB];
            int lcnt[NB];
            Box acc;
            int cnt = 0;
            for (int b = 0; b < NB - 1; b++) {
                acc.grow(bb[b]);
                cnt += bc[b];
                lacc[b] = acc;
                lcnt[b] = cnt;
            }
            acc = Box();
            cnt = 0;
            for (int b = NB - 1; b > 0; b--) {
                acc.grow(bb[b]);
                cnt += bc[b];
                if (lcnt[b - 1] == 0 || cnt == 0) continue;
                const float cost = lacc[b - 1].area() * lcnt[b - 1] + acc.area() * cnt;
                if (cost < best) { best = cost; bestAxis = k; bestSplit = b; }
            }
        }
        if (bestSplit <= 0) {
            // every centroid equal (no bin split leaves both sides
            // non-empty): a leaf of all of them, which the device build
            // (bvh_device.hip) makes too
            outIndex = first;
            outCount = n;
            return;
        }
        const float scale = NB / (cb.hi[bestAxis] - cb.lo[bestAxis]);
        auto it = std::partition(idx.begin() + first, idx.begin() + first + n, [&](int t) {
            return std::min(std::max(int((cen[3 * t + bestAxis] - cb.lo[bestAxis]) * scale), 0), NB - 1) < bestSplit;
        });
        const int mid = int(it - idx.begin());
        const int node = int(bnodes.size());
        bnodes.emplace_back();
        int li, lc, ri, rc;
        Box lb, rb;
        build(first, mid - first, depth + 1, li, lc, lb);
        build(mid, first + n - mid, depth + 1, ri, rc, rb);
        pad(lb);
        pad(rb);
        BNode &bn = bnodes[node];
        bn.box[0] = lb; bn.box[1] = rb;
        bn.index[0] = li; bn.index[1] = ri;
        bn.count[0] = lc; bn.count[1] = rc;
        outIndex = node;
        outCount = 0;
    }

    // 4-wide collapse: a node takes the two children of binary node b and
    // repeatedly opens its internal child of largest area until it holds four.
    struct Slot { Box box; int index, count; };
    std::vector<float> nodes;  // 32 floats per 4-wide node
    int depth4 = 0, stackNeed = 0;
    int collapse(int b, int depth, int pushed) {
        Slot s[4];
        int n = 2;
        for (int k = 0; k < 2; k++) s[k] = Slot{bnodes[b].box[k], bnodes[b].index[k], bnodes[b].count[k]};
        while (n < 4) {
            int open = -1;
            float best = -1;
            for (int k = 0; k < n; k++)
                if (s[k].count == 0 && s[k].box.area() > best) { best = s[k].box.area(); open = k; }
            if (open < 0) break;
            const BNode &c = bnodes[s[open].index];
            s[open] = Slot{c.box[0], c.index[0], c.count[0]};
            s[n++] = Slot{c.box[1], c.index[1], c.count[1]};
        }
        const int me = int(nodes.size() / 32);
        nodes.resize(nodes.size() + 32);
        depth4 = std::max(depth4, depth);
        // a traversal visiting this node pushes at most n - 1 siblings
        const int pushHere = pushed + (n - 1);
        stackNeed = std::max(stackNeed, pushHere);
        int child[4], count[4];
        for (int k = 0; k < 4; k++) {
            if (k >= n) { child[k] = 0; count[k] = -1; continue; }
            if (s[k].count > 0) { child[k] = s[k].index; count[k] = s[k].count; }
            else { child[k] = collapse(s[k].index, depth + 1, pushHere); count[k] = 0; }
        }
        float *nd = &nodes[32 * size_t(me)];
        for (int k = 0; k < 4; k++) {
            const Box e;  // empty slot: an inverted box
            const Box &bx = k < n ? s[k].box : e;
            nd[0 + k] = bx.lo[0]; nd[4 + k] = bx.hi[0];
            nd[8 + k] = bx.lo[1]; nd[12 + k] = bx.hi[1];
            nd[16 + k] = bx.lo[2]; nd[20 + k] = bx.hi[2];
        }
        std::memcpy(nd + 24, child, 16);
        std::memcpy(nd + 28, count, 16);
        return me;
    }
};
}  // namespace

int BuildBvh(int nt, const float *triP, int maxDepth, BvhOut &out) {
    Builder b;
    b.triP = triP;
    b.maxDepth = maxDepth;
    b.idx.resize(nt);
    b.tb.resize(nt);
    b.cen.resize(3 * size_t(nt));
    for (int i = 0; i < nt; i++) {
        b.idx[i] = i;
        Box bx;
        for (int v = 0; v < 3; v++) bx.grow(triP + 9 * size_t(i) + 3 * v);
        b.tb[i] = bx;
        for (int k = 0; k < 3; k++) b.cen[3 * size_t(i) + k] = 0.5f * (bx.lo[k] + bx.hi[k]);
    }
    b.bnodes.reserve(size_t(nt) / 2 + 2);
    int rc = nt;
    if (nt > b.maxLeaf) {
        int ri;
        Box rb;
        b.build(0, nt, 0, ri, rc, rb);
    }
    if (rc > 0) {
        // the whole scene is one leaf (a few triangles, or all centroids
        // equal): a root whose two children are that leaf (testing a
        // triangle twice gives the same hit)
        BNode r;
        Box lb = b.rangeBox(0, nt);
        Builder::pad(lb);
        r.box[0] = r.box[1] = lb;
        r.index[0] = r.index[1] = 0;
        r.count[0] = r.count[1] = nt;
        b.bnodes.assign(1, r);
    }
    b.nodes.reserve(b.bnodes.size() * 32 / 2 + 32);
    b.collapse(0, 0, 0);
    out.nodes = std::move(b.nodes);
    out.order = std::move(b.idx);
    out.depth = b.depthReached;
    out.depth4 = b.depth4;
    out.stackNeed = b.stackNeed;
    return 0;
}

}  // namespace drmlt

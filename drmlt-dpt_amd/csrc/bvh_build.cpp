// Binned-SAH BVH builder (host, once per scene) replacing the reference's
// Embree build (scene.cpp:42-54). Output layout is what the gfx950 traversal
// reads: 4 float4 per internal node -- both child boxes plus
// (left, right, leftCount, rightCount); a child with count > 0 is a leaf whose
// triangles are [index, index + count) in the reordered triangle arrays.
// Tree depth is capped below the device traversal stack (kBvhStack).
#include "bvh_build.h"

#include <algorithm>
#include <cmath>
#include <cstring>

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

struct Builder {
    const float *triP;
    std::vector<int> idx;
    std::vector<Box> tb;
    std::vector<float> cen;
    std::vector<float> nodes;  // 16 floats per node
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
        int mid = first + n / 2;
        if (ext > 0) {
            const int NB = 16;
            Box bb[NB];
            int bc[NB] = {0};
            const float scale = NB / ext;
            auto binOf = [&](int t) {
                int b = int((cen[3 * t + axis] - cb.lo[axis]) * scale);
                return std::min(std::max(b, 0), NB - 1);
            };
            for (int i = first; i < first + n; i++) {
                int b = binOf(idx[i]);
                bc[b]++;
                bb[b].grow(tb[idx[i]]);
            }
            float best = 3e38f;
            int bestSplit = -1;
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
                float cost = lacc[b - 1].area() * lcnt[b - 1] + acc.area() * cnt;
                if (cost < best) { best = cost; bestSplit = b; }
            }
            if (bestSplit > 0) {
                auto it = std::partition(idx.begin() + first, idx.begin() + first + n, [&](int t) { return binOf(t) < bestSplit; });
                mid = int(it - idx.begin());
                if (mid == first || mid == first + n) mid = first + n / 2;
            }
        }
        if (mid == first + n / 2) {
            std::nth_element(idx.begin() + first, idx.begin() + mid, idx.begin() + first + n,
                             [&](int a, int b) { return cen[3 * a + axis] < cen[3 * b + axis]; });
        }
        int node = int(nodes.size() / 16);
        nodes.resize(nodes.size() + 16);
        int li, lc, ri, rc;
        Box lb, rb;
        build(first, mid - first, depth + 1, li, lc, lb);
        build(mid, first + n - mid, depth + 1, ri, rc, rb);
        pad(lb);
        pad(rb);
        float *nd = &nodes[16 * node];
        nd[0] = lb.lo[0]; nd[1] = lb.lo[1]; nd[2] = lb.lo[2]; nd[3] = lb.hi[0];
        nd[4] = lb.hi[1]; nd[5] = lb.hi[2]; nd[6] = rb.lo[0]; nd[7] = rb.lo[1];
        nd[8] = rb.lo[2]; nd[9] = rb.hi[0]; nd[10] = rb.hi[1]; nd[11] = rb.hi[2];
        int ch[4] = {li, ri, lc, rc};
        std::memcpy(nd + 12, ch, 16);
        outIndex = node;
        outCount = 0;
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
    b.nodes.reserve(size_t(nt) * 16 / 2 + 16);
    if (nt <= b.maxLeaf) {
        // a root whose two children are the same leaf (testing a triangle twice
        // gives the same hit; every child slot of a node is then a leaf)
        b.nodes.resize(16);
        Box lb = b.rangeBox(0, nt);
        Builder::pad(lb);
        float *nd = &b.nodes[0];
        nd[0] = lb.lo[0]; nd[1] = lb.lo[1]; nd[2] = lb.lo[2]; nd[3] = lb.hi[0];
        nd[4] = lb.hi[1]; nd[5] = lb.hi[2];
        nd[6] = lb.lo[0]; nd[7] = lb.lo[1]; nd[8] = lb.lo[2]; nd[9] = lb.hi[0]; nd[10] = lb.hi[1]; nd[11] = lb.hi[2];
        int ch[4] = {0, 0, nt, nt};
        std::memcpy(nd + 12, ch, 16);
    } else {
        int ri, rc;
        Box rb;
        b.build(0, nt, 0, ri, rc, rb);
    }
    out.nodes = std::move(b.nodes);
    out.order = std::move(b.idx);
    out.depth = b.depthReached;
    return 0;
}

}  // namespace drmlt

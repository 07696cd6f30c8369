// Device BVH build replacing the reference's Embree build (scene.cpp:42-54):
// the binned-SAH tree of the host builder (bvh_build.cpp: 32 bins over the
// centroid extent of each axis, leaves of at most 4 triangles, a depth cap,
// child boxes padded) built top-down one tree level per launch -- one
// 256-thread workgroup per node, or, while a level has fewer than 256 nodes,
// one workgroup per 8192-triangle chunk of a node -- then collapsed into the
// 4-wide layout the traversal reads one 4-wide level per launch. The split
// decisions use the host builder's arithmetic in the same order (no
// contraction), so the tree
// has the same nodes; nodes are numbered in allocation order instead of depth
// first, which leaves every query's answer unchanged (the traversal's closest
// hit breaks ties by scene order).
#include "bvh_device.h"

#include <chrono>
#include <vector>

namespace drmlt {
namespace {

constexpr int kBins = 32;
constexpr int kLeaf = 4;
constexpr int kWG = 256;

// one binary node: two child slots (padded box lo.xyz hi.xyz, index, count:
// count > 0 leaf range [index, index + count), 0 internal node `index`)
struct BNodeD {
    float box[2][6];
    int index[2], count[2];
};
// a node to build: the triangle range [first, first + count) of the level's
// index array, its depth, and the parent slot it fills (2 node + side; -1: root)
struct BTask {
    int first, count, depth, parent;
};
// a 4-wide node to write from binary node b
struct CTask {
    int b, depth, pushed, out;
};

// order-preserving float <-> int for LDS atomic min / max
__device__ inline int ford(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ inline float fdec(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7fffffff); }

__device__ inline float wmin(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ inline float wmax(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// Box::area of the host builder, without contraction
__device__ inline float box_area(const float *lo, const float *hi) {
#pragma clang fp contract(off)
    float d[3];
    for (int k = 0; k < 3; k++) d[k] = fmaxf(hi[k] - lo[k], 0.0f);
    return 2.0f * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]);
}
// the host builder's split cost: area x count of both sides
__device__ inline float sah_cost(const float *llo, const float *lhi, int lc, const float *rlo, const float *rhi, int rc) {
#pragma clang fp contract(off)
    return box_area(llo, lhi) * float(lc) + box_area(rlo, rhi) * float(rc);
}
// Builder::pad of the host builder
__device__ inline void box_pad(float *lo, float *hi) {
#pragma clang fp contract(off)
    for (int k = 0; k < 3; k++) {
        const float p = 1e-5f * (fabsf(lo[k]) + fabsf(hi[k])) + 1e-7f;
        lo[k] -= p;
        hi[k] += p;
    }
}
__device__ inline int bin_of(float c, float lo, float scale) {
#pragma clang fp contract(off)
    return min(max(int((c - lo) * scale), 0), kBins - 1);
}

// The split of a node from its bins (counts cnt[k * kBins + b], boxes
// bb[(k * kBins + b) * 6 + j] as ford values), by one wave: the host
// builder's choice -- over the axes with extent and the splits b = 31 .. 1
// (left bins below b), the first in that order of the lowest cost
// area x count of both sides -- with each candidate's boxes and counts (the
// host builder's running minima, maxima and sums, which do not depend on the
// order they are taken in) from scans over the bins, one bin per lane, and
// the first lowest cost found by a wave reduction over (cost, position in
// that order). axis / split -1: no split leaves both sides non-empty.
__device__ void sah_split_wave(int lane, const float *ext, const int *cnt3, const int *bb3, int &axisOut, int &splitOut) {
    constexpr int kCand = 3 * (kBins - 1);
    static_assert(kBins == 32, "one bin per lane of a half wave");
    float bestC = 3e38f;
    int bestI = kCand;
    // pass 0: axes 0 and 1 on the two half waves; pass 1: axis 2. Lane b of a
    // half wave holds bin b; running minima, maxima and sums over the bins
    // below and from b come from scans within the half wave.
    for (int pass = 0; pass < 2; pass++) {
        const int k = 2 * pass + (lane >> 5), b = lane & 31;
        const bool act = k < 3;
        const float e = k == 0 ? ext[0] : k == 1 ? ext[1] : k == 2 ? ext[2] : 0.0f;
        float plo[3], phi[3], slo[3], shi[3];
        int pc = act ? cnt3[k * kBins + b] : 0, sc = pc;
        for (int j = 0; j < 3; j++) {
            plo[j] = act ? fdec(bb3[(k * kBins + b) * 6 + j]) : 3e38f;
            phi[j] = act ? fdec(bb3[(k * kBins + b) * 6 + 3 + j]) : -3e38f;
            slo[j] = plo[j];
            shi[j] = phi[j];
        }
        for (int d = 1; d < 32; d <<= 1) {
            const bool up = b >= d, down = b + d < 32;
            const int tpc = __shfl_up(pc, d, 32), tsc = __shfl_down(sc, d, 32);
            if (up) pc += tpc;
            if (down) sc += tsc;
            for (int j = 0; j < 3; j++) {
                const float a0 = __shfl_up(plo[j], d, 32), a1 = __shfl_up(phi[j], d, 32);
                const float c0 = __shfl_down(slo[j], d, 32), c1 = __shfl_down(shi[j], d, 32);
                if (up) { plo[j] = fminf(plo[j], a0); phi[j] = fmaxf(phi[j], a1); }
                if (down) { slo[j] = fminf(slo[j], c0); shi[j] = fmaxf(shi[j], c1); }
            }
        }
        // the split before bin b: left = bins 0 .. b-1 (lane b-1's prefix), right = bins b .. 31
        float llo[3], lhi[3];
        const int lc = __shfl_up(pc, 1, 32);
        for (int j = 0; j < 3; j++) { llo[j] = __shfl_up(plo[j], 1, 32); lhi[j] = __shfl_up(phi[j], 1, 32); }
        if (act && e > 0 && b >= 1 && lc != 0 && sc != 0) {
            const float cost = sah_cost(llo, lhi, lc, slo, shi, sc);
            const int c = k * (kBins - 1) + (kBins - 1 - b);
            if (cost < 3e38f && (cost < bestC || (cost == bestC && c < bestI))) { bestC = cost; bestI = c; }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const float oc = __shfl_xor(bestC, o);
        const int oi = __shfl_xor(bestI, o);
        if (oc < bestC || (oc == bestC && oi < bestI)) { bestC = oc; bestI = oi; }
    }
    axisOut = bestI < kCand ? bestI / (kBins - 1) : -1;
    splitOut = bestI < kCand ? kBins - 1 - bestI % (kBins - 1) : -1;
}

// Per triangle: box of its corners and its centroid.
__global__ void __launch_bounds__(256) k_bvh_prep(int nt, const float *triP, float4 *blo, float4 *bhi, float4 *cen, int *idx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    const float *p = triP + 9 * size_t(i);
    float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
    for (int v = 0; v < 3; v++)
        for (int k = 0; k < 3; k++) {
            lo[k] = fminf(lo[k], p[3 * v + k]);
            hi[k] = fmaxf(hi[k], p[3 * v + k]);
        }
    blo[i] = make_float4(lo[0], lo[1], lo[2], 0.0f);
    bhi[i] = make_float4(hi[0], hi[1], hi[2], 0.0f);
    cen[i] = make_float4(0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2]), 0.0f);
    idx[i] = i;
}

// One tree level: workgroup t builds task t (Builder::build of the host
// builder, one call).
__global__ void __launch_bounds__(kWG) k_bvh_level(const BTask *tasks, const int *idxIn, int *idxOut, int *order, const float4 *blo,
                                                  const float4 *bhi, const float4 *cen, BNodeD *bn, int *nodeCount, BTask *next,
                                                  int *nextCount, int maxDepth, int *depthMax, int cap, int *overflow) {
    const BTask T = tasks[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ float red[12][kWG / 64];
    __shared__ float box[12];  // range box lo, hi; centroid box lo, hi
    __shared__ int bcnt[3][kBins];
    __shared__ int bbox[3][kBins][6];
    __shared__ int dec[4];     // axis, split, mid, node
    __shared__ int wsum[2][kWG / 64];
    // 1. the range's box and its centroids' box
    float v[12];
    for (int k = 0; k < 3; k++) { v[k] = 3e38f; v[3 + k] = -3e38f; v[6 + k] = 3e38f; v[9 + k] = -3e38f; }
    for (int i = tid; i < T.count; i += kWG) {
        const int t = idxIn[T.first + i];
        const float4 a = blo[t], b = bhi[t], c = cen[t];
        const float al[3] = {a.x, a.y, a.z}, bh[3] = {b.x, b.y, b.z}, cc[3] = {c.x, c.y, c.z};
        for (int k = 0; k < 3; k++) {
            v[k] = fminf(v[k], al[k]); v[3 + k] = fmaxf(v[3 + k], bh[k]);
            v[6 + k] = fminf(v[6 + k], cc[k]); v[9 + k] = fmaxf(v[9 + k], cc[k]);
        }
    }
    for (int k = 0; k < 12; k++) {
        const bool isMin = (k % 6) < 3;
        const float r = isMin ? wmin(v[k]) : wmax(v[k]);
        if (lane == 0) red[k][wave] = r;
    }
    __syncthreads();
    if (tid < 12) {
        const bool isMin = (tid % 6) < 3;
        float r = red[tid][0];
        for (int w = 1; w < kWG / 64; w++) r = isMin ? fminf(r, red[tid][w]) : fmaxf(r, red[tid][w]);
        box[tid] = r;
    }
    __syncthreads();
    bool leaf = T.count <= kLeaf || T.depth >= maxDepth;
    if (tid == 0) {
        atomicMax(depthMax, T.depth);
        if (T.parent >= 0) {
            float lo[3] = {box[0], box[1], box[2]}, hi[3] = {box[3], box[4], box[5]};
            box_pad(lo, hi);
            BNodeD &P = bn[T.parent >> 1];
            const int side = T.parent & 1;
            for (int k = 0; k < 3; k++) { P.box[side][k] = lo[k]; P.box[side][3 + k] = hi[k]; }
        }
    }
    if (leaf) {
        for (int i = tid; i < T.count; i += kWG) order[T.first + i] = idxIn[T.first + i];
        if (tid == 0 && T.parent >= 0) {
            bn[T.parent >> 1].index[T.parent & 1] = T.first;
            bn[T.parent >> 1].count[T.parent & 1] = T.count;
        }
        if (tid == 0 && T.parent < 0) atomicAdd(overflow + 1, 1);  // the root is one leaf: the host builder's case
        return;
    }
    // 2. bins over the centroid extent of each axis
    const float clo[3] = {box[6], box[7], box[8]}, chi[3] = {box[9], box[10], box[11]};
    float ext[3], scale[3];
    for (int k = 0; k < 3; k++) {
        ext[k] = chi[k] - clo[k];
        scale[k] = ext[k] > 0 ? float(kBins) / ext[k] : 0.0f;
    }
    for (int i = tid; i < 3 * kBins; i += kWG) {
        const int k = i / kBins, b = i % kBins;
        bcnt[k][b] = 0;
        for (int j = 0; j < 3; j++) { bbox[k][b][j] = ford(3e38f); bbox[k][b][3 + j] = ford(-3e38f); }
    }
    __syncthreads();
    for (int i = tid; i < T.count; i += kWG) {
        const int t = idxIn[T.first + i];
        const float4 a = blo[t], bb = bhi[t], c = cen[t];
        const float al[3] = {a.x, a.y, a.z}, bh[3] = {bb.x, bb.y, bb.z}, cc[3] = {c.x, c.y, c.z};
        for (int k = 0; k < 3; k++) {
            if (!(ext[k] > 0)) continue;
            const int b = bin_of(cc[k], clo[k], scale[k]);
            atomicAdd(&bcnt[k][b], 1);
            for (int j = 0; j < 3; j++) {
                atomicMin(&bbox[k][b][j], ford(al[j]));
                atomicMax(&bbox[k][b][3 + j], ford(bh[j]));
            }
        }
    }
    __syncthreads();
    // 3. the split, by the first wave
    if (wave == 0) {
        int ax, sp;
        sah_split_wave(lane, ext, &bcnt[0][0], &bbox[0][0][0], ax, sp);
        if (tid == 0) { dec[0] = ax; dec[1] = sp; }
    }
    __syncthreads();
    const int axis = dec[0], split = dec[1];
    // no bin split leaves both sides non-empty (every centroid equal): a leaf
    // of all of them, as the host builder makes
    leaf = split <= 0;
    if (leaf) {
        for (int i = tid; i < T.count; i += kWG) order[T.first + i] = idxIn[T.first + i];
        if (tid == 0 && T.parent >= 0) {
            bn[T.parent >> 1].index[T.parent & 1] = T.first;
            bn[T.parent >> 1].count[T.parent & 1] = T.count;
        }
        if (tid == 0 && T.parent < 0) atomicAdd(overflow + 1, 1);  // the root is one leaf: the host builder's case
        return;
    }
    // 4. stable partition into idxOut: left of the split, then right
    int mid;
    {
        int nl = 0;
        for (int i = tid; i < T.count; i += kWG) {
            const float4 c = cen[idxIn[T.first + i]];
            const float cc = axis == 0 ? c.x : axis == 1 ? c.y : c.z;
            nl += bin_of(cc, clo[axis], scale[axis]) < split ? 1 : 0;
        }
        for (int o = 32; o > 0; o >>= 1) nl += __shfl_xor(nl, o);
        if (lane == 0) wsum[0][wave] = nl;
        __syncthreads();
        int nLeft = 0;
        for (int w = 0; w < kWG / 64; w++) nLeft += wsum[0][w];
        mid = T.first + nLeft;
        __syncthreads();
        int lbase = 0, rbase = 0;
        for (int c0 = 0; c0 < T.count; c0 += kWG) {
            const int i = c0 + tid;
            int t = -1;
            bool left = false;
            if (i < T.count) {
                t = idxIn[T.first + i];
                const float4 c = cen[t];
                const float cc = axis == 0 ? c.x : axis == 1 ? c.y : c.z;
                left = bin_of(cc, clo[axis], scale[axis]) < split;
            }
            const unsigned long long ml = __ballot(left), mr = __ballot(i < T.count && !left);
            const unsigned long long below = (1ull << lane) - 1ull;
            if (lane == 0) { wsum[0][wave] = __popcll(ml); wsum[1][wave] = __popcll(mr); }
            __syncthreads();
            int lo = 0, ro = 0, lt = 0, rt = 0;
            for (int w = 0; w < kWG / 64; w++) {
                if (w < wave) { lo += wsum[0][w]; ro += wsum[1][w]; }
                lt += wsum[0][w]; rt += wsum[1][w];
            }
            if (i < T.count) {
                if (left) idxOut[T.first + lbase + lo + __popcll(ml & below)] = t;
                else idxOut[mid + rbase + ro + __popcll(mr & below)] = t;
            }
            lbase += lt;
            rbase += rt;
            __syncthreads();
        }
    }
    // 5. this node and its two children's tasks
    if (tid == 0) {
        const int node = T.parent < 0 ? 0 : atomicAdd(nodeCount, 1);
        const int q = atomicAdd(nextCount, 2);
        if (node >= cap || q + 1 >= cap) {
            atomicAdd(overflow, 1);  // not reached: a split leaves more than four triangles per internal node
        } else {
            if (T.parent >= 0) {
                bn[T.parent >> 1].index[T.parent & 1] = node;
                bn[T.parent >> 1].count[T.parent & 1] = 0;
            }
            next[q] = BTask{T.first, mid - T.first, T.depth + 1, 2 * node};
            next[q + 1] = BTask{mid, T.first + T.count - mid, T.depth + 1, 2 * node + 1};
        }
    }
}

// ---- levels with many small nodes (the bottom of the tree): one wave per
// node, four per workgroup, the steps of k_bvh_level with wave reductions and
// a wave-local partition (the same values and the same order)
constexpr int kWaveTasks = 4096;  // levels with at least this many nodes are built this way
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
// Internal nodes are not numbered here: flag[t] = 1 and mid[t] record them,
// and k_bvh_emit numbers them in task order after a scan of the flags (one
// atomic counter per node would serialise hundreds of thousands of nodes).
__global__ void __launch_bounds__(kWG) k_bvh_level_wave(const BTask *tasks, int ntasks, const int *idxIn, int *idxOut, int *order,
                                                       const float4 *blo, const float4 *bhi, const float4 *cen, BNodeD *bn, int *flag, int *midOut,
                                                       int maxDepth, int *depthMax, int *overflow) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = int(blockIdx.x) * (kWG / 64) + wave;
    __shared__ int bcntS[kWG / 64][3][kBins];
    __shared__ int bboxS[kWG / 64][3][kBins][6];
    if (t >= ntasks) return;  // the whole wave; nothing below waits on other waves
    const BTask T = tasks[t];
    int (*bcnt)[kBins] = bcntS[wave];
    int (*bbox)[kBins][6] = bboxS[wave];
    // 1. the range's box and its centroids' box
    float v[12];
    for (int k = 0; k < 3; k++) { v[k] = 3e38f; v[3 + k] = -3e38f; v[6 + k] = 3e38f; v[9 + k] = -3e38f; }
    for (int i = lane; i < T.count; i += 64) {
        const int q = idxIn[T.first + i];
        const float4 a = blo[q], b = bhi[q], c = cen[q];
        const float al[3] = {a.x, a.y, a.z}, bh[3] = {b.x, b.y, b.z}, cc[3] = {c.x, c.y, c.z};
        for (int k = 0; k < 3; k++) {
            v[k] = fminf(v[k], al[k]); v[3 + k] = fmaxf(v[3 + k], bh[k]);
            v[6 + k] = fminf(v[6 + k], cc[k]); v[9 + k] = fmaxf(v[9 + k], cc[k]);
        }
    }
    float box[12];
    for (int k = 0; k < 12; k++) box[k] = (k % 6) < 3 ? wmin(v[k]) : wmax(v[k]);
    bool leaf = T.count <= kLeaf || T.depth >= maxDepth;
    if (lane == 0) {
        if (t == 0) atomicMax(depthMax, T.depth);  // every node of a level has the same depth
        if (T.parent >= 0) {
            float lo[3] = {box[0], box[1], box[2]}, hi[3] = {box[3], box[4], box[5]};
            box_pad(lo, hi);
            BNodeD &P = bn[T.parent >> 1];
            const int side = T.parent & 1;
            for (int k = 0; k < 3; k++) { P.box[side][k] = lo[k]; P.box[side][3 + k] = hi[k]; }
        }
    }
    int axis = -1, split = -1;
    float clo[3], ext[3], scale[3];
    if (!leaf) {
        // 2. bins over the centroid extent of each axis
        for (int k = 0; k < 3; k++) {
            clo[k] = box[6 + k];
            ext[k] = box[9 + k] - clo[k];
            scale[k] = ext[k] > 0 ? float(kBins) / ext[k] : 0.0f;
        }
        for (int i = lane; i < 3 * kBins; i += 64) {
            const int k = i / kBins, b = i % kBins;
            bcnt[k][b] = 0;
            for (int j = 0; j < 3; j++) { bbox[k][b][j] = ford(3e38f); bbox[k][b][3 + j] = ford(-3e38f); }
        }
        wave_sync();
        for (int i = lane; i < T.count; i += 64) {
            const int q = idxIn[T.first + i];
            const float4 a = blo[q], bb = bhi[q], c = cen[q];
            const float al[3] = {a.x, a.y, a.z}, bh[3] = {bb.x, bb.y, bb.z}, cc[3] = {c.x, c.y, c.z};
            for (int k = 0; k < 3; k++) {
                if (!(ext[k] > 0)) continue;
                const int b = bin_of(cc[k], clo[k], scale[k]);
                atomicAdd(&bcnt[k][b], 1);
                for (int j = 0; j < 3; j++) {
                    atomicMin(&bbox[k][b][j], ford(al[j]));
                    atomicMax(&bbox[k][b][3 + j], ford(bh[j]));
                }
            }
        }
        wave_sync();
        // 3. the split
        sah_split_wave(lane, ext, &bcnt[0][0], &bbox[0][0][0], axis, split);
        leaf = split <= 0;
    }
    if (leaf) {
        for (int i = lane; i < T.count; i += 64) order[T.first + i] = idxIn[T.first + i];
        if (lane == 0) {
            flag[t] = 0;
            if (T.parent >= 0) {
                bn[T.parent >> 1].index[T.parent & 1] = T.first;
                bn[T.parent >> 1].count[T.parent & 1] = T.count;
            } else {
                atomicAdd(overflow + 1, 1);  // the root is one leaf: the host builder's case
            }
        }
        return;
    }
    // 4. stable partition into idxOut: left of the split, then right
    auto isLeft = [&](int q) {
        const float4 c = cen[q];
        const float cc = axis == 0 ? c.x : axis == 1 ? c.y : c.z;
        return bin_of(cc, clo[axis], scale[axis]) < split;
    };
    int nl = 0;
    for (int i = lane; i < T.count; i += 64) nl += isLeft(idxIn[T.first + i]) ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) nl += __shfl_xor(nl, o);
    const int mid = T.first + nl;
    int lbase = 0, rbase = 0;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int c0 = 0; c0 < T.count; c0 += 64) {
        const int i = c0 + lane;
        int q = -1;
        bool left = false;
        if (i < T.count) {
            q = idxIn[T.first + i];
            left = isLeft(q);
        }
        const unsigned long long ml = __ballot(left), mr = __ballot(i < T.count && !left);
        if (i < T.count) {
            if (left) idxOut[T.first + lbase + __popcll(ml & below)] = q;
            else idxOut[mid + rbase + __popcll(mr & below)] = q;
        }
        lbase += __popcll(ml);
        rbase += __popcll(mr);
    }
    // 5. an internal node: numbered by k_bvh_emit
    if (lane == 0) {
        flag[t] = 1;
        midOut[t] = mid;
    }
}

// Exclusive scan of the 0/1 flags of n tasks in blocks of kScanBlock:
// position within the block (local) and each block's count (blockSum).
constexpr int kScanBlock = 1024;
__global__ void __launch_bounds__(kScanBlock) k_scan_local(int n, const int *flag, int *local, int *blockSum) {
    __shared__ int ws[kScanBlock / 64];
    const int i = int(blockIdx.x) * kScanBlock + int(threadIdx.x), lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool f = i < n && flag[i] != 0;
    const unsigned long long m = __ballot(f);
    if (lane == 0) ws[wave] = __popcll(m);
    __syncthreads();
    int off = 0;
    for (int w = 0; w < wave; w++) off += ws[w];
    if (i < n) local[i] = off + __popcll(m & ((1ull << lane) - 1ull));
    if (threadIdx.x == 0) {
        int s = 0;
        for (int w = 0; w < kScanBlock / 64; w++) s += ws[w];
        blockSum[blockIdx.x] = s;
    }
}
// The blocks' exclusive offsets in place and the total (one thread: at most
// maxTasks / kScanBlock blocks).
__global__ void k_scan_sums(int nb, int *blockSum, int *total) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int s = 0;
    for (int b = 0; b < nb; b++) {
        const int v = blockSum[b];
        blockSum[b] = s;
        s += v;
    }
    *total = s;
}
// The internal nodes of a wave-built level in task order: node
// nodeCount + rank, children tasks 2 rank and 2 rank + 1.
__global__ void __launch_bounds__(256) k_bvh_emit(const BTask *tasks, int ntasks, const int *flag, const int *midIn, const int *local,
                                                 const int *blockOff, const int *nodeCount, BNodeD *bn, BTask *next, int cap, int *overflow) {
    const int t = int(blockIdx.x) * 256 + int(threadIdx.x);
    if (t >= ntasks || !flag[t]) return;
    const BTask T = tasks[t];
    const int rank = blockOff[t / kScanBlock] + local[t], node = *nodeCount + rank, mid = midIn[t];
    if (node >= cap || 2 * rank + 1 >= cap || T.parent < 0) {
        atomicAdd(overflow, 1);  // not reached: the root is built by the chunked levels
        return;
    }
    bn[T.parent >> 1].index[T.parent & 1] = node;
    bn[T.parent >> 1].count[T.parent & 1] = 0;
    next[2 * rank] = BTask{T.first, mid - T.first, T.depth + 1, 2 * node};
    next[2 * rank + 1] = BTask{mid, T.first + T.count - mid, T.depth + 1, 2 * node + 1};
}
// Counters after an emitted level: nodes += total, next tasks = 2 total.
__global__ void k_bvh_emit_counts(const int *total, int *nodeCount, int *nextCount) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    *nodeCount += *total;
    *nextCount = 2 * *total;
}

// ---- levels with few nodes (the top of the tree): several workgroups per
// node, each over a chunk of its triangle range; per-node boxes and bins meet
// in global memory through atomic min / max / add, which give the values
// the one-workgroup kernel computes, and the partition keeps the chunks'
// order, so the tree is the same.
constexpr int kChunk = 8192;    // triangles per workgroup
constexpr int kWideTasks = 256; // levels with fewer nodes are built this way
// a chunk [first, first + count) of task `task`, whose first chunk is chunk `firstChunk`
struct WChunk {
    int task, first, count, firstChunk;
};
// per task: range and centroid boxes (12 ford values), bins (counts and ford boxes), decision
struct WTask {
    int box[12];
    int cnt[3][kBins];
    int bb[3][kBins][6];
    int axis, split, mid, leaf;
};

__global__ void __launch_bounds__(256) k_wide_init(int ntasks, WTask *w) {
    const int t = blockIdx.x;
    if (t >= ntasks) return;
    WTask &W = w[t];
    for (int i = threadIdx.x; i < 12; i += blockDim.x) W.box[i] = ford((i % 6) < 3 ? 3e38f : -3e38f);
    for (int i = threadIdx.x; i < 3 * kBins; i += blockDim.x) {
        const int k = i / kBins, b = i % kBins;
        W.cnt[k][b] = 0;
        for (int j = 0; j < 3; j++) { W.bb[k][b][j] = ford(3e38f); W.bb[k][b][3 + j] = ford(-3e38f); }
    }
}

// 1. the range's box and its centroids' box, chunk by chunk
__global__ void __launch_bounds__(kWG) k_wide_box(const WChunk *chunks, const int *idxIn, const float4 *blo, const float4 *bhi, const float4 *cen,
                                                 WTask *w) {
    const WChunk C = chunks[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ float red[12][kWG / 64];
    float v[12];
    for (int k = 0; k < 3; k++) { v[k] = 3e38f; v[3 + k] = -3e38f; v[6 + k] = 3e38f; v[9 + k] = -3e38f; }
    for (int i = tid; i < C.count; i += kWG) {
        const int t = idxIn[C.first + i];
        const float4 a = blo[t], b = bhi[t], c = cen[t];
        const float al[3] = {a.x, a.y, a.z}, bh[3] = {b.x, b.y, b.z}, cc[3] = {c.x, c.y, c.z};
        for (int k = 0; k < 3; k++) {
            v[k] = fminf(v[k], al[k]); v[3 + k] = fmaxf(v[3 + k], bh[k]);
            v[6 + k] = fminf(v[6 + k], cc[k]); v[9 + k] = fmaxf(v[9 + k], cc[k]);
        }
    }
    for (int k = 0; k < 12; k++) {
        const bool isMin = (k % 6) < 3;
        const float r = isMin ? wmin(v[k]) : wmax(v[k]);
        if (lane == 0) red[k][wave] = r;
    }
    __syncthreads();
    if (tid < 12) {
        const bool isMin = (tid % 6) < 3;
        float r = red[tid][0];
        for (int q = 1; q < kWG / 64; q++) r = isMin ? fminf(r, red[tid][q]) : fmaxf(r, red[tid][q]);
        if (isMin) atomicMin(&w[C.task].box[tid], ford(r));
        else atomicMax(&w[C.task].box[tid], ford(r));
    }
}

// the bin parameters of a task from its centroid box (the one-workgroup
// kernel's arithmetic)
__device__ inline void wide_bin_params(const WTask &W, float *clo, float *ext, float *scale) {
    for (int k = 0; k < 3; k++) {
        clo[k] = fdec(W.box[6 + k]);
        const float chi = fdec(W.box[9 + k]);
        ext[k] = chi - clo[k];
        scale[k] = ext[k] > 0 ? float(kBins) / ext[k] : 0.0f;
    }
}

// 2. bins over the centroid extent of each axis, chunk by chunk
__global__ void __launch_bounds__(kWG) k_wide_bins(const WChunk *chunks, const int *idxIn, const float4 *blo, const float4 *bhi, const float4 *cen,
                                                  WTask *w) {
    const WChunk C = chunks[blockIdx.x];
    const int tid = threadIdx.x;
    __shared__ int bcnt[3][kBins];
    __shared__ int bbox[3][kBins][6];
    WTask &W = w[C.task];
    float clo[3], ext[3], scale[3];
    wide_bin_params(W, clo, ext, scale);
    for (int i = tid; i < 3 * kBins; i += kWG) {
        const int k = i / kBins, b = i % kBins;
        bcnt[k][b] = 0;
        for (int j = 0; j < 3; j++) { bbox[k][b][j] = ford(3e38f); bbox[k][b][3 + j] = ford(-3e38f); }
    }
    __syncthreads();
    for (int i = tid; i < C.count; i += kWG) {
        const int t = idxIn[C.first + i];
        const float4 a = blo[t], bb = bhi[t], c = cen[t];
        const float al[3] = {a.x, a.y, a.z}, bh[3] = {bb.x, bb.y, bb.z}, cc[3] = {c.x, c.y, c.z};
        for (int k = 0; k < 3; k++) {
            if (!(ext[k] > 0)) continue;
            const int b = bin_of(cc[k], clo[k], scale[k]);
            atomicAdd(&bcnt[k][b], 1);
            for (int j = 0; j < 3; j++) {
                atomicMin(&bbox[k][b][j], ford(al[j]));
                atomicMax(&bbox[k][b][3 + j], ford(bh[j]));
            }
        }
    }
    __syncthreads();
    for (int i = tid; i < 3 * kBins; i += kWG) {
        const int k = i / kBins, b = i % kBins;
        if (bcnt[k][b] == 0) continue;
        atomicAdd(&W.cnt[k][b], bcnt[k][b]);
        for (int j = 0; j < 3; j++) {
            atomicMin(&W.bb[k][b][j], bbox[k][b][j]);
            atomicMax(&W.bb[k][b][3 + j], bbox[k][b][3 + j]);
        }
    }
}

// 3. per task (one wave): the parent's box slot, a leaf, or the split, this
// node and its two children's tasks (steps 3 and 5 of the one-workgroup
// kernel); the left count comes from the bins the partition uses
__global__ void __launch_bounds__(64) k_wide_split(const BTask *tasks, int ntasks, const int *idxIn, int *order, WTask *w, BNodeD *bn,
                                                   int *nodeCount, BTask *next, int *nextCount, int maxDepth, int *depthMax, int cap,
                                                   int *overflow) {
    const int t = blockIdx.x, tid = threadIdx.x;
    if (t >= ntasks) return;
    const BTask T = tasks[t];
    WTask &W = w[t];
    __shared__ int dec[2];
    float clo[3], ext[3], scale[3];
    wide_bin_params(W, clo, ext, scale);
    bool leaf = T.count <= kLeaf || T.depth >= maxDepth;
    if (tid == 0) {
        atomicMax(depthMax, T.depth);
        if (T.parent >= 0) {
            float lo[3] = {fdec(W.box[0]), fdec(W.box[1]), fdec(W.box[2])}, hi[3] = {fdec(W.box[3]), fdec(W.box[4]), fdec(W.box[5])};
            box_pad(lo, hi);
            BNodeD &P = bn[T.parent >> 1];
            const int side = T.parent & 1;
            for (int k = 0; k < 3; k++) { P.box[side][k] = lo[k]; P.box[side][3 + k] = hi[k]; }
        }
        dec[0] = -1;
        dec[1] = -1;
    }
    if (!leaf) {
        int ax, sp;
        sah_split_wave(tid, ext, &W.cnt[0][0], &W.bb[0][0][0], ax, sp);
        if (tid == 0) { dec[0] = ax; dec[1] = sp; }
    }
    __syncthreads();
    const int axis = dec[0], split = dec[1];
    leaf = leaf || split <= 0;
    if (leaf) {
        for (int i = tid; i < T.count; i += 64) order[T.first + i] = idxIn[T.first + i];
        if (tid == 0) {
            if (T.parent >= 0) {
                bn[T.parent >> 1].index[T.parent & 1] = T.first;
                bn[T.parent >> 1].count[T.parent & 1] = T.count;
            } else {
                atomicAdd(overflow + 1, 1);  // the root is one leaf: the host builder's case
            }
            W.leaf = 1;
        }
        return;
    }
    if (tid == 0) {
        int nLeft = 0;
        for (int b = 0; b < split; b++) nLeft += W.cnt[axis][b];
        const int mid = T.first + nLeft;
        W.axis = axis;
        W.split = split;
        W.mid = mid;
        W.leaf = 0;
        const int node = T.parent < 0 ? 0 : atomicAdd(nodeCount, 1);
        const int q = atomicAdd(nextCount, 2);
        if (node >= cap || q + 1 >= cap) {
            atomicAdd(overflow, 1);  // not reached: a split leaves more than four triangles per internal node
        } else {
            if (T.parent >= 0) {
                bn[T.parent >> 1].index[T.parent & 1] = node;
                bn[T.parent >> 1].count[T.parent & 1] = 0;
            }
            next[q] = BTask{T.first, mid - T.first, T.depth + 1, 2 * node};
            next[q + 1] = BTask{mid, T.first + T.count - mid, T.depth + 1, 2 * node + 1};
        }
    }
}

__device__ inline bool wide_left(const WTask &W, const float *clo, const float *scale, const float4 &c) {
    const float cc = W.axis == 0 ? c.x : W.axis == 1 ? c.y : c.z;
    return bin_of(cc, clo[W.axis], scale[W.axis]) < W.split;
}

// 4a. left count of each chunk
__global__ void __launch_bounds__(kWG) k_wide_count(const WChunk *chunks, const int *idxIn, const float4 *cen, const WTask *w, int *leftCnt) {
    const WChunk C = chunks[blockIdx.x];
    const WTask &W = w[C.task];
    if (W.leaf) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ int ws[kWG / 64];
    float clo[3], ext[3], scale[3];
    wide_bin_params(W, clo, ext, scale);
    int nl = 0;
    for (int i = tid; i < C.count; i += kWG) nl += wide_left(W, clo, scale, cen[idxIn[C.first + i]]) ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) nl += __shfl_xor(nl, o);
    if (lane == 0) ws[wave] = nl;
    __syncthreads();
    if (tid == 0) {
        int s = 0;
        for (int q = 0; q < kWG / 64; q++) s += ws[q];
        leftCnt[blockIdx.x] = s;
    }
}

// 4b. stable partition into idxOut, each chunk after the task's earlier chunks
__global__ void __launch_bounds__(kWG) k_wide_scatter(const WChunk *chunks, const BTask *tasks, const int *idxIn, int *idxOut, const float4 *cen,
                                                     const WTask *w, const int *leftCnt) {
    const WChunk C = chunks[blockIdx.x];
    const WTask &W = w[C.task];
    if (W.leaf) return;
    const BTask T = tasks[C.task];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ int wsum[2][kWG / 64];
    float clo[3], ext[3], scale[3];
    wide_bin_params(W, clo, ext, scale);
    int lbase = 0;
    for (int j = C.firstChunk; j < int(blockIdx.x); j++) lbase += leftCnt[j];
    int rbase = (C.first - T.first) - lbase;
    for (int c0 = 0; c0 < C.count; c0 += kWG) {
        const int i = c0 + tid;
        int t = -1;
        bool left = false;
        if (i < C.count) {
            t = idxIn[C.first + i];
            left = wide_left(W, clo, scale, cen[t]);
        }
        const unsigned long long ml = __ballot(left), mr = __ballot(i < C.count && !left);
        const unsigned long long below = (1ull << lane) - 1ull;
        if (lane == 0) { wsum[0][wave] = __popcll(ml); wsum[1][wave] = __popcll(mr); }
        __syncthreads();
        int lo = 0, ro = 0, lt = 0, rt = 0;
        for (int q = 0; q < kWG / 64; q++) {
            if (q < wave) { lo += wsum[0][q]; ro += wsum[1][q]; }
            lt += wsum[0][q]; rt += wsum[1][q];
        }
        if (i < C.count) {
            if (left) idxOut[T.first + lbase + lo + __popcll(ml & below)] = t;
            else idxOut[W.mid + rbase + ro + __popcll(mr & below)] = t;
        }
        lbase += lt;
        rbase += rt;
        __syncthreads();
    }
}

// One 4-wide level: thread t writes the 4-wide node of task t (the host
// builder's collapse: open the internal child of largest area until four
// slots are used).
__global__ void __launch_bounds__(256) k_bvh_collapse(const CTask *tasks, int n, const BNodeD *bn, int nb, float4 *nodes, int *nodeCount,
                                                     CTask *next, int *nextCount, int *depthMax, int *stackMax, int *overflow) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const CTask T = tasks[t];
    float sb[4][6];
    int si[4], sc[4];
    int ns = 2;
    for (int k = 0; k < 2; k++) {
        for (int j = 0; j < 6; j++) sb[k][j] = bn[T.b].box[k][j];
        si[k] = bn[T.b].index[k];
        sc[k] = bn[T.b].count[k];
    }
    while (ns < 4) {
        int open = -1;
        float best = -1;
        for (int k = 0; k < ns; k++) {
            const float a = box_area(sb[k], sb[k] + 3);
            if (sc[k] == 0 && a > best) { best = a; open = k; }
        }
        if (open < 0) break;
        if (si[open] < 0 || si[open] >= nb) { atomicAdd(overflow, 1); return; }  // not reached: every slot names a built node
        const BNodeD &c = bn[si[open]];
        for (int j = 0; j < 6; j++) { sb[open][j] = c.box[0][j]; sb[ns][j] = c.box[1][j]; }
        si[ns] = c.index[1]; sc[ns] = c.count[1];
        si[open] = c.index[0]; sc[open] = c.count[0];
        ns++;
    }
    atomicMax(depthMax, T.depth);
    const int pushHere = T.pushed + (ns - 1);
    atomicMax(stackMax, pushHere);
    int internal = 0;
    for (int k = 0; k < ns; k++) internal += sc[k] == 0 ? 1 : 0;
    const int base = internal ? atomicAdd(nodeCount, internal) : 0;
    const int qb = internal ? atomicAdd(nextCount, internal) : 0;
    if (base + internal > nb || qb + internal > nb || T.out >= nb) { atomicAdd(overflow, 1); return; }  // not reached
    int child[4], count[4];
    int j = 0;
    for (int k = 0; k < 4; k++) {
        if (k >= ns) { child[k] = 0; count[k] = -1; continue; }
        if (sc[k] > 0) { child[k] = si[k]; count[k] = sc[k]; continue; }
        child[k] = base + j;
        count[k] = 0;
        next[qb + j] = CTask{si[k], T.depth + 1, pushHere, base + j};
        j++;
    }
    float nd[32];
    for (int k = 0; k < 4; k++) {
        const bool used = k < ns;
        nd[0 + k] = used ? sb[k][0] : 3e38f; nd[4 + k] = used ? sb[k][3] : -3e38f;
        nd[8 + k] = used ? sb[k][1] : 3e38f; nd[12 + k] = used ? sb[k][4] : -3e38f;
        nd[16 + k] = used ? sb[k][2] : 3e38f; nd[20 + k] = used ? sb[k][5] : -3e38f;
        nd[24 + k] = __int_as_float(child[k]);
        nd[28 + k] = __int_as_float(count[k]);
    }
    float4 *o = nodes + 8 * size_t(T.out);
    for (int q = 0; q < 8; q++) o[q] = make_float4(nd[4 * q], nd[4 * q + 1], nd[4 * q + 2], nd[4 * q + 3]);
}

// Leaf-order triangle records: p0 with the scene index in .w, e1, e2.
__global__ void __launch_bounds__(256) k_bvh_tris(int nt, const int *order, const float4 *s0, const float4 *s1, const float4 *s2, float4 *tri) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nt) return;
    const int i = min(max(order[k], 0), nt - 1);  // every position is written; the clamp keeps a defect from reading out of bounds
    const float4 p = s0[i];
    tri[3 * size_t(k)] = make_float4(p.x, p.y, p.z, __int_as_float(i));
    tri[3 * size_t(k) + 1] = s1[i];
    tri[3 * size_t(k) + 2] = s2[i];
}

}  // namespace

hipError_t BuildBvhDevice(int nt, const float *triP, const float4 *s0, const float4 *s1, const float4 *s2, int maxDepth, hipStream_t s,
                          DeviceBvh &out) {
    const auto t0 = std::chrono::steady_clock::now();
    out = DeviceBvh();
    if (nt <= kLeaf) return hipErrorInvalidValue;  // the host builder's single-leaf root covers these
    hipError_t e = hipSuccess;
    float4 *blo = nullptr, *bhi = nullptr, *cen = nullptr;
    int *idxA = nullptr, *idxB = nullptr, *order = nullptr, *counters = nullptr;
    BTask *tA = nullptr, *tB = nullptr;
    BNodeD *bn = nullptr;
    CTask *cA = nullptr, *cB = nullptr;
    WTask *wt = nullptr;
    WChunk *wc = nullptr;
    int *leftCnt = nullptr;
    int *wflag = nullptr, *wmid = nullptr, *wlocal = nullptr, *wblock = nullptr;
    std::vector<BTask> hostTasks;
    std::vector<WChunk> hostChunks;
    // binary nodes: at most nt - 1 (every split leaves both sides non-empty)
    const size_t maxB = size_t(nt), maxTasks = size_t(nt), maxChunks = size_t(nt) / kChunk + kWideTasks + 1;
#define BVH_CHK(x) do { e = (x); if (e != hipSuccess) goto done; } while (0)
    BVH_CHK(hipMalloc(&blo, sizeof(float4) * nt));
    BVH_CHK(hipMalloc(&bhi, sizeof(float4) * nt));
    BVH_CHK(hipMalloc(&cen, sizeof(float4) * nt));
    BVH_CHK(hipMalloc(&idxA, sizeof(int) * nt));
    BVH_CHK(hipMalloc(&idxB, sizeof(int) * nt));
    BVH_CHK(hipMalloc(&order, sizeof(int) * nt));
    BVH_CHK(hipMalloc(&tA, sizeof(BTask) * maxTasks));
    BVH_CHK(hipMalloc(&tB, sizeof(BTask) * maxTasks));
    BVH_CHK(hipMalloc(&bn, sizeof(BNodeD) * maxB));
    BVH_CHK(hipMalloc(&counters, sizeof(int) * 8));  // 0 binary nodes, 1 next tasks, 2 depth, 3 4-wide nodes, 4 depth4, 5 stack, 6 overflow
    BVH_CHK(hipMalloc(&wt, sizeof(WTask) * kWideTasks));
    BVH_CHK(hipMalloc(&wc, sizeof(WChunk) * maxChunks));
    BVH_CHK(hipMalloc(&leftCnt, sizeof(int) * maxChunks));
    BVH_CHK(hipMalloc(&wflag, sizeof(int) * maxTasks));
    BVH_CHK(hipMalloc(&wmid, sizeof(int) * maxTasks));
    BVH_CHK(hipMalloc(&wlocal, sizeof(int) * maxTasks));
    BVH_CHK(hipMalloc(&wblock, sizeof(int) * (maxTasks / kScanBlock + 2)));
    {
        BVH_CHK(hipMemsetAsync(order, 0, sizeof(int) * nt, s));
        BVH_CHK(hipMemsetAsync(bn, 0, sizeof(BNodeD) * maxB, s));
        hipLaunchKernelGGL(k_bvh_prep, dim3((nt + 255) / 256), dim3(256), 0, s, nt, triP, blo, bhi, cen, idxA);
        BVH_CHK(hipGetLastError());
        const int init[8] = {1, 0, 0, 1, 0, 0, 0, 0};  // the root is binary node 0 and 4-wide node 0
        BVH_CHK(hipMemcpyAsync(counters, init, sizeof(init), hipMemcpyHostToDevice, s));
        const BTask root{0, nt, 0, -1};
        BVH_CHK(hipMemcpyAsync(tA, &root, sizeof(root), hipMemcpyHostToDevice, s));
        int ntasks = 1, overflow = 0;
        for (int level = 0; ntasks > 0; level++) {
            if (level > maxDepth + 1) { e = hipErrorUnknown; goto done; }  // not reached: tasks at the depth cap are leaves
            BVH_CHK(hipMemsetAsync(counters + 1, 0, sizeof(int), s));
            if (ntasks < kWideTasks) {
                // few nodes: each over chunks of kChunk triangles, one workgroup per chunk
                hostTasks.resize(size_t(ntasks));
                BVH_CHK(hipMemcpyAsync(hostTasks.data(), tA, sizeof(BTask) * size_t(ntasks), hipMemcpyDeviceToHost, s));
                BVH_CHK(hipStreamSynchronize(s));
                hostChunks.clear();
                for (int t = 0; t < ntasks; t++) {
                    const BTask &T = hostTasks[size_t(t)];
                    const int first = int(hostChunks.size());
                    for (int c0 = 0; c0 < T.count; c0 += kChunk) hostChunks.push_back(WChunk{t, T.first + c0, std::min(kChunk, T.count - c0), first});
                }
                const int nch = int(hostChunks.size());
                if (size_t(nch) > maxChunks || nch == 0) { e = hipErrorUnknown; goto done; }  // not reached: the counts sum to at most nt
                BVH_CHK(hipMemcpyAsync(wc, hostChunks.data(), sizeof(WChunk) * size_t(nch), hipMemcpyHostToDevice, s));
                hipLaunchKernelGGL(k_wide_init, dim3(ntasks), dim3(256), 0, s, ntasks, wt);
                hipLaunchKernelGGL(k_wide_box, dim3(nch), dim3(kWG), 0, s, wc, idxA, blo, bhi, cen, wt);
                hipLaunchKernelGGL(k_wide_bins, dim3(nch), dim3(kWG), 0, s, wc, idxA, blo, bhi, cen, wt);
                hipLaunchKernelGGL(k_wide_split, dim3(ntasks), dim3(64), 0, s, tA, ntasks, idxA, order, wt, bn, counters, tB, counters + 1,
                                   maxDepth, counters + 2, int(maxTasks), counters + 6);
                hipLaunchKernelGGL(k_wide_count, dim3(nch), dim3(kWG), 0, s, wc, idxA, cen, wt, leftCnt);
                hipLaunchKernelGGL(k_wide_scatter, dim3(nch), dim3(kWG), 0, s, wc, tA, idxA, idxB, cen, wt, leftCnt);
            } else if (ntasks >= kWaveTasks) {
                // many small nodes: one wave each, internal nodes numbered in task order by a scan
                const int nb = (ntasks + kScanBlock - 1) / kScanBlock;
                hipLaunchKernelGGL(k_bvh_level_wave, dim3((ntasks + kWG / 64 - 1) / (kWG / 64)), dim3(kWG), 0, s, tA, ntasks, idxA, idxB, order,
                                   blo, bhi, cen, bn, wflag, wmid, maxDepth, counters + 2, counters + 6);
                hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(kScanBlock), 0, s, ntasks, wflag, wlocal, wblock);
                hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(64), 0, s, nb, wblock, wblock + nb);
                hipLaunchKernelGGL(k_bvh_emit, dim3((ntasks + 255) / 256), dim3(256), 0, s, tA, ntasks, wflag, wmid, wlocal, wblock, counters, bn, tB,
                                   int(maxTasks), counters + 6);
                hipLaunchKernelGGL(k_bvh_emit_counts, dim3(1), dim3(64), 0, s, wblock + nb, counters, counters + 1);
            } else {
                hipLaunchKernelGGL(k_bvh_level, dim3(ntasks), dim3(kWG), 0, s, tA, idxA, idxB, order, blo, bhi, cen, bn, counters, tB,
                                   counters + 1, maxDepth, counters + 2, int(maxTasks), counters + 6);
            }
            BVH_CHK(hipGetLastError());
            BVH_CHK(hipMemcpyAsync(&ntasks, counters + 1, sizeof(int), hipMemcpyDeviceToHost, s));
            BVH_CHK(hipMemcpyAsync(&overflow, counters + 6, sizeof(int), hipMemcpyDeviceToHost, s));
            BVH_CHK(hipStreamSynchronize(s));
            if (overflow || size_t(ntasks) > maxTasks) { e = hipErrorUnknown; goto done; }
            std::swap(tA, tB);
            std::swap(idxA, idxB);
        }
        int rootLeaf = 0;
        BVH_CHK(hipMemcpy(&rootLeaf, counters + 7, sizeof(int), hipMemcpyDeviceToHost));
        if (rootLeaf) { e = hipErrorInvalidValue; goto done; }
        int nb = 0;
        BVH_CHK(hipMemcpy(&nb, counters, sizeof(int), hipMemcpyDeviceToHost));
        // 4-wide collapse: at most nb 4-wide nodes
        BVH_CHK(hipMalloc(&out.nodes, sizeof(float4) * 8 * size_t(nb)));
        BVH_CHK(hipMalloc(&cA, sizeof(CTask) * size_t(nb)));
        BVH_CHK(hipMalloc(&cB, sizeof(CTask) * size_t(nb)));
        const CTask croot{0, 0, 0, 0};
        BVH_CHK(hipMemcpyAsync(cA, &croot, sizeof(croot), hipMemcpyHostToDevice, s));
        int nc = 1;
        for (int level = 0; nc > 0; level++) {
            if (level > maxDepth + 1) { e = hipErrorUnknown; goto done; }  // not reached: 4-wide levels never outnumber binary ones
            BVH_CHK(hipMemsetAsync(counters + 1, 0, sizeof(int), s));
            hipLaunchKernelGGL(k_bvh_collapse, dim3((nc + 255) / 256), dim3(256), 0, s, cA, nc, bn, nb, out.nodes, counters + 3, cB,
                               counters + 1, counters + 4, counters + 5, counters + 6);
            BVH_CHK(hipGetLastError());
            BVH_CHK(hipMemcpyAsync(&nc, counters + 1, sizeof(int), hipMemcpyDeviceToHost, s));
            BVH_CHK(hipMemcpyAsync(&overflow, counters + 6, sizeof(int), hipMemcpyDeviceToHost, s));
            BVH_CHK(hipStreamSynchronize(s));
            if (overflow || nc > nb) { e = hipErrorUnknown; goto done; }
            std::swap(cA, cB);
        }
        BVH_CHK(hipMalloc(&out.tri, sizeof(float4) * 3 * size_t(nt)));
        hipLaunchKernelGGL(k_bvh_tris, dim3((nt + 255) / 256), dim3(256), 0, s, nt, order, s0, s1, s2, out.tri);
        BVH_CHK(hipGetLastError());
        int c[8];
        BVH_CHK(hipMemcpyAsync(c, counters, sizeof(c), hipMemcpyDeviceToHost, s));
        BVH_CHK(hipStreamSynchronize(s));
        out.numNodes = c[3];
        out.depth = c[2];
        out.depth4 = c[4];
        out.stackNeed = c[5];
    }
done:
#undef BVH_CHK
    for (void *p : {(void *)blo, (void *)bhi, (void *)cen, (void *)idxA, (void *)idxB, (void *)order, (void *)tA, (void *)tB, (void *)bn,
                    (void *)counters, (void *)cA, (void *)cB, (void *)wt, (void *)wc, (void *)leftCnt, (void *)wflag, (void *)wmid,
                    (void *)wlocal, (void *)wblock})
        if (p) (void)hipFree(p);
    if (e != hipSuccess) {
        if (out.nodes) (void)hipFree(out.nodes);
        if (out.tri) (void)hipFree(out.tri);
        out.nodes = nullptr;
        out.tri = nullptr;
    }
    out.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return e;
}

}  // namespace drmlt

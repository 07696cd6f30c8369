// Device scene: triangle soup in BVH leaf order for traversal, the same
// triangles in scene order for re-intersection and shading, materials, area
// lights and the pinhole camera. Ray queries replace the reference's Embree
// calls (scene.cpp:80-111); re-intersection and area sampling follow
// trianglemesh.cpp:30-79 and 303-363; the camera follows camera.cpp:38-74.
#pragma once
#include "dev_math.h"

namespace drmlt {

constexpr int kBvhStack = 48;    // traversal stack entries per lane (the builder keeps every ray within it)
constexpr int kTraceLanes = 64;  // workgroup size of every kernel that traces rays

// Traversal stacks of a workgroup's lanes in LDS, interleaved by lane (entry k
// of lane t at [k * kTraceLanes + t]). Every traversal of a lane reuses its one
// stack, whatever the number of call sites the path code inlines.
__shared__ int g_trace_stack[kBvhStack * kTraceLanes];

struct DevCamera {
    float s2c[16], c2s[16], c2w[16], w2c[16];
    float dist, nearClip, farClip;
    int W, H;
};

struct DevSceneData {
    // traversal data (leaf order): per triangle three float4 p0 | e1 | e2,
    // interleaved (48 bytes), .w of p0 holding the scene-order index
    const float4 *tri;
    // 4-wide BVH: 8 float4 per node: lo.x[4], hi.x[4], lo.y[4], hi.y[4], lo.z[4], hi.z[4], child[4], count[4]
    // (count > 0 leaf triangle range, 0 internal node, < 0 empty slot)
    const float4 *nodes;
    // scene order
    const float4 *sp0, *se1, *se2;  // p0, e1 = p1 - p0, e2 = p2 - p0 (float, as the reference computes them)
    const float *sn;                // 9 floats per triangle: n0, n1, n2
    const int *tmat, *tlight;
    const float *mats;              // 16 floats per material
    int nt, nlights;
    const int *lightTris;           // first, count
    const float *lightEmit;         // rgb, weight
    const float *lightArea;         // total area per light
    const float *lightCdf;          // per light: count+1 cdf entries, at lightCdfOff[l]
    const int *lightCdfOff;
    const float *lightFunc;         // per light: triangle areas (func), at lightCdfOff[l] - l
    const float *pickCdf;           // nlights + 1
    float lightWeightSum;
    DevCamera cam;
};
// Scene of the counting kernels: the same data plus the launch's work
// counters (nodes visited, triangles tested, rays, state bytes).
struct DevSceneCounted : DevSceneData {
    unsigned long long *counters;
};

// The path code is compiled twice: as the product kernels (DevScene, no
// counting code at all) and as the counting kernels behind drmlt_count_work
// (DRMLT_COUNT_WORK defined, DevScene = DevSceneCounted).
#ifdef DRMLT_COUNT_WORK
using DevScene = DevSceneCounted;
#define DRMLT_COUNT(stmt) stmt
#else
struct DevScene : DevSceneData {};
#define DRMLT_COUNT(stmt)
#endif

// What a traversal reads: the 4-wide nodes and the leaf-order triangles (and,
// in the counting build, the work counters).
struct BvhRef {
    const float4 *nodes, *tri;
    DRMLT_COUNT(unsigned long long *counters;)
};
DEV BvhRef bvh_of(const DevScene &S) {
    BvhRef b;
    b.nodes = S.nodes; b.tri = S.tri;
    DRMLT_COUNT(b.counters = S.counters;)
    return b;
}
#ifdef DRMLT_COUNT_WORK
DEV void count_work(const BvhRef &S, unsigned nodes, unsigned tris) {
    atomicAdd(S.counters + 0, (unsigned long long)nodes);
    atomicAdd(S.counters + 1, (unsigned long long)tris);
    atomicAdd(S.counters + 2, 1ull);
}
#endif

// ---------------------------------------------------------------- traversal
// The traversal reads through global-address-space pointers: it is reached
// through a function call, where generic pointers would become flat loads
// (slower, and their completion is waited for together with LDS traffic).
typedef float vfloat4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) vfloat4 gvfloat4;
DEV float4 gload(const float4 *p, size_t i) {
    const vfloat4 v = ((gvfloat4 *)p)[i];
    return make_float4(v.x, v.y, v.z, v.w);
}

// Moller-Trumbore with the same operation order as the CPU restatement; no FMA
// contraction so that the hit search is bit-identical to a float CPU search.
// Written without branches (every quantity is computed, the tests combined at
// the end), which gives the same decision and t as testing in sequence, and
// keeps the triangle's loads from being deferred into a branch.
DEV bool tri_hit(const float4 &A, const float4 &B, const float4 &Cc, const float o[3], const float d[3], float tmin, float tmax,
                 float &t) {
#pragma clang fp contract(off)
    const float e1x = B.x, e1y = B.y, e1z = B.z, e2x = Cc.x, e2y = Cc.y, e2z = Cc.z;
    const float s1x = d[1] * e2z - d[2] * e2y, s1y = d[2] * e2x - d[0] * e2z, s1z = d[0] * e2y - d[1] * e2x;
    const float div = s1x * e1x + s1y * e1y + s1z * e1z;
    const float inv = 1.0f / div;
    const float sx = o[0] - A.x, sy = o[1] - A.y, sz = o[2] - A.z;
    const float u = (sx * s1x + sy * s1y + sz * s1z) * inv;
    const float s2x = sy * e1z - sz * e1y, s2y = sz * e1x - sx * e1z, s2z = sx * e1y - sy * e1x;
    const float v = (d[0] * s2x + d[1] * s2y + d[2] * s2z) * inv;
    t = (e2x * s2x + e2y * s2y + e2z * s2z) * inv;
    return div != 0.0f && u >= 0.0f && v >= 0.0f && u + v <= 1.0f && t > tmin && t < tmax;
}

// Ray interval against a box; a slab whose bounds come out NaN (origin on
// the plane of a zero direction component) does not narrow it.
DEV bool slab(float lox, float loy, float loz, float hix, float hiy, float hiz, const float o[3], const float inv[3], float tmin,
              float tmax, float &tnear) {
    float t0 = tmin, t1 = tmax;
    float a, b;
    bool ok;
#define DRMLT_SLAB(lo, hi, k)                             \
    a = (lo - o[k]) * inv[k];                             \
    b = (hi - o[k]) * inv[k];                             \
    ok = !(a != a) && !(b != b);                          \
    t0 = ok ? fmaxf(t0, fminf(a, b)) : t0;                \
    t1 = ok ? fminf(t1, fmaxf(a, b)) : t1;
    DRMLT_SLAB(lox, hix, 0)
    DRMLT_SLAB(loy, hiy, 1)
    DRMLT_SLAB(loz, hiz, 2)
#undef DRMLT_SLAB
    tnear = t0;
    return t0 <= t1;
}

// The same test for a ray whose inverse direction is finite in every
// component: no slab bound can be NaN then, so the guards are left out (the
// box decisions are identical).
DEV bool slab_finite(float lox, float loy, float loz, float hix, float hiy, float hiz, const float o[3], const float inv[3], float tmin,
                     float tmax, float &tnear) {
    const float ax = (lox - o[0]) * inv[0], bx = (hix - o[0]) * inv[0];
    const float ay = (loy - o[1]) * inv[1], by = (hiy - o[1]) * inv[1];
    const float az = (loz - o[2]) * inv[2], bz = (hiz - o[2]) * inv[2];
    const float t0 = fmaxf(fmaxf(fmaxf(tmin, fminf(ax, bx)), fminf(ay, by)), fminf(az, bz));
    const float t1 = fminf(fminf(fminf(tmax, fmaxf(ax, bx)), fmaxf(ay, by)), fmaxf(az, bz));
    tnear = t0;
    return t0 <= t1;
}

// next representable float above a positive finite x
DEV float next_up(float x) { return __int_as_float(__float_as_int(x) + 1); }

// One 4-wide node: the ray's interval against each child box (slab) and the
// child records.
struct Node4 {
    float4 lx, hx, ly, hy, lz, hz;
    int4 child, count;
};
DEV Node4 load_node4(const BvhRef &S, int node) {
    const size_t n = 8 * size_t(node);
    Node4 r;
    r.lx = gload(S.nodes, n); r.hx = gload(S.nodes, n + 1); r.ly = gload(S.nodes, n + 2);
    r.hy = gload(S.nodes, n + 3); r.lz = gload(S.nodes, n + 4); r.hz = gload(S.nodes, n + 5);
    const float4 c = gload(S.nodes, n + 6), k = gload(S.nodes, n + 7);
    r.child = make_int4(__float_as_int(c.x), __float_as_int(c.y), __float_as_int(c.z), __float_as_int(c.w));
    r.count = make_int4(__float_as_int(k.x), __float_as_int(k.y), __float_as_int(k.z), __float_as_int(k.w));
    return r;
}
DEV float f4at(const float4 &v, int k) { return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w; }
DEV int i4at(const int4 &v, int k) { return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w; }

// Triangles tested per batch of a leaf: their loads are issued together.
#ifndef DRMLT_TRI_BATCH
#define DRMLT_TRI_BATCH 2
#endif
constexpr int kTriBatch = DRMLT_TRI_BATCH;

// State of one ray's traversal between node visits.
struct TraceState {
    float o[3], d[3], inv[3];
    float tmin, tmax, tbest;
    int best, sp;
    bool any, done;
    bool finite;  // every inverse direction component finite: slab_finite applies
    Node4 N;  // the node the next step visits (its fetch already issued)
    DRMLT_COUNT(unsigned nvis; unsigned ntri;)
};
DEV void trace_begin(const BvhRef &S, TraceState &t, float ox, float oy, float oz, float dx, float dy, float dz, float tmin, float tmax,
                     bool any) {
    t.o[0] = ox; t.o[1] = oy; t.o[2] = oz;
    t.d[0] = dx; t.d[1] = dy; t.d[2] = dz;
    t.inv[0] = 1.0f / dx; t.inv[1] = 1.0f / dy; t.inv[2] = 1.0f / dz;
    t.tmin = tmin; t.tmax = tmax; t.tbest = tmax;
    t.best = -1; t.sp = 0;
    t.any = any; t.done = false;
    t.finite = isfinite(t.inv[0]) && isfinite(t.inv[1]) && isfinite(t.inv[2]);
    DRMLT_COUNT(t.nvis = 0; t.ntri = 0;)
    t.N = load_node4(S, 0);
}

// One traversal for both query kinds, so that lanes asking for a closest hit
// and lanes asking for visibility share the loop.
// any = false: closest hit in (tmin, tmax); ties in t go to the smaller
// scene-order index; the result is t.best (scene-order triangle index or -1)
// with the hit distance t.tbest.
// any = true: t.best is a triangle with a hit in (tmin, tmax) (the first
// found) or -1; the ray interval is not narrowed, so the answer is the same
// boolean as a search of every triangle.
// One step visits one node: the four child boxes are tested, the next node to
// visit is chosen (the nearest internal child in front of the best hit so
// far, the others pushed farthest first, else the top of the stack) and its
// fetch issued, and only then the node's leaves are tested, so the two memory
// latencies overlap. A child chosen before the leaves shortened the ray may
// turn out to lie behind the best hit: visiting it costs one fetch, its own
// box tests cull it. t.done is set when the traversal has ended.
DEV void trace_step(const BvhRef &S, TraceState &t) {
    const float *o = t.o, *d = t.d, *inv = t.inv;
    const float tmin = t.tmin, tmax = t.tmax;
    const bool any = t.any;
    int *stack = g_trace_stack + threadIdx.x;  // entry k at stack[k * kTraceLanes]
    int sp = t.sp, best = t.best;
    float tbest = t.tbest;
    const Node4 &N = t.N;
    DRMLT_COUNT(t.nvis++;)
    float tn[4];
    bool hit[4];
    if (t.finite) {
#pragma unroll
        for (int k = 0; k < 4; k++)
            hit[k] = i4at(N.count, k) >= 0 &&
                     slab_finite(f4at(N.lx, k), f4at(N.ly, k), f4at(N.lz, k), f4at(N.hx, k), f4at(N.hy, k), f4at(N.hz, k), o, inv, tmin, tbest, tn[k]);
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++)
            hit[k] = i4at(N.count, k) >= 0 &&
                     slab(f4at(N.lx, k), f4at(N.ly, k), f4at(N.lz, k), f4at(N.hx, k), f4at(N.hy, k), f4at(N.hz, k), o, inv, tmin, tbest, tn[k]);
    }
    // internal children in front of the best hit, nearest first: a
    // sorting network on (entry distance, child) in registers (an
    // insertion sort over arrays would index them dynamically and put
    // them in scratch memory); other slots get a key above every
    // distance and sort last
    const float kNone = __int_as_float(0x7f800000);
    float k0 = (hit[0] && N.count.x == 0 && tn[0] <= tbest) ? fminf(tn[0], 3.4e38f) : kNone;
    float k1 = (hit[1] && N.count.y == 0 && tn[1] <= tbest) ? fminf(tn[1], 3.4e38f) : kNone;
    float k2 = (hit[2] && N.count.z == 0 && tn[2] <= tbest) ? fminf(tn[2], 3.4e38f) : kNone;
    float k3 = (hit[3] && N.count.w == 0 && tn[3] <= tbest) ? fminf(tn[3], 3.4e38f) : kNone;
    int c0 = N.child.x, c1 = N.child.y, c2 = N.child.z, c3 = N.child.w;
    const int nh = int(k0 != kNone) + int(k1 != kNone) + int(k2 != kNone) + int(k3 != kNone);
#define DRMLT_CSWAP(ka, ca, kb, cb)                          \
{                                                        \
    const bool sw = kb < ka;                             \
    const float tk = sw ? kb : ka;                       \
    const int tc = sw ? cb : ca;                         \
    kb = sw ? ka : kb; cb = sw ? ca : cb;                \
    ka = tk; ca = tc;                                    \
}
    DRMLT_CSWAP(k0, c0, k1, c1)
    DRMLT_CSWAP(k2, c2, k3, c3)
    DRMLT_CSWAP(k0, c0, k2, c2)
    DRMLT_CSWAP(k1, c1, k3, c3)
    DRMLT_CSWAP(k1, c1, k2, c2)
#undef DRMLT_CSWAP
    int next = -1;
    if (nh > 0) {
        // farthest first onto the stack, visit the nearest
        if (nh > 3) stack[kTraceLanes * sp++] = c3;
        if (nh > 2) stack[kTraceLanes * sp++] = c2;
        if (nh > 1) stack[kTraceLanes * sp++] = c1;
        next = c0;
    } else if (sp > 0) {
        next = stack[kTraceLanes * --sp];
    }
    // the hit leaves of the node as a queue of up to four triangle
    // ranges in registers (named, not an array, so nothing is indexed
    // dynamically)
    int qf0 = 0, qe0 = 0, qf1 = 0, qe1 = 0, qf2 = 0, qe2 = 0, qf3 = 0, qe3 = 0;
#define DRMLT_ENQ(k, cnt, first)                                                                        \
if (hit[k] && (cnt) > 0) {                                                                          \
    DRMLT_COUNT(t.ntri += (cnt);)                                                                     \
    if (qe0 == 0) { qf0 = (first); qe0 = (first) + (cnt); }                                         \
    else if (qe1 == 0) { qf1 = (first); qe1 = (first) + (cnt); }                                    \
    else if (qe2 == 0) { qf2 = (first); qe2 = (first) + (cnt); }                                    \
    else { qf3 = (first); qe3 = (first) + (cnt); }                                                  \
}
    DRMLT_ENQ(0, N.count.x, N.child.x)
    DRMLT_ENQ(1, N.count.y, N.child.y)
    DRMLT_ENQ(2, N.count.z, N.child.z)
    DRMLT_ENQ(3, N.count.w, N.child.w)
#undef DRMLT_ENQ
    // the next node's fetch and the first triangle batch go out together
    Node4 NN;
    if (next >= 0) NN = load_node4(S, next);
    float4 A[kTriBatch], B[kTriBatch], Cc[kTriBatch];
    int t0 = qf0, end = qe0;
    bool have = end > 0;
#define DRMLT_LOAD_BATCH()                                                                              \
_Pragma("unroll") for (int j = 0; j < kTriBatch; j++) {                                             \
    const int t = min(t0 + j, end - 1); /* past the end: the last triangle again, not tested */     \
    A[j] = gload(S.tri, 3 * size_t(t));                                                             \
    B[j] = gload(S.tri, 3 * size_t(t) + 1);                                                         \
    Cc[j] = gload(S.tri, 3 * size_t(t) + 2);                                                        \
}
    if (have) { DRMLT_LOAD_BATCH() }
    bool found = false;
    while (have) {
#pragma unroll
        for (int j = 0; j < kTriBatch; j++) {
            float th;
            if (!found && t0 + j < end && tri_hit(A[j], B[j], Cc[j], o, d, tmin, (any || best < 0) ? tmax : next_up(tbest), th)) {
                const int id = __float_as_int(A[j].w);
                if (any) {
                    best = id;
                    found = true;
                } else if (th < tbest || (th == tbest && (best < 0 || id < best))) {
                    tbest = th;
                    best = id;
                }
            }
        }
        // the next batch: further in this leaf, else the next leaf of the queue
        t0 += kTriBatch;
        if (t0 >= end) {
            t0 = qf1; end = qe1;
            qf1 = qf2; qe1 = qe2; qf2 = qf3; qe2 = qe3; qe3 = 0;
        }
        have = !found && end > 0;
        if (have) { DRMLT_LOAD_BATCH() }
    }
#undef DRMLT_LOAD_BATCH
    t.sp = sp;
    t.best = best;
    t.tbest = tbest;
    if (found || next < 0) {
        t.done = true;
        DRMLT_COUNT(count_work(S, t.nvis, t.ntri);)
        return;
    }
    t.N = NN;
}

DEV int trace_impl(const BvhRef &S, const float o[3], const float d[3], float tmin, float tmax, bool any, float &tbest) {
    TraceState t;
    trace_begin(S, t, o[0], o[1], o[2], d[0], d[1], d[2], tmin, tmax, any);
    while (!t.done) trace_step(S, t);
    tbest = t.tbest;
    return t.best;
}

// The traversal is emitted once per kernel as a real call (not inlined at
// each of the path code's call sites), so every trace of a kernel runs the
// same code; the stage kernels additionally reach it from a single site per
// round (dev_walk.h). The counting build inlines it (it needs the counters
// and is not timed).
struct TraceHit {
    int prim;
    float t;
};
#if defined(DRMLT_COUNT_WORK) || defined(DRMLT_TRACE_INLINE)
DEV TraceHit trace_call(const BvhRef S, float ox, float oy, float oz, float dx, float dy, float dz, float tmin, float tmax, int any) {
#else
__device__ __noinline__ TraceHit trace_call(const BvhRef S, float ox, float oy, float oz, float dx, float dy, float dz, float tmin,
                                            float tmax, int any) {
#endif
    const float o[3] = {ox, oy, oz}, d[3] = {dx, dy, dz};
    TraceHit h;
    h.prim = trace_impl(S, o, d, tmin, tmax, any != 0, h.t);
    return h;
}
DEV int trace_closest(const DevScene &S, const float o[3], const float d[3], float tmin, float tmax, float &tbest) {
    const TraceHit h = trace_call(bvh_of(S), o[0], o[1], o[2], d[0], d[1], d[2], tmin, tmax, 0);
    tbest = h.t;
    return h.prim;
}
DEV bool trace_any(const DevScene &S, const float o[3], const float d[3], float tmin, float tmax) {
    return trace_call(bvh_of(S), o[0], o[1], o[2], d[0], d[1], d[2], tmin, tmax, 1).prim >= 0;
}

// ---------------------------------------------------------------- geometry in T
template <class T> struct Isect {
    V3<T> pos, gn, sn;
    int prim;
};
template <class T> struct RayT {
    V3<T> org, dir;
    float minT, maxT;
};

DEV V3<float> f3(const float4 &a) { return V3<float>(a.x, a.y, a.z); }

// trianglemesh.cpp:30-79 (re-intersect the named triangle; no u >= 0 or t check)
template <class T> DEV bool re_intersect(const DevScene &S, int prim, const RayT<T> &r, Isect<T> &is) {
    const V3<T> p0 = lift<T>(f3(S.sp0[prim])), e1 = lift<T>(f3(S.se1[prim])), e2 = lift<T>(f3(S.se2[prim]));
    is.gn = normalize(cross(e1, e2));
    V3<T> s1 = cross(r.dir, e2);
    T div = dot(s1, e1);
    if (val(div) == 0) return false;
    T inv = T(1) / div;
    V3<T> sv = r.org - p0;
    T u = dot(sv, s1) * inv;
    V3<T> s2 = cross(sv, e1);
    T v = dot(r.dir, s2) * inv;
    if (!(val(v) >= 0 && val(u + v) <= 1)) return false;
    T t = dot(e2, s2) * inv;
    T w = T(1) - u - v;
    is.pos = r.org + t * r.dir;
    const float *n = S.sn + 9 * prim;
    is.sn = normalize(w * vconst<T>(n[0], n[1], n[2]) + u * vconst<T>(n[3], n[4], n[5]) + v * vconst<T>(n[6], n[7], n[8]));
    if (val(dot(is.gn, is.sn)) < 0) is.gn = -is.gn;
    is.prim = prim;
    return true;
}

// path.cpp:95-104: closest hit on the float ray, then re-intersect in T
template <class T> DEV bool intersect(const DevScene &S, const RayT<T> &r, Isect<T> &is) {
    V3<float> o = lower(r.org), d = lower(r.dir);
    float of[3] = {o.x, o.y, o.z}, df[3] = {d.x, d.y, d.z};
    float tb;
    int prim = trace_closest(S, of, df, r.minT, r.maxT, tb);
    if (prim < 0) return false;
    return re_intersect(S, prim, r, is);
}
// scene.cpp:95-111
template <class T> DEV bool occluded(const DevScene &S, const V3<T> &org, const V3<T> &dir, T dist) {
    V3<float> o = lower(org), d = lower(dir);
    float of[3] = {o.x, o.y, o.z}, df[3] = {d.x, d.y, d.z};
    return trace_any(S, of, df, kIsectEps, (1.0f - kShadowEps) * float(val(dist)));
}

// ---------------------------------------------------------------- camera
template <class T> DEV V3<T> xform_point(const float *m, const V3<T> &p) {
    T x = T(m[0]) * p.x + T(m[1]) * p.y + T(m[2]) * p.z + T(m[3]);
    T y = T(m[4]) * p.x + T(m[5]) * p.y + T(m[6]) * p.z + T(m[7]);
    T z = T(m[8]) * p.x + T(m[9]) * p.y + T(m[10]) * p.z + T(m[11]);
    T w = T(m[12]) * p.x + T(m[13]) * p.y + T(m[14]) * p.z + T(m[15]);
    T iw = T(1) / w;
    return V3<T>(x * iw, y * iw, z * iw);
}
template <class T> DEV V3<T> xform_vec(const float *m, const V3<T> &v) {
    return V3<T>(T(m[0]) * v.x + T(m[1]) * v.y + T(m[2]) * v.z, T(m[4]) * v.x + T(m[5]) * v.y + T(m[6]) * v.z,
                 T(m[8]) * v.x + T(m[9]) * v.y + T(m[10]) * v.z);
}
// camera.cpp:38-48
template <class T> DEV void sample_primary(const DevCamera &c, const T &sx, const T &sy, RayT<T> &r) {
    V3<T> o = xform_point(c.s2c, V3<T>(sx, sy, T(0)));
    V3<T> d = normalize(o);
    float invZ = 1.0f / float(val(d.z));
    r.org = xform_point(c.c2w, V3<T>(T(0), T(0), T(0)));
    r.dir = xform_vec(c.c2w, d);
    r.minT = c.nearClip * invZ;
    r.maxT = c.farClip * invZ;
}
// camera.cpp:58-74
template <class T> DEV bool project_point(const DevCamera &c, const V3<T> &p, T &sx, T &sy) {
    V3<T> cp = xform_point(c.w2c, p);
    if (val(cp.z) < c.nearClip || val(cp.z) > c.farClip) return false;
    V3<T> rp = xform_point(c.c2s, cp);
    if (val(rp.x) < 0 || val(rp.x) > 1 || val(rp.y) < 0 || val(rp.y) > 1) return false;
    sx = rp.x;
    sy = rp.y;
    return true;
}

// ---------------------------------------------------------------- lights
// distribution.h:46-52 (SampleDiscrete: upper_bound over the normalised cdf)
DEV int sample_discrete(const float *cdf, int count, float u) {
    int lo = 0, hi = count + 1;  // first index with cdf[idx] > u
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (cdf[mid] > u) hi = mid; else lo = mid + 1;
    }
    int off = lo - 1;
    return off < 0 ? 0 : (off > count - 1 ? count - 1 : off);
}
DEV float pick_light_prob(const DevScene &S, int li) { return S.lightEmit[4 * li + 3] / S.lightWeightSum; }
DEV int pick_light(const DevScene &S, float u) { return sample_discrete(S.pickCdf, S.nlights, u); }
DEV int pick_light_tri(const DevScene &S, int li, float u) {
    return S.lightTris[2 * li] + sample_discrete(S.lightCdf + S.lightCdfOff[li], S.lightTris[2 * li + 1], u);
}

// trianglemesh.cpp:303-363
template <class T> DEV void sample_tri(const DevScene &S, int li, int prim, const T &u0, const T &u1, V3<T> &pos, V3<T> &nrm, float &pdf) {
    const T a = s_sqrt(T(1) - u0);
    const T b1 = T(1) - a, b2 = a * u1;
    pos = lift<T>(f3(S.sp0[prim])) + lift<T>(f3(S.se1[prim])) * b1 + lift<T>(f3(S.se2[prim])) * b2;
    const float *n = S.sn + 9 * prim;
    nrm = normalize(vconst<T>(n[0], n[1], n[2]) * (T(1) - b1 - b2) + vconst<T>(n[3], n[4], n[5]) * b1 + vconst<T>(n[6], n[7], n[8]) * b2);
    pdf = 1.0f / S.lightArea[li];
}
// trianglemesh.cpp:236-272 (float only: used when a generated path is converted)
DEV void get_sample_param(const DevScene &S, int prim, const V3<float> &pos, float &u0, float &u1) {
    const V3<float> e0 = pos - f3(S.sp0[prim]), e1 = f3(S.se1[prim]), e2 = f3(S.se2[prim]);
    const float d11 = dot(e1, e1), d12 = dot(e1, e2), d22 = dot(e2, e2), d01 = dot(e0, e1), d02 = dot(e0, e2);
    const float invDen = 1.0f / (d11 * d22 - d12 * d12);
    const float b1 = (d22 * d01 - d12 * d02) * invDen, b2 = (d11 * d02 - d12 * d01) * invDen;
    const float a = 1.0f - b1;
    u0 = 1.0f - a * a;
    u1 = b2 / a;
}
// arealight.cpp:28-54
template <class T> DEV bool light_sample_direct(const DevScene &S, int li, int prim, const V3<T> &pos, const T &u0, const T &u1,
                                                V3<T> &toL, T &dist, V3<T> &contrib, T &cosL, T &dPdf, T &emPdf) {
    V3<T> pl, nl;
    float shapePdf;
    sample_tri(S, li, prim, u0, u1, pl, nl, shapePdf);
    toL = pl - pos;
    T d2 = length2(toL);
    dist = s_sqrt(d2);
    toL = toL * (T(1) / dist);
    cosL = -dot(toL, nl);
    if (val(cosL) > kCosEps) {
        const float *e = S.lightEmit + 4 * li;
        T k = cosL / (d2 * T(shapePdf));
        contrib = V3<T>(k * T(e[0]), k * T(e[1]), k * T(e[2]));
        dPdf = T(shapePdf) * d2 / cosL;
        emPdf = T(shapePdf) * cosL * T(kINVPI);
        return true;
    }
    return false;
}
// arealight.cpp:56-75
template <class T> DEV void light_emission(const DevScene &S, int li, const V3<T> &dir, const V3<T> &nl, V3<T> &em, T &dPdf, T &emPdf) {
    T c = -dot(nl, dir);
    if (val(c) > 0) {
        const float *e = S.lightEmit + 4 * li;
        em = vconst<T>(e[0], e[1], e[2]);
        dPdf = T(1.0f / S.lightArea[li]);
        emPdf = c * dPdf * T(kINVPI);
    } else {
        em = vconst<T>(0, 0, 0);
        dPdf = T(0);
        emPdf = T(0);
    }
}

}  // namespace drmlt

// ORACLE (test infrastructure only): scene, camera, triangle intersection and
// area light of the reference, restated on flat triangle arrays.
#pragma once
#include "orc_core.h"

namespace orc {

// 4x4 row-major matrix helpers (transform.cpp / transform.h)
struct M4 { double m[4][4]; };
inline M4 M4Mul(const M4 &a, const M4 &b) {
    M4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double s = 0;
            for (int k = 0; k < 4; k++) s += a.m[i][k] * b.m[k][j];
            r.m[i][j] = s;
        }
    return r;
}
inline M4 M4Inverse(const M4 &in) {
    double a[4][8];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 8; j++) a[i][j] = j < 4 ? in.m[i][j] : (j - 4 == i ? 1.0 : 0.0);
    for (int c = 0; c < 4; c++) {
        int p = c;
        for (int r = c + 1; r < 4; r++) if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        for (int j = 0; j < 8; j++) std::swap(a[c][j], a[p][j]);
        double d = a[c][c];
        for (int j = 0; j < 8; j++) a[c][j] /= d;
        for (int r = 0; r < 4; r++)
            if (r != c) {
                double f = a[r][c];
                for (int j = 0; j < 8; j++) a[r][j] -= f * a[c][j];
            }
    }
    M4 o;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) o.m[i][j] = a[i][j + 4];
    return o;
}

template <class F> struct CameraT {
    F s2c[4][4], c2s[4][4], c2w[4][4], w2c[4][4];
    F dist, nearClip, farClip;
    int W, H;
};

// camera.cpp:11-28 with transform.cpp:4-8 (Scale), Translate, 79-90 (Perspective),
// 44-77 (LookAt). camToWorld is built directly from LookAt (the reference
// decomposes it into translation + quaternion, which is the same rigid map).
template <class F> inline CameraT<F> MakeCamera(const float *cam) {
    // cam: pos(3) target(3) up(3) fov W H near far
    double pos[3] = {cam[0], cam[1], cam[2]}, tgt[3] = {cam[3], cam[4], cam[5]}, up[3] = {cam[6], cam[7], cam[8]};
    double fov = cam[9];
    int W = int(cam[10]), H = int(cam[11]);
    double nearC = cam[12], farC = cam[13];
    double aspect = double(W) / double(H);
    M4 S = {{{-0.5, 0, 0, 0}, {0, -0.5 * aspect, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}}};
    M4 T = {{{1, 0, 0, -1.0}, {0, 1, 0, -1.0 / aspect}, {0, 0, 1, 0}, {0, 0, 0, 1}}};
    double recip = 1.0 / (farC - nearC);
    double cot = 1.0 / std::tan((M_PI / 180.0) * (fov / 2.0));
    M4 P = {{{cot, 0, 0, 0}, {0, cot, 0, 0}, {0, 0, farC * recip, -nearC * farC * recip}, {0, 0, 1, 0}}};
    M4 c2s = M4Mul(M4Mul(S, T), P);
    M4 s2c = M4Inverse(c2s);
    double d[3] = {tgt[0] - pos[0], tgt[1] - pos[1], tgt[2] - pos[2]};
    double dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (auto &x : d) x /= dl;
    double ul = std::sqrt(up[0] * up[0] + up[1] * up[1] + up[2] * up[2]);
    double un[3] = {up[0] / ul, up[1] / ul, up[2] / ul};
    double left[3] = {un[1] * d[2] - un[2] * d[1], un[2] * d[0] - un[0] * d[2], un[0] * d[1] - un[1] * d[0]};
    double ll = std::sqrt(left[0] * left[0] + left[1] * left[1] + left[2] * left[2]);
    for (auto &x : left) x /= ll;
    double nu[3] = {d[1] * left[2] - d[2] * left[1], d[2] * left[0] - d[0] * left[2], d[0] * left[1] - d[1] * left[0]};
    M4 c2w = {{{left[0], nu[0], d[0], pos[0]}, {left[1], nu[1], d[1], pos[1]}, {left[2], nu[2], d[2], pos[2]}, {0, 0, 0, 1}}};
    M4 w2c = M4Inverse(c2w);
    CameraT<F> c;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            c.s2c[i][j] = F(s2c.m[i][j]);
            c.c2s[i][j] = F(c2s.m[i][j]);
            c.c2w[i][j] = F(c2w.m[i][j]);
            c.w2c[i][j] = F(w2c.m[i][j]);
        }
    c.dist = F(W / (2.0 * std::tan((fov / 2.0) * (M_PI / 180.0))));
    c.nearClip = F(nearC);
    c.farClip = F(farC);
    c.W = W;
    c.H = H;
    return c;
}
// transform.h:21-30
template <class F> inline V3<F> XformPoint(const F m[4][4], const V3<F> &p) {
    F x = m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z + m[0][3];
    F y = m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z + m[1][3];
    F z = m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z + m[2][3];
    F w = m[3][0] * p.x + m[3][1] * p.y + m[3][2] * p.z + m[3][3];
    F iw = F(1) / w;
    return V3<F>(x * iw, y * iw, z * iw);
}
template <class F> inline V3<F> XformVector(const F m[4][4], const V3<F> &v) {
    return V3<F>(m[0][0] * v.x + m[0][1] * v.y + m[0][2] * v.z, m[1][0] * v.x + m[1][1] * v.y + m[1][2] * v.z,
                 m[2][0] * v.x + m[2][1] * v.y + m[2][2] * v.z);
}

template <class F> struct Ray { V3<F> org, dir; F minT, maxT; };

// camera.cpp:38-48
template <class F> inline void SamplePrimary(const CameraT<F> &c, F sx, F sy, Ray<F> &r) {
    V3<F> o = XformPoint(c.s2c, V3<F>(sx, sy, F(0)));
    V3<F> d = Normalize(o);
    F invZ = F(1) / d.z;
    r.org = XformPoint(c.c2w, V3<F>());
    r.dir = XformVector(c.c2w, d);
    r.minT = c.nearClip * invZ;
    r.maxT = c.farClip * invZ;
}
// camera.cpp:58-74
template <class F> inline bool ProjectPoint(const CameraT<F> &c, const V3<F> &p, F &sx, F &sy) {
    V3<F> cp = XformPoint(c.w2c, p);
    if (cp.z < c.nearClip || cp.z > c.farClip) return false;
    V3<F> rp = XformPoint(c.c2s, cp);
    if (rp.x < F(0) || rp.x > F(1) || rp.y < F(0) || rp.y > F(1)) return false;
    sx = rp.x;
    sy = rp.y;
    return true;
}

// distribution.h:8-57 (PiecewiseConstant1D)
template <class F> struct Dist1D {
    std::vector<F> func, cdf;
    F funcInt = 0;
    int count = 0;
    void build(const F *f, int n) {
        count = n;
        func.assign(f, f + n);
        cdf.assign(n + 1, F(0));
        for (int i = 1; i < n + 1; i++) cdf[i] = cdf[i - 1] + func[i - 1] / F(n);
        funcInt = cdf[n];
        if (funcInt == F(0)) {
            for (int i = 1; i < n + 1; i++) cdf[i] = F(i) / F(n);
        } else {
            for (int i = 1; i < n + 1; i++) cdf[i] /= funcInt;
        }
    }
    int SampleDiscrete(F u) const {
        auto it = std::upper_bound(cdf.begin(), cdf.end(), u);
        int off = int(it - cdf.begin() - 1);
        return std::min(std::max(off, 0), count - 1);
    }
    F Pmf(int off) const { return func[off] / (funcInt * F(count)); }
};

template <class F> struct LightT {
    int firstTri, numTris;
    V3<F> emission;
    F weight;
    F totalArea;
    Dist1D<F> areaDist;
};

struct BvhNode { float lo[3], hi[3]; int left, right, first, count; };

template <class F> struct SceneT {
    int nt = 0;
    std::vector<V3<F>> p0, e1, e2, n0, n1, n2;  // per triangle
    std::vector<int> triMat, triLight;
    std::vector<Material<F>> mats;
    std::vector<LightT<F>> lights;
    Dist1D<F> lightDist;
    F lightWeightSum = 0;
    CameraT<F> cam;
    // float copies for the hit search (the reference's Embree works in float,
    // scene.cpp:61-78)
    std::vector<float> fp0, fe1, fe2;
    std::vector<BvhNode> bvh;
    std::vector<int> bvhTris;
    bool useBvh = true;
};

// Moller-Trumbore on float data; the same arithmetic as trianglemesh.cpp:30-56
// plus the u >= 0 and (tmin, tmax) checks that a hit search needs.
inline bool TriHitF(const float *p0, const float *e1, const float *e2, const float o[3], const float d[3], float tmin,
                    float tmax, float &t, float &u, float &v) {
    float s1[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    float div = s1[0] * e1[0] + s1[1] * e1[1] + s1[2] * e1[2];
    if (div == 0.0f) return false;
    float inv = 1.0f / div;
    float s[3] = {o[0] - p0[0], o[1] - p0[1], o[2] - p0[2]};
    u = (s[0] * s1[0] + s[1] * s1[1] + s[2] * s1[2]) * inv;
    if (!(u >= 0.0f)) return false;
    float s2[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
    v = (d[0] * s2[0] + d[1] * s2[1] + d[2] * s2[2]) * inv;
    if (!(v >= 0.0f) || !(u + v <= 1.0f)) return false;
    t = (e2[0] * s2[0] + e2[1] * s2[1] + e2[2] * s2[2]) * inv;
    return t > tmin && t < tmax;
}

// Simple median-split BVH for the CPU baseline (not the product's builder).
template <class F> inline void BuildOracleBvh(SceneT<F> &s) {
    int n = s.nt;
    std::vector<int> idx(n);
    for (int i = 0; i < n; i++) idx[i] = i;
    std::vector<float> cen(3 * n), lo(3 * n), hi(3 * n);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            float a = s.fp0[3 * i + k], b = a + s.fe1[3 * i + k], c = a + s.fe2[3 * i + k];
            lo[3 * i + k] = std::min(a, std::min(b, c));
            hi[3 * i + k] = std::max(a, std::max(b, c));
            cen[3 * i + k] = 0.5f * (lo[3 * i + k] + hi[3 * i + k]);
        }
    s.bvh.clear();
    s.bvh.reserve(2 * n);
    struct Item { int node, first, count; };
    std::vector<Item> stack;
    s.bvh.push_back(BvhNode());
    stack.push_back({0, 0, n});
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        BvhNode nd;
        for (int k = 0; k < 3; k++) { nd.lo[k] = 1e30f; nd.hi[k] = -1e30f; }
        for (int i = it.first; i < it.first + it.count; i++)
            for (int k = 0; k < 3; k++) {
                nd.lo[k] = std::min(nd.lo[k], lo[3 * idx[i] + k]);
                nd.hi[k] = std::max(nd.hi[k], hi[3 * idx[i] + k]);
            }
        // pad boxes by a relative epsilon so float rounding never culls a hit
        for (int k = 0; k < 3; k++) {
            float pad = 1e-5f * (std::fabs(nd.lo[k]) + std::fabs(nd.hi[k])) + 1e-7f;
            nd.lo[k] -= pad;
            nd.hi[k] += pad;
        }
        if (it.count <= 4) {
            nd.left = nd.right = -1;
            nd.first = it.first;
            nd.count = it.count;
            s.bvh[it.node] = nd;
            continue;
        }
        int axis = 0;
        float ext = -1;
        for (int k = 0; k < 3; k++)
            if (nd.hi[k] - nd.lo[k] > ext) { ext = nd.hi[k] - nd.lo[k]; axis = k; }
        int mid = it.first + it.count / 2;
        std::nth_element(idx.begin() + it.first, idx.begin() + mid, idx.begin() + it.first + it.count,
                         [&](int a, int b) { return cen[3 * a + axis] < cen[3 * b + axis]; });
        nd.left = int(s.bvh.size());
        nd.right = nd.left + 1;
        nd.first = nd.count = 0;
        s.bvh[it.node] = nd;
        s.bvh.push_back(BvhNode());
        s.bvh.push_back(BvhNode());
        stack.push_back({nd.left, it.first, mid - it.first});
        stack.push_back({nd.right, mid, it.first + it.count - mid});
    }
    s.bvhTris = idx;
}

inline bool BoxHit(const BvhNode &n, const float o[3], const float inv[3], float tmin, float tmax) {
    float t0 = tmin, t1 = tmax;
    for (int k = 0; k < 3; k++) {
        float a = (n.lo[k] - o[k]) * inv[k], b = (n.hi[k] - o[k]) * inv[k];
        if (a > b) std::swap(a, b);
        if (!(a <= t1 && b >= t0)) {
            if (std::isnan(a) || std::isnan(b)) continue;  // origin on a slab plane with zero direction
            return false;
        }
        t0 = std::max(t0, a);
        t1 = std::min(t1, b);
    }
    return true;
}

template <class F> inline int TraceClosest(const SceneT<F> &s, const float o[3], const float d[3], float tmin, float tmax, float &tb) {
    if (!s.useBvh || s.bvh.empty()) {
        int best = -1;
        tb = tmax;
        for (int i = 0; i < s.nt; i++) {
            float t, u, v;
            if (TriHitF(&s.fp0[3 * i], &s.fe1[3 * i], &s.fe2[3 * i], o, d, tmin, tmax, t, u, v))
                if (t < tb || (t == tb && i < best)) { tb = t; best = i; }
        }
        return best;
    }
    float inv[3] = {1.0f / d[0], 1.0f / d[1], 1.0f / d[2]};
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    int best = -1;
    tb = tmax;
    while (sp) {
        const BvhNode &n = s.bvh[stack[--sp]];
        if (!BoxHit(n, o, inv, tmin, tb)) continue;
        if (n.left < 0) {
            for (int j = n.first; j < n.first + n.count; j++) {
                int i = s.bvhTris[j];
                float t, u, v;
                if (TriHitF(&s.fp0[3 * i], &s.fe1[3 * i], &s.fe2[3 * i], o, d, tmin, best < 0 ? tmax : std::nextafter(tb, 3e38f), t, u, v))
                    if (t < tb || (t == tb && (best < 0 || i < best))) { tb = t; best = i; }
            }
        } else {
            stack[sp++] = n.left;
            stack[sp++] = n.right;
        }
    }
    return best;
}

template <class F> inline bool TraceAny(const SceneT<F> &s, const float o[3], const float d[3], float tmin, float tmax) {
    if (!s.useBvh || s.bvh.empty()) {
        for (int i = 0; i < s.nt; i++) {
            float t, u, v;
            if (TriHitF(&s.fp0[3 * i], &s.fe1[3 * i], &s.fe2[3 * i], o, d, tmin, tmax, t, u, v)) return true;
        }
        return false;
    }
    float inv[3] = {1.0f / d[0], 1.0f / d[1], 1.0f / d[2]};
    int stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const BvhNode &n = s.bvh[stack[--sp]];
        if (!BoxHit(n, o, inv, tmin, tmax)) continue;
        if (n.left < 0) {
            for (int j = n.first; j < n.first + n.count; j++) {
                int i = s.bvhTris[j];
                float t, u, v;
                if (TriHitF(&s.fp0[3 * i], &s.fe1[3 * i], &s.fe2[3 * i], o, d, tmin, tmax, t, u, v)) return true;
            }
        } else {
            stack[sp++] = n.left;
            stack[sp++] = n.right;
        }
    }
    return false;
}

template <class F> struct Isect {
    V3<F> pos, gn, sn;
    int prim = -1;
};

// trianglemesh.cpp:58-79: re-intersect the named triangle to get position and
// normals (no u >= 0 or t-range check, as in the reference).
template <class F> inline bool ReIntersect(const SceneT<F> &s, int prim, const Ray<F> &r, Isect<F> &is) {
    const V3<F> &p0 = s.p0[prim], &e1 = s.e1[prim], &e2 = s.e2[prim];
    is.gn = Normalize(Cross(e1, e2));
    V3<F> s1 = Cross(r.dir, e2);
    F div = Dot(s1, e1);
    if (div == F(0)) return false;
    F inv = F(1) / div;
    V3<F> sv = r.org - p0;
    F u = Dot(sv, s1) * inv;
    V3<F> s2 = Cross(sv, e1);
    F v = Dot(r.dir, s2) * inv;
    if (!(v >= F(0) && u + v <= F(1))) return false;
    F t = Dot(e2, s2) * inv;
    F w = F(1) - u - v;
    is.pos = r.org + t * r.dir;
    is.sn = Normalize(w * s.n0[prim] + u * s.n1[prim] + v * s.n2[prim]);
    if (Dot(is.gn, is.sn) < F(0)) is.gn = -is.gn;
    is.prim = prim;
    return true;
}

// path.cpp:95-104 + scene.cpp:80-93
template <class F> inline bool Intersect(const SceneT<F> &s, const Ray<F> &r, Isect<F> &is) {
    float o[3] = {float(r.org.x), float(r.org.y), float(r.org.z)};
    float d[3] = {float(r.dir.x), float(r.dir.y), float(r.dir.z)};
    float tb;
    float maxT = std::isinf(double(r.maxT)) ? std::numeric_limits<float>::infinity() : float(r.maxT);
    int prim = TraceClosest(s, o, d, float(r.minT), maxT, tb);
    if (prim < 0) return false;
    return ReIntersect(s, prim, r, is);
}
// scene.cpp:95-111
template <class F> inline bool Occluded(const SceneT<F> &s, const V3<F> &org, const V3<F> &dir, F dist) {
    float o[3] = {float(org.x), float(org.y), float(org.z)};
    float d[3] = {float(dir.x), float(dir.y), float(dir.z)};
    float minT = float(kIsectEps), maxT = float((F(1) - F(kShadowEps)) * dist);
    return TraceAny(s, o, d, minT, maxT);
}

// trianglemesh.cpp:303-363 (area sampling of one triangle, pdf = 1/totalArea)
template <class F> inline void SampleTri(const SceneT<F> &s, const LightT<F> &L, int prim, F u0, F u1, V3<F> &pos, V3<F> &nrm, F &pdf) {
    const F a = std::sqrt(F(1) - u0);
    const F b1 = F(1) - a, b2 = a * u1;
    pos = s.p0[prim] + (s.e1[prim] * b1) + (s.e2[prim] * b2);
    nrm = Normalize(s.n0[prim] * (F(1) - b1 - b2) + s.n1[prim] * b1 + s.n2[prim] * b2);
    pdf = F(1) / L.totalArea;
}
// trianglemesh.cpp:236-272
template <class F> inline void GetSampleParam(const SceneT<F> &s, int prim, const V3<F> &pos, F &u0, F &u1) {
    const V3<F> e0 = pos - s.p0[prim], &e1 = s.e1[prim], &e2 = s.e2[prim];
    const F d11 = Dot(e1, e1), d12 = Dot(e1, e2), d22 = Dot(e2, e2), d01 = Dot(e0, e1), d02 = Dot(e0, e2);
    const F invDen = F(1) / (d11 * d22 - d12 * d12);
    const F b1 = (d22 * d01 - d12 * d02) * invDen, b2 = (d11 * d02 - d12 * d01) * invDen;
    const F a = F(1) - b1;
    u0 = F(1) - sq(a);
    u1 = b2 / a;
}

// arealight.cpp:28-54
template <class F> inline bool LightSampleDirect(const SceneT<F> &s, const LightT<F> &L, int prim, const V3<F> &pos, F u0, F u1,
                                                 V3<F> &dirToLight, F &dist, V3<F> &contrib, F &cosAtLight, F &directPdf, F &emissionPdf) {
    V3<F> pl, nl;
    F shapePdf;
    SampleTri(s, L, prim, u0, u1, pl, nl, shapePdf);
    dirToLight = pl - pos;
    F d2 = LengthSquared(dirToLight);
    dist = std::sqrt(d2);
    dirToLight = dirToLight / dist;
    cosAtLight = -Dot(dirToLight, nl);
    if (cosAtLight > C<F>::CosEps) {
        contrib = (cosAtLight / (d2 * shapePdf)) * L.emission;
        directPdf = shapePdf * d2 / cosAtLight;
        emissionPdf = shapePdf * cosAtLight * C<F>::INVPI;
        return true;
    }
    return false;
}
// arealight.cpp:56-75
template <class F> inline void LightEmission(const LightT<F> &L, const V3<F> &dir, const V3<F> &nl, V3<F> &em, F &directPdf, F &emissionPdf) {
    F c = -Dot(nl, dir);
    if (c > F(0)) {
        em = L.emission;
        directPdf = F(1) / L.totalArea;
        emissionPdf = c * directPdf * C<F>::INVPI;
    } else {
        em = V3<F>();
        directPdf = 0;
        emissionPdf = 0;
    }
}
// arealight.cpp:77-104
template <class F> inline void LightEmit(const SceneT<F> &s, const LightT<F> &L, int prim, F p0, F p1, F d0, F d1, Ray<F> &r,
                                         V3<F> &em, F &cosAtLight, F &emissionPdf, F &directPdf) {
    V3<F> nrm;
    F shapePdf;
    SampleTri(s, L, prim, p0, p1, r.org, nrm, shapePdf);
    V3<F> d = SampleCosHemisphere(d0, d1);
    V3<F> b0, b1;
    CoordinateSystem(nrm, b0, b1);
    r.dir = d.x * b0 + d.y * b1 + d.z * nrm;
    em = L.emission * (F(M_PI) / shapePdf);
    cosAtLight = d.z;
    emissionPdf = d.z * C<F>::INVPI * shapePdf;
    directPdf = shapePdf;
}

template <class F> inline F PickLightProb(const SceneT<F> &s, int li) { return s.lights[li].weight / s.lightWeightSum; }

// Scene from flat arrays (the scene format the product also reads).
template <class F> inline void LoadScene(SceneT<F> &s, int nt, const float *triP, const float *triN, const int *triMat, const int *triLight,
                                         int nm, const float *mats, int nl, const int *lightTris, const float *lightEmit, const float *cam) {
    s.nt = nt;
    s.p0.resize(nt); s.e1.resize(nt); s.e2.resize(nt); s.n0.resize(nt); s.n1.resize(nt); s.n2.resize(nt);
    s.fp0.resize(3 * nt); s.fe1.resize(3 * nt); s.fe2.resize(3 * nt);
    s.triMat.assign(triMat, triMat + nt);
    s.triLight.assign(triLight, triLight + nt);
    for (int i = 0; i < nt; i++) {
        const float *p = triP + 9 * i, *n = triN + 9 * i;
        // e1 = p1 - p0, e2 = p2 - p0 evaluated in float, as trianglemesh.cpp:204-205 does
        float e1f[3] = {p[3] - p[0], p[4] - p[1], p[5] - p[2]};
        float e2f[3] = {p[6] - p[0], p[7] - p[1], p[8] - p[2]};
        for (int k = 0; k < 3; k++) { s.fp0[3 * i + k] = p[k]; s.fe1[3 * i + k] = e1f[k]; s.fe2[3 * i + k] = e2f[k]; }
        s.p0[i] = V3<F>(p[0], p[1], p[2]);
        s.e1[i] = V3<F>(F(p[3]) - F(p[0]), F(p[4]) - F(p[1]), F(p[5]) - F(p[2]));
        s.e2[i] = V3<F>(F(p[6]) - F(p[0]), F(p[7]) - F(p[1]), F(p[8]) - F(p[2]));
        s.n0[i] = V3<F>(n[0], n[1], n[2]);
        s.n1[i] = V3<F>(n[3], n[4], n[5]);
        s.n2[i] = V3<F>(n[6], n[7], n[8]);
    }
    s.mats.resize(nm);
    for (int i = 0; i < nm; i++) {
        const float *m = mats + 16 * i;
        Material<F> &M = s.mats[i];
        M.type = int(m[0]);
        M.a = V3<F>(m[1], m[2], m[3]);
        M.b = V3<F>(m[4], m[5], m[6]);
        M.expo = m[7];
        M.alpha = m[8];
        M.eta = m[9];
        M.invEta = m[10];
        M.ksWeight = m[11];
        M.twoSided = m[12] != 0.0f;
    }
    s.lights.resize(nl);
    std::vector<F> w(nl);
    s.lightWeightSum = 0;
    for (int li = 0; li < nl; li++) {
        LightT<F> &L = s.lights[li];
        L.firstTri = lightTris[2 * li];
        L.numTris = lightTris[2 * li + 1];
        L.emission = V3<F>(lightEmit[4 * li], lightEmit[4 * li + 1], lightEmit[4 * li + 2]);
        L.weight = lightEmit[4 * li + 3];
        std::vector<F> area(L.numTris);
        L.totalArea = 0;
        for (int k = 0; k < L.numTris; k++) {
            int t = L.firstTri + k;
            area[k] = F(0.5) * Length(Cross(s.e1[t], s.e2[t]));
            L.totalArea += area[k];
        }
        L.areaDist.build(area.data(), L.numTris);
        w[li] = L.weight;
        s.lightWeightSum += L.weight;
    }
    s.lightDist.build(w.data(), nl);
    s.cam = MakeCamera<F>(cam);
    BuildOracleBvh(s);
}

}  // namespace orc

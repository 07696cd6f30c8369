// ORACLE (test infrastructure only): the reference's bidirectional path space
// (path.cpp) restated: subpath generation, fixed-topology re-evaluation
// (PerturbPathBidir), MIS bookkeeping and the primary-sample layout.
//
// Lens bookkeeping (lensContrib, lensScore, lensVertexPos) is tracked as the
// reference tracks it; the h2mc integrator's lens mutation reads it.
// Simplification relative to the reference: scenes are static, so `time`
// carries no information.
#pragma once
#include "orc_scene.h"

namespace orc {

constexpr int VMAX = 16;  // surface vertices per subpath
constexpr int DMAX = 40;  // primary-sample dimensions

// Uniform / normal sources. Generation draws uniforms in a fixed documented
// order so the product can be fed the same stream in parity tests.
template <class F> struct Src {
    virtual F uni() = 0;
    virtual ~Src() {}
};
template <class F> struct ArraySrc : Src<F> {
    const float *u; int n; int i = 0;
    ArraySrc(const float *u_, int n_) : u(u_), n(n_) {}
    F uni() override { return i < n ? F(u[i++]) : F(0.5); }
};

// path.h:30-50
template <class F> struct SurfV {
    int prim = -1;
    F rnd[2] = {0, 0};
    F disc = 0;
    F useAbs = 0;
    int dlLight = 0, dlPrim = -1;
    F dlRnd[2] = {0, 0};
    F rr = 1;
};
template <class F> struct PathS {
    int c = 0, l = 0;  // camera / light subpath vertex counts (SubpathContrib camDepth/lightDepth)
    F screen[2] = {0, 0};
    F lPos[2] = {0, 0}, lDir[2] = {0, 0};
    int lLight = 0, lPrim = -1;
    V3<F> lensPos;  // lensVertexPos: the second camera vertex of the last evaluation (path.h:50-107)
    std::vector<SurfV<F>> cam, lgt;
};
// path.h:12-24
template <class F> struct Contrib {
    int c = 0, l = 0;
    F screen[2] = {0, 0};
    V3<F> contrib;
    F lsScore = 0, ssScore = 0, lensScore = 0, mis = 0;
};

template <class F> struct Opts {
    F roughThr = F(0.03);  // dptoptions.h:19
    bool lightCoord = true;  // dptoptions.h:28
    bool chadMisQuirk = false;  // path.cpp:2598-2604 (AD twin only)
};

// path.cpp:433-444
template <class F> struct BState {
    Isect<F> isect;
    V3<F> wi;
    F prev = 0, thisw = 0;  // accMISWPrev, accMISWThis
    V3<F> thr;
    V3<F> lensC;  // lensContrib
    F ssJ = 0, lcJ = 0, lastPdf = 0;
};

template <class F> inline F MIS(F p) { return p * p; }  // path.cpp:33-37

// path.cpp:39-58
template <class F> inline F SNC(bool adjoint, const V3<F> &wi, const Isect<F> &is, const V3<F> &wo) {
    const F cwi = Dot(is.sn, wi), cwo = Dot(is.sn, wo);
    const F wg = Dot(is.gn, wi), og = Dot(is.gn, wo);
    if (wg * cwi <= F(0) || og * cwo <= F(0)) return F(0);
    if (adjoint) return std::fabs((og * cwi) / (wg * cwo));
    return F(1);
}

template <class F> inline const Material<F> &MatOf(const SceneT<F> &s, int prim) { return s.mats[s.triMat[prim]]; }

// path.cpp:455-474
template <class F> inline void EmitFromCamera(const SceneT<F> &s, const F screen[2], Ray<F> &r, BState<F> &st) {
    Ray<F> cr;
    SamplePrimary(s.cam, F(0.5), F(0.5), cr);
    SamplePrimary(s.cam, screen[0], screen[1], r);
    const F cosAtCam = Dot(cr.dir, r.dir);
    const F ipcd = s.cam.dist / cosAtCam;
    const F pdfW = sq(ipcd) / cosAtCam;
    const F npix = F(s.cam.W * s.cam.H);
    st.thr = V3<F>(1, 1, 1);
    st.prev = MIS(npix / pdfW);
    st.thisw = 0;
    st.ssJ = 1;
    st.lensC = V3<F>(1, 1, 1);
}
// path.cpp:489-505
template <class F> inline void EmitFromLight(const SceneT<F> &s, const Opts<F> &o, F pick, const PathS<F> &p, Ray<F> &r, BState<F> &st) {
    const LightT<F> &L = s.lights[p.lLight];
    F cosL, emPdf, dPdf;
    LightEmit(s, L, p.lPrim, p.lPos[0], p.lPos[1], p.lDir[0], p.lDir[1], r, st.thr, cosL, emPdf, dPdf);
    emPdf *= pick;
    dPdf *= pick;
    st.thr = st.thr * (F(1) / pick);
    st.prev = MIS(dPdf / emPdf);
    st.thisw = o.chadMisQuirk ? F(0) : MIS(cosL / emPdf);
    st.ssJ = 1;
    st.lensC = V3<F>(0, 0, 0);
}
// path.cpp:507-517
template <class F> inline void ConvertMIS(const Ray<F> &r, BState<F> &st) {
    st.prev *= MIS(DistanceSquared(r.org, st.isect.pos));
    F invC = F(1) / MIS(std::fabs(Dot(r.dir, st.isect.sn)));
    st.prev *= invC;
    st.thisw *= invC;
}

// path.cpp:519-613. Returns false when the reference clears the contribution
// list (a degenerate zero-length edge). prevLensC: the lensContrib of the
// light-subpath state before this vertex.
template <class F> inline bool ConnectToCamera(const SceneT<F> &s, const Opts<F> &o, bool occl, int lgtDepth, const BState<F> &st,
                                               const SurfV<F> &v, const V3<F> &prevLensC, const V3<F> &prevPos, std::vector<Contrib<F>> &out) {
    Ray<F> cr;
    SamplePrimary(s.cam, F(0.5), F(0.5), cr);
    V3<F> toCam = cr.org - st.isect.pos;
    if (-Dot(cr.dir, toCam) <= F(0)) return true;
    F sx, sy;
    if (!ProjectPoint(s.cam, st.isect.pos, sx, sy)) return true;
    const F d2 = LengthSquared(toCam), d = std::sqrt(d2);
    toCam = toCam * (F(1) / d);
    if (occl && Occluded(s, st.isect.pos, toCam, d)) return true;
    const Material<F> &m = MatOf(s, v.prim);
    BsdfOut<F> b;
    BsdfEval(m, true, st.wi, st.isect.sn, toCam, b);
    if (b.contrib.isZero()) return true;
    F fac = SNC(true, st.wi, st.isect, toCam);
    if (fac <= F(0)) return true;
    b.contrib = b.contrib * fac;
    V3<F> lensC(0, 0, 0);
    if (Roughness(m, v.disc) > o.roughThr && lgtDepth >= 1) {
        BsdfOut<F> e;
        BsdfEval(m, false, toCam, st.isect.sn, st.wi, e);
        const F dd = DistanceSquared(st.isect.pos, prevPos);
        if (dd <= F(0)) { out.clear(); return false; }
        lensC = e.contrib.mul(prevLensC) * (F(1) / dd);
    }
    const F cosAtCam = -Dot(cr.dir, toCam);
    const F ipcd = s.cam.dist / cosAtCam;
    const F i2sa = sq(ipcd) / cosAtCam;
    const F i2surf = i2sa * std::fabs(b.cosWo) / d2;
    const F npix = F(s.cam.W * s.cam.H);
    const F wLight = MIS(i2surf / npix) * (st.prev + st.thisw * MIS(b.revPdf));
    const F mis = F(1) / (wLight + F(1));
    const F surf2img = b.cosWo / i2surf;
    V3<F> c = (mis * b.contrib / (npix * surf2img)).mul(st.thr);
    const F score = Luminance(c);
    if (score > F(0)) {
        Contrib<F> k;
        k.c = 1; k.l = 2 + lgtDepth; k.screen[0] = sx; k.screen[1] = sy;
        k.contrib = c; k.lsScore = score; k.ssScore = score * st.ssJ; k.lensScore = Luminance(lensC); k.mis = mis;
        out.push_back(k);
    }
    return true;
}

// path.cpp:615-724 (BSDFSampling<adjoint, perturb>) at subpath vertex `depth`.
// `next` may alias `cur`.
template <class F> inline bool BSDFSampling(const SceneT<F> &s, const Opts<F> &o, bool adjoint, bool perturb, int depth, const BState<F> &cur,
                                            SurfV<F> &v, BState<F> &next, V3<F> &dir, V3<F> &bc) {
    const Material<F> &m = MatOf(s, v.prim);
    BsdfOut<F> b;
    v.useAbs = Roughness(m, v.disc) > o.roughThr ? F(1) : F(0);
    if (!perturb || v.useAbs == F(0)) {
        if (!BsdfSample(m, adjoint, cur.wi, cur.isect.sn, v.rnd[0], v.rnd[1], v.disc, dir, b)) return false;
        bc = b.contrib;
        if (v.useAbs == F(1)) {
            F jac;
            ToSphericalCoord(dir, v.rnd[0], v.rnd[1], jac);
            next.lcJ = F(1) / jac;
            jac *= b.pdf;
            next.ssJ = cur.ssJ * jac;
        } else {
            next.lcJ = b.pdf;
        }
    } else {
        F jac;
        dir = SampleSphere(v.rnd[0], v.rnd[1], jac);
        BsdfEval(m, adjoint, cur.wi, cur.isect.sn, dir, b);
        if (b.contrib.isZero() || b.pdf <= F(0)) return false;
        bc = b.contrib * (F(1) / b.pdf);
        next.lcJ = F(1) / jac;
        jac *= b.pdf;
        next.ssJ = cur.ssJ * jac;
    }
    F fac = SNC(adjoint, cur.wi, cur.isect, dir);
    if (fac <= F(0)) return false;
    // lensContrib (path.cpp:664-711); bc is still contrib / pdf here
    const bool grazing = std::fabs(b.cosWo) < F(1e-2) || std::fabs(Dot(dir, cur.isect.gn)) < F(1e-2);
    if (v.useAbs == F(1)) {
        if (adjoint) {
            next.lensC = grazing ? V3<F>(0, 0, 0) : bc * b.pdf;
        } else if (depth == 0) {
            next.lensC = grazing ? V3<F>(0, 0, 0) : cur.lensC.mul(bc * b.pdf);
        } else if (depth == 1) {
            BsdfOut<F> imp;
            BsdfEval(m, true, dir, cur.isect.sn, cur.wi, imp);
            if (std::fabs(imp.cosWo) < F(1e-2) || std::fabs(Dot(cur.wi, cur.isect.gn)) < F(1e-2)) {
                next.lensC = V3<F>(0, 0, 0);
            } else {
                next.lensC = g_plant == 21 ? cur.lensC : cur.lensC.mul(imp.contrib * SNC(true, dir, cur.isect, cur.wi));
            }
        } else {
            next.lensC = cur.lensC;
        }
    } else {
        next.lensC = (adjoint || depth <= 1) ? V3<F>(0, 0, 0) : cur.lensC;
    }
    bc = bc * fac;
    if (adjoint) next.lensC = next.lensC * fac;
    next.lastPdf = b.pdf;
    const F newThis = MIS(b.cosWo / b.pdf) * (cur.thisw * MIS(b.revPdf) + cur.prev);
    next.thisw = newThis;
    next.prev = MIS(F(1) / b.pdf);
    next.thr = cur.thr.mul(bc);
    return true;
}

// path.cpp:726-810 (bidirectional DirectLighting)
template <class F> inline void DirectLighting(const SceneT<F> &s, const Opts<F> &o, bool occl, int camDepth, const BState<F> &st,
                                              const F screen[2], F pick, const SurfV<F> &v, std::vector<Contrib<F>> &out) {
    const Material<F> &m = MatOf(s, v.prim);
    const LightT<F> &L = s.lights[v.dlLight];
    V3<F> toL, lc;
    F dist, cosL, dPdf, emPdf;
    if (!LightSampleDirect(s, L, v.dlPrim, st.isect.pos, v.dlRnd[0], v.dlRnd[1], toL, dist, lc, cosL, dPdf, emPdf)) return;
    if (occl && Occluded(s, st.isect.pos, toL, dist)) return;
    BsdfOut<F> b;
    BsdfEval(m, false, st.wi, st.isect.sn, toL, b);
    if (b.contrib.isZero()) return;
    F fac = SNC(false, st.wi, st.isect, toL);
    if (fac <= F(0)) return;
    b.contrib = b.contrib * fac;
    V3<F> lensC = st.lensC;
    if (camDepth == 1) {
        if (Roughness(m, v.disc) > o.roughThr) {
            BsdfOut<F> imp;
            BsdfEval(m, true, toL, st.isect.sn, st.wi, imp);
            lensC = lensC.mul(imp.contrib * SNC(true, toL, st.isect, st.wi));
        } else {
            lensC = V3<F>(0, 0, 0);
        }
    }
    V3<F> c = st.thr.mul(b.contrib).mul(lc) * (F(1) / pick);
    const F wL = MIS(b.pdf / (pick * dPdf));
    const F wC = MIS(emPdf * b.cosWo / (dPdf * cosL)) * (st.prev + st.thisw * MIS(b.revPdf));
    const F mis = F(1) / (wL + F(1) + wC);
    c = c * mis;
    const F score = Luminance(c);
    if (score > F(0)) {
        Contrib<F> k;
        k.c = 2 + camDepth; k.l = 1; k.screen[0] = screen[0]; k.screen[1] = screen[1];
        k.contrib = c; k.lsScore = score; k.ssScore = score * st.ssJ; k.mis = mis;
        k.lensScore = camDepth >= 1 ? Luminance(lensC) : F(0);
        out.push_back(k);
    }
}

// path.cpp:812-860 (bidirMIS = true)
template <class F> inline void HandleHitLight(const SceneT<F> &s, int camDepth, int light, const Ray<F> &r, const F screen[2],
                                              const BState<F> &st, std::vector<Contrib<F>> &out) {
    const LightT<F> &L = s.lights[light];
    V3<F> em;
    F dPdf, emPdf;
    LightEmission(L, r.dir, st.isect.sn, em, dPdf, emPdf);
    if (em.sum() > F(0)) {
        V3<F> c = st.thr.mul(em);
        F mis = F(1);
        if (camDepth > 0) {
            F pick = PickLightProb(s, light);
            dPdf *= pick;
            emPdf *= pick;
            F wC = MIS(dPdf) * st.prev + MIS(emPdf) * st.thisw;
            mis = F(1) / (F(1) + wC);
            c = c * mis;
        }
        F score = Luminance(c);
        if (score > F(0)) {
            Contrib<F> k;
            k.c = 2 + camDepth; k.l = 0; k.screen[0] = screen[0]; k.screen[1] = screen[1];
            k.contrib = c; k.lsScore = score; k.ssScore = score * st.ssJ; k.mis = mis;
            k.lensScore = camDepth >= 2 ? Luminance(st.lensC) : F(0);
            out.push_back(k);
        }
    }
}

// path.cpp:862-976
template <class F> inline void ConnectVertex(const SceneT<F> &s, const Opts<F> &o, bool occl, int camDepth, int lgtDepth, const BState<F> &ls,
                                             const SurfV<F> &lv, const BState<F> &cs, const SurfV<F> &cv, const F screen[2],
                                             std::vector<Contrib<F>> &out) {
    V3<F> toL = ls.isect.pos - cs.isect.pos;
    const F d2 = LengthSquared(toL), d = std::sqrt(d2);
    toL = toL * (F(1) / d);
    if (occl && Occluded(s, cs.isect.pos, toL, d)) return;
    const Material<F> &cm = MatOf(s, cv.prim);
    BsdfOut<F> cb;
    BsdfEval(cm, false, cs.wi, cs.isect.sn, toL, cb);
    if (cb.contrib.isZero()) return;
    F cf = SNC(false, cs.wi, cs.isect, toL);
    if (cf <= F(0)) return;
    cb.contrib = cb.contrib * cf;
    const Material<F> &lm = MatOf(s, lv.prim);
    BsdfOut<F> lb;
    BsdfEval(lm, true, ls.wi, ls.isect.sn, -toL, lb);
    if (lb.contrib.isZero()) return;
    F lf = SNC(true, ls.wi, ls.isect, -toL);
    if (lf <= F(0)) return;
    lb.contrib = lb.contrib * lf;
    const F G = F(1) / d2;
    V3<F> lensC = cs.lensC;
    if (camDepth == 0) {
        const bool camAbs = Roughness(cm, cv.disc) > o.roughThr, lgtAbs = Roughness(lm, lv.disc) > o.roughThr;
        lensC = (camAbs && lgtAbs) ? lb.contrib.mul(cb.contrib) * G : V3<F>(0, 0, 0);
    } else if (camDepth == 1) {
        if (Roughness(cm, cv.disc) > o.roughThr) {
            BsdfOut<F> imp;
            BsdfEval(cm, true, toL, cs.isect.sn, cs.wi, imp);
            lensC = lensC.mul(imp.contrib * SNC(true, toL, cs.isect, cs.wi));
        } else {
            lensC = V3<F>(0, 0, 0);
        }
    }
    const F camPdfA = cb.pdf * lb.cosWo * G;
    const F lgtPdfA = lb.pdf * cb.cosWo * G;
    const F wL = MIS(camPdfA) * (ls.prev + ls.thisw * MIS(lb.revPdf));
    const F wC = MIS(lgtPdfA) * (cs.prev + cs.thisw * MIS(cb.revPdf));
    const F mis = F(1) / (wL + F(1) + wC);
    V3<F> c = ls.thr.mul(cs.thr).mul(cb.contrib).mul(lb.contrib) * G;
    c = c * mis;
    const F ssJ = ls.ssJ * cs.ssJ;
    const F score = Luminance(c);
    if (score > F(0)) {
        Contrib<F> k;
        k.c = 2 + camDepth; k.l = 2 + lgtDepth; k.screen[0] = screen[0]; k.screen[1] = screen[1];
        k.contrib = c; k.lsScore = score; k.ssScore = score * ssJ; k.lensScore = Luminance(lensC); k.mis = mis;
        out.push_back(k);
    }
}

// path.cpp:1430-1478 (bidirectional LightCoordinateSampling)
template <class F> inline bool LightCoordinateSampling(const SceneT<F> &s, const Opts<F> &o, bool occl, int camDepth, const SurfV<F> &cur,
                                                       const SurfV<F> &nxt, BState<F> &st, V3<F> &dir, V3<F> &bc) {
    const Material<F> &m = MatOf(s, cur.prim);
    const int li = s.triLight[nxt.prim];
    V3<F> np, nn;
    F shapePdf;
    SampleTri(s, s.lights[li], nxt.prim, cur.rnd[0], cur.rnd[1], np, nn, shapePdf);
    dir = np - st.isect.pos;
    const F d2 = LengthSquared(dir), d = std::sqrt(d2);
    dir = dir * (F(1) / d);
    if (occl && Occluded(s, st.isect.pos, dir, d)) return false;
    BsdfOut<F> b;
    BsdfEval(m, false, st.wi, st.isect.sn, dir, b);
    if (b.contrib.isZero()) return false;
    bc = b.contrib * (F(1) / b.pdf);
    st.thr = st.thr.mul(bc);
    st.ssJ *= std::fabs(Dot(dir, nn) * (F(1) / d2)) * b.pdf;
    st.thisw = MIS(b.cosWo / b.pdf) * (st.thisw * MIS(b.revPdf) + st.prev);
    st.prev = MIS(F(1) / b.pdf);
    if (camDepth == 1) {  // path.cpp:1462-1476
        if (Roughness(m, cur.disc) > o.roughThr) {
            BsdfOut<F> imp;
            BsdfEval(m, true, dir, st.isect.sn, st.wi, imp);
            st.lensC = st.lensC.mul(imp.contrib * SNC(true, dir, st.isect, st.wi));
        } else {
            st.lensC = V3<F>(0, 0, 0);
        }
    }
    return true;
}

// Primary-sample layout: GetPathPos, path.cpp:2393-2439 (static scenes: no time dim)
template <class F> inline int Dim(int c, int l) { return std::max(c + l - 1, 2) * 2; }  // path.h:132-139 + path.cpp:2286-2289
template <class F> inline void GetPathPos(const PathS<F> &p, F *u) {
    int k = 0;
    if (p.l > 1) {
        u[k++] = p.lPos[0]; u[k++] = p.lPos[1]; u[k++] = p.lDir[0]; u[k++] = p.lDir[1];
        for (int d = 0; d < (int)p.lgt.size(); d++) {
            if (d == (int)p.lgt.size() - 1) break;
            u[k++] = p.lgt[d].rnd[0]; u[k++] = p.lgt[d].rnd[1];
        }
        if (p.c == 1) return;
    }
    u[k++] = p.screen[0]; u[k++] = p.screen[1];
    for (int d = 0; d < (int)p.cam.size(); d++) {
        if (d == (int)p.cam.size() - 1) {
            if (p.l == 1) { u[k++] = p.cam[d].dlRnd[0]; u[k++] = p.cam[d].dlRnd[1]; }
            return;
        }
        u[k++] = p.cam[d].rnd[0]; u[k++] = p.cam[d].rnd[1];
    }
}
template <class F> inline void SetPathPos(PathS<F> &p, const F *u) {
    int k = 0;
    if (p.l > 1) {
        p.lPos[0] = u[k++]; p.lPos[1] = u[k++]; p.lDir[0] = u[k++]; p.lDir[1] = u[k++];
        for (int d = 0; d < (int)p.lgt.size() - 1; d++) { p.lgt[d].rnd[0] = u[k++]; p.lgt[d].rnd[1] = u[k++]; }
        if (p.c == 1) return;
    }
    p.screen[0] = u[k++]; p.screen[1] = u[k++];
    for (int d = 0; d < (int)p.cam.size(); d++) {
        if (d == (int)p.cam.size() - 1) {
            if (p.l == 1) { p.cam[d].dlRnd[0] = u[k++]; p.cam[d].dlRnd[1] = u[k++]; }
            return;
        }
        p.cam[d].rnd[0] = u[k++]; p.cam[d].rnd[1] = u[k++];
    }
}

// path.cpp:1271-1274
template <class F> inline void Perturb(F &v, const F *off, int &k) { v = Modulo(v + off[k++], F(1)); }

// path.cpp:1480-1628. discOff[0..VMAX) are the N(0, discreteStdDev) draws for
// camera surface vertices, discOff[VMAX..2VMAX) for light surface vertices.
// occl = false and fixedPrims = true give the fixed-topology function the
// reference's AD twin differentiates (triangles read from the state, no
// occlusion tests, path.cpp:2987-3175).
template <class F> inline void PerturbPathBidir(const SceneT<F> &s, const Opts<F> &o, const F *off, const F *discOff, PathS<F> &p,
                                                std::vector<Contrib<F>> &out, bool occl = true, bool fixedPrims = false) {
    int k = 0;
    BState<F> ls;
    V3<F> lastLightOrg;
    auto isectAt = [&](SurfV<F> &v, const Ray<F> &r, Isect<F> &is) -> bool {
        if (fixedPrims) return ReIntersect(s, v.prim, r, is);
        if (!Intersect(s, r, is)) return false;
        v.prim = is.prim;
        return true;
    };
    if (p.l > 1) {
        const F pick = PickLightProb(s, p.lLight);
        Ray<F> r;
        Perturb(p.lPos[0], off, k); Perturb(p.lPos[1], off, k);
        Perturb(p.lDir[0], off, k); Perturb(p.lDir[1], off, k);
        EmitFromLight(s, o, pick, p, r, ls);
        r.minT = F(kIsectEps);
        r.maxT = std::numeric_limits<F>::infinity();
        V3<F> prevLensC(0, 0, 0);
        for (int d = 0; d < (int)p.lgt.size(); d++) {
            SurfV<F> &v = p.lgt[d];
            if (!isectAt(v, r, ls.isect)) return;
            ls.wi = -r.dir;
            v.disc = Modulo(v.disc + discOff[VMAX + d], F(1));
            ConvertMIS(r, ls);
            if (d == (int)p.lgt.size() - 1 && p.c == 1) {
                ConnectToCamera(s, o, occl, d, ls, v, prevLensC, r.org, out);
                return;
            }
            if (d == (int)p.lgt.size() - 1) break;
            Perturb(v.rnd[0], off, k); Perturb(v.rnd[1], off, k);
            V3<F> bc;
            if (!BSDFSampling(s, o, true, true, d, ls, v, ls, r.dir, bc)) return;
            ls.thr = ls.thr * v.rr;
            prevLensC = ls.lensC;
            r.org = ls.isect.pos;
        }
    }
    Perturb(p.screen[0], off, k); Perturb(p.screen[1], off, k);
    Ray<F> r;
    BState<F> cs;
    EmitFromCamera(s, p.screen, r, cs);
    bool lcp = false;
    for (int d = 0; d < (int)p.cam.size(); d++) {
        SurfV<F> &v = p.cam[d];
        bool hit = isectAt(v, r, cs.isect);
        cs.wi = -r.dir;
        if (hit) ConvertMIS(r, cs);
        if (d == (int)p.cam.size() - 1 && p.l == 0) {
            if (hit) {
                const int li = s.triLight[v.prim];
                if (li >= 0) HandleHitLight(s, d, li, r, p.screen, cs, out);
            }
            return;
        }
        if (!hit) return;
        v.disc = Modulo(v.disc + discOff[d], F(1));
        if (d == 1) {
            p.lensPos = cs.isect.pos;
            const F dd = DistanceSquared(cs.isect.pos, r.org);
            if (dd <= F(0)) { out.clear(); return; }
            cs.lensC = cs.lensC * (g_plant == 20 ? F(1) : F(1) / dd);  // 20: planted mistake (tests)
        }
        if (d == (int)p.cam.size() - 1) {
            if (p.l == 1) {
                const F pick = PickLightProb(s, v.dlLight);
                Perturb(v.dlRnd[0], off, k); Perturb(v.dlRnd[1], off, k);
                DirectLighting(s, o, occl, d, cs, p.screen, pick, v, out);
            } else {
                ConnectVertex(s, o, occl, d, (int)p.lgt.size() - 1, ls, p.lgt.back(), cs, v, p.screen, out);
            }
            return;
        }
        Perturb(v.rnd[0], off, k); Perturb(v.rnd[1], off, k);
        if (o.lightCoord && d == (int)p.cam.size() - 2 && p.l == 0) {
            const int nprim = p.cam.back().prim;
            if (nprim >= 0 && s.triLight[nprim] >= 0) lcp = true;
        }
        V3<F> bc;
        if (lcp) {
            if (!LightCoordinateSampling(s, o, occl, d, v, p.cam.back(), cs, r.dir, bc)) return;
        } else {
            if (!BSDFSampling(s, o, false, true, d, cs, v, cs, r.dir, bc)) return;
        }
        cs.thr = cs.thr * v.rr;
        r.org = cs.isect.pos;
        r.minT = F(kIsectEps);
        r.maxT = std::numeric_limits<F>::infinity();
    }
}

// path.cpp:1835-1877 (PerturbLensLastVertex): the light subpath's next-to-last
// vertex `cur` aims at the point the camera ray through `screen` meets on the
// last vertex `nxt` (found by a scene query; on nxt's own triangle when
// fixedPrims). Its bsdfRndParam becomes that direction's spherical coordinates.
template <class F> inline bool PerturbLensLastVertex(const SceneT<F> &s, const Opts<F> &o, bool occl, bool fixedPrims, const F screen[2],
                                                     BState<F> &st, SurfV<F> &cur, SurfV<F> &nxt, V3<F> &dir) {
    const Material<F> &m = MatOf(s, cur.prim);
    cur.useAbs = Roughness(m, cur.disc) > o.roughThr ? F(1) : F(0);
    if (cur.useAbs == F(0)) return false;
    Ray<F> r;
    SamplePrimary(s.cam, screen[0], screen[1], r);
    Isect<F> ni;
    if (fixedPrims) {
        if (!ReIntersect(s, nxt.prim, r, ni)) return false;
    } else {
        if (!Intersect(s, r, ni)) return false;
        nxt.prim = ni.prim;
    }
    dir = ni.pos - st.isect.pos;
    const F dist = Length(dir);
    dir = dir * (F(1) / dist);
    if (occl && Occluded(s, st.isect.pos, dir, dist)) return false;
    F jac;
    ToSphericalCoord(dir, cur.rnd[0], cur.rnd[1], jac);
    BsdfOut<F> b;
    BsdfEval(m, true, st.wi, st.isect.sn, dir, b);
    if (b.contrib.isZero() || b.pdf <= F(0)) return false;
    const F fac = SNC(true, dir, st.isect, st.wi);
    if (fac <= F(0)) return false;
    b.contrib = b.contrib * fac;
    st.lensC = b.contrib;
    const V3<F> bc = b.contrib * (F(1) / b.pdf);
    st.ssJ *= b.pdf * std::fabs(std::sin(std::acos(dir.z)));
    st.thisw = MIS(b.cosWo / b.pdf) * (st.thisw * MIS(b.revPdf) + st.prev);
    st.prev = MIS(F(1) / b.pdf);
    st.thr = st.thr.mul(bc);
    return true;
}
// path.cpp:1879-1912 (PerturbLensFirstVertex): the first camera vertex aims at
// the second one's position of the last evaluation (lensVertexPos).
template <class F> inline bool PerturbLensFirstVertex(const SceneT<F> &s, const Opts<F> &o, bool occl, const V3<F> &lensPos, BState<F> &st,
                                                      SurfV<F> &v, V3<F> &dir) {
    const Material<F> &m = MatOf(s, v.prim);
    v.useAbs = Roughness(m, v.disc) > o.roughThr ? F(1) : F(0);
    if (v.useAbs == F(0)) return false;
    dir = lensPos - st.isect.pos;
    const F dist = Length(dir);
    dir = dir * (F(1) / dist);
    if (occl && Occluded(s, st.isect.pos, dir, dist)) return false;
    F jac;
    ToSphericalCoord(dir, v.rnd[0], v.rnd[1], jac);
    BsdfOut<F> b;
    BsdfEval(m, false, st.wi, st.isect.sn, dir, b);
    if (b.contrib.isZero() || b.pdf <= F(0) || std::fabs(b.cosWo) < F(1e-2) || std::fabs(Dot(dir, st.isect.gn)) < F(1e-2)) return false;
    st.lensC = st.lensC.mul(b.contrib);
    const V3<F> bc = b.contrib * (F(1) / b.pdf);
    st.ssJ *= b.pdf * std::fabs(std::sin(std::acos(dir.z)));
    st.thisw = MIS(b.cosWo / b.pdf) * (st.thisw * MIS(b.revPdf) + st.prev);
    st.prev = MIS(F(1) / b.pdf);
    st.thr = st.thr.mul(bc);
    return true;
}
// path.cpp:1914-2067 (PerturbLensBidir): the state re-evaluated with its
// screen position moved to `screen`. Every vertex keeps its triangle except
// the one the new camera ray meets (the first camera vertex; for c = 1 the
// last light vertex), which a scene query finds; the first camera vertex (the
// next-to-last light vertex for c = 1) then aims at the fixed next vertex.
// occl = false and fixedPrims = true give the function whose derivatives in
// `screen` the reference's lens Gaussian takes (PathFuncMode::Lens,
// path.cpp:2987-3175).
template <class F> inline void PerturbLensBidir(const SceneT<F> &s, const Opts<F> &o, const F screen[2], PathS<F> &p,
                                                std::vector<Contrib<F>> &out, bool occl = true, bool fixedPrims = false) {
    BState<F> ls;
    if (p.l > 1) {
        const F pick = PickLightProb(s, p.lLight);
        Ray<F> r;
        EmitFromLight(s, o, pick, p, r, ls);
        r.minT = F(kIsectEps);
        r.maxT = std::numeric_limits<F>::infinity();
        V3<F> prevLensC(0, 0, 0);
        const int nl = (int)p.lgt.size();
        for (int d = 0; d < nl; d++) {
            SurfV<F> &v = p.lgt[d];
            if (!ReIntersect(s, v.prim, r, ls.isect)) return;
            ls.wi = -r.dir;
            ConvertMIS(r, ls);
            if (d == nl - 1 && p.c == 1) {
                ConnectToCamera(s, o, occl, d, ls, v, prevLensC, r.org, out);
                return;
            }
            if (d == nl - 1) break;
            V3<F> bc;
            if (d == nl - 2 && p.c == 1) {
                if (!PerturbLensLastVertex(s, o, occl, fixedPrims, screen, ls, v, p.lgt[nl - 1], r.dir)) return;
            } else {
                if (!BSDFSampling(s, o, true, true, d, ls, v, ls, r.dir, bc)) return;
            }
            ls.thr = ls.thr * v.rr;
            prevLensC = ls.lensC;
            r.org = ls.isect.pos;
        }
    }
    p.screen[0] = screen[0];
    p.screen[1] = screen[1];
    Ray<F> r;
    BState<F> cs;
    EmitFromCamera(s, p.screen, r, cs);
    bool lcp = false;
    const int nc = (int)p.cam.size();
    for (int d = 0; d < nc; d++) {
        SurfV<F> &v = p.cam[d];
        bool hit = true;
        if (d == 0 && !fixedPrims) {
            hit = Intersect(s, r, cs.isect);
            v.prim = hit ? cs.isect.prim : -1;
        } else if (v.prim < 0) {
            hit = false;
        } else if (!ReIntersect(s, v.prim, r, cs.isect)) {
            return;
        }
        cs.wi = -r.dir;
        if (hit) ConvertMIS(r, cs);
        if (d == nc - 1 && p.l == 0) {
            const int li = hit ? s.triLight[v.prim] : -1;
            if (li >= 0) HandleHitLight(s, d, li, r, p.screen, cs, out);
            return;
        }
        if (!hit) return;
        if (d == 1) {
            p.lensPos = cs.isect.pos;
            const F dd = DistanceSquared(cs.isect.pos, r.org);
            if (dd <= F(0)) { out.clear(); return; }
            cs.lensC = cs.lensC * (g_plant == 20 ? F(1) : F(1) / dd);  // 20: planted mistake (tests)
        }
        if (d == nc - 1) {
            if (p.l == 1) DirectLighting(s, o, false, d, cs, p.screen, PickLightProb(s, v.dlLight), v, out);
            else ConnectVertex(s, o, occl && d == 0, d, (int)p.lgt.size() - 1, ls, p.lgt.back(), cs, v, p.screen, out);
            return;
        }
        if (d == nc - 2 && p.l == 0) {
            const int nprim = p.cam.back().prim;
            if (nprim >= 0 && s.triLight[nprim] >= 0) lcp = true;
        }
        V3<F> bc;
        if (d == 0) {
            if (!PerturbLensFirstVertex(s, o, occl, p.lensPos, cs, v, r.dir)) return;
        } else if (lcp) {
            if (!LightCoordinateSampling(s, o, false, d, v, p.cam.back(), cs, r.dir, bc)) return;
        } else {
            if (!BSDFSampling(s, o, false, true, d, cs, v, cs, r.dir, bc)) return;
        }
        cs.thr = cs.thr * v.rr;
        r.org = cs.isect.pos;
        r.minT = F(kIsectEps);
        r.maxT = std::numeric_limits<F>::infinity();
    }
}

// path.cpp:475-487 (uniform order: pos(2), dir(2), light pick, triangle pick)
template <class F> inline void EmitFromLightInit(const SceneT<F> &s, PathS<F> &p, F &pick, Src<F> &g) {
    p.lPos[0] = g.uni(); p.lPos[1] = g.uni();
    p.lDir[0] = g.uni(); p.lDir[1] = g.uni();
    p.lLight = s.lightDist.SampleDiscrete(g.uni());
    pick = PickLightProb(s, p.lLight);
    const LightT<F> &L = s.lights[p.lLight];
    p.lPrim = L.firstTri + L.areaDist.SampleDiscrete(g.uni());
}
// path.cpp:178-186 (uniform order: light pick, rnd(2), triangle pick)
template <class F> inline void DirectLightingInit(const SceneT<F> &s, SurfV<F> &v, F &pick, Src<F> &g) {
    v.dlLight = s.lightDist.SampleDiscrete(g.uni());
    pick = PickLightProb(s, v.dlLight);
    v.dlRnd[0] = g.uni(); v.dlRnd[1] = g.uni();
    const LightT<F> &L = s.lights[v.dlLight];
    v.dlPrim = L.firstTri + L.areaDist.SampleDiscrete(g.uni());
}

// path.cpp:1123-1269 (bidirMIS = true). Uniform order: screen(2) is drawn
// after the light subpath; per vertex bsdfDiscrete before bsdfRndParam.
template <class F> inline void GenerateSubpath(const SceneT<F> &s, const Opts<F> &o, int camLen, int lgtLen, PathS<F> &p,
                                               std::vector<Contrib<F>> &out, Src<F> &g) {
    p.c = camLen; p.l = lgtLen;
    p.cam.clear(); p.lgt.clear();
    Ray<F> r;
    BState<F> ls;
    if (lgtLen > 1) {
        F pick = 1;
        EmitFromLightInit(s, p, pick, g);
        EmitFromLight(s, o, pick, p, r, ls);
        r.minT = F(kIsectEps);
        r.maxT = std::numeric_limits<F>::infinity();
        V3<F> prevLensC(0, 0, 0);
        for (int d = 0;; d++) {
            p.lgt.push_back(SurfV<F>());
            SurfV<F> &v = p.lgt.back();
            if (!Intersect(s, r, ls.isect)) return;
            v.prim = ls.isect.prim;
            v.disc = g.uni();
            ls.wi = -r.dir;
            ConvertMIS(r, ls);
            if (d + 2 == lgtLen) {
                if (camLen == 1) {
                    ConnectToCamera(s, o, true, d, ls, v, prevLensC, r.org, out);
                    return;
                }
                break;
            }
            v.rnd[0] = g.uni(); v.rnd[1] = g.uni();
            V3<F> bc;
            if (!BSDFSampling(s, o, true, false, d, ls, v, ls, r.dir, bc)) return;
            v.rr = 1;
            prevLensC = ls.lensC;
            r.org = ls.isect.pos;
        }
    }
    BState<F> cs;
    p.screen[0] = g.uni(); p.screen[1] = g.uni();
    EmitFromCamera(s, p.screen, r, cs);
    for (int d = 0;; d++) {
        p.cam.push_back(SurfV<F>());
        SurfV<F> &v = p.cam.back();
        bool hit = Intersect(s, r, cs.isect);
        if (hit) v.prim = cs.isect.prim;
        cs.wi = -r.dir;
        if (hit) ConvertMIS(r, cs);
        if (d + 2 >= camLen && lgtLen == 0) {
            if (hit && s.triLight[v.prim] >= 0) {
                const int li = s.triLight[v.prim];
                if (d > 1) {
                    SurfV<F> &pv = p.cam[p.cam.size() - 2];
                    GetSampleParam(s, v.prim, cs.isect.pos, pv.rnd[0], pv.rnd[1]);
                    V3<F> toPrev = cs.isect.pos - r.org;
                    const F d2 = LengthSquared(toPrev), id2 = F(1) / d2, id = std::sqrt(id2);
                    toPrev = toPrev * id;
                    cs.ssJ *= std::fabs(Dot(toPrev, cs.isect.sn) * id2) * (cs.lcJ * (F(1) / s.lights[li].totalArea));
                }
                HandleHitLight(s, d, li, r, p.screen, cs, out);
            }
            return;
        }
        if (!hit) return;
        if (d == 1) {
            p.lensPos = cs.isect.pos;
            const F dd = DistanceSquared(cs.isect.pos, r.org);
            if (dd <= F(0)) return;
            cs.lensC = cs.lensC * (g_plant == 20 ? F(1) : F(1) / dd);  // 20: planted mistake (tests)
        }
        v.disc = g.uni();
        if (d + 2 == camLen) {
            if (lgtLen == 1) {
                F pick = 1;
                DirectLightingInit(s, v, pick, g);
                DirectLighting(s, o, true, d, cs, p.screen, pick, v, out);
            } else {
                ConnectVertex(s, o, true, d, lgtLen - 2, ls, p.lgt.back(), cs, v, p.screen, out);
            }
            return;
        }
        v.rnd[0] = g.uni(); v.rnd[1] = g.uni();
        V3<F> bc;
        if (!BSDFSampling(s, o, false, false, d, cs, v, cs, r.dir, bc)) return;
        v.rr = 1;
        r.org = cs.isect.pos;
        r.minT = F(kIsectEps);
        r.maxT = std::numeric_limits<F>::infinity();
    }
}

// path.cpp:400-415 (RussianRoulette): one uniform per call
template <class F> inline bool RussianRoulette(int depth, const V3<F> &bc, F &rrW, V3<F> &thr, Src<F> &g) {
    F p = F(1);
    if (depth >= 3) p = std::min(bc.maxCoeff(), F(0.95));
    if (g.uni() > p) return false;
    rrW = F(1) / p;
    thr = thr * rrW;
    return true;
}

// path.cpp:978-1121 (GeneratePathBidir, all (s,t) strategies). Light-subpath
// states after the first are value-initialised in the reference, so a vertex
// sampled with relative parameterisation leaves ssJacobian == 0 there.
template <class F> inline void GeneratePathBidir(const SceneT<F> &s, const Opts<F> &o, int minDepth, int maxDepth, PathS<F> &p,
                                                 std::vector<Contrib<F>> &out, Src<F> &g) {
    p.cam.clear(); p.lgt.clear();
    std::vector<BState<F>> lsv(1);
    F pick = 1;
    EmitFromLightInit(s, p, pick, g);
    Ray<F> r;
    EmitFromLight(s, o, pick, p, r, lsv[0]);
    r.minT = F(kIsectEps);
    r.maxT = std::numeric_limits<F>::infinity();
    V3<F> prevLensC(0, 0, 0);
    for (int d = 0;; d++) {
        p.lgt.push_back(SurfV<F>());
        SurfV<F> &v = p.lgt.back();
        bool hit = Intersect(s, r, lsv[d].isect);
        if (!hit) { lsv.pop_back(); p.lgt.pop_back(); break; }
        v.prim = lsv[d].isect.prim;
        v.disc = g.uni();
        lsv[d].wi = -r.dir;
        ConvertMIS(r, lsv[d]);
        if (d + 2 >= minDepth) {
            if (!ConnectToCamera(s, o, true, d, lsv[d], v, prevLensC, r.org, out)) {}
        }
        if (maxDepth != -1 && d + 2 >= maxDepth) break;
        lsv.push_back(BState<F>());
        v.rnd[0] = g.uni(); v.rnd[1] = g.uni();
        V3<F> bc;
        if (!BSDFSampling(s, o, true, false, d, lsv[d], v, lsv[d + 1], r.dir, bc)) { lsv.pop_back(); break; }
        if (!RussianRoulette(d, bc, v.rr, lsv[d + 1].thr, g)) { lsv.pop_back(); break; }
        prevLensC = lsv[d + 1].lensC;
        r.org = lsv[d].isect.pos;
    }
    BState<F> cs;
    p.screen[0] = g.uni(); p.screen[1] = g.uni();
    EmitFromCamera(s, p.screen, r, cs);
    for (int d = 0;; d++) {
        p.cam.push_back(SurfV<F>());
        SurfV<F> &v = p.cam.back();
        bool hit = Intersect(s, r, cs.isect);
        if (hit) v.prim = cs.isect.prim;
        cs.wi = -r.dir;
        if (hit) ConvertMIS(r, cs);
        if (d + 1 >= minDepth) {
            const int li = hit ? s.triLight[v.prim] : -1;
            if (li >= 0) {
                if (o.lightCoord && d > 1) {
                    SurfV<F> &pv = p.cam[p.cam.size() - 2];
                    GetSampleParam(s, v.prim, cs.isect.pos, pv.rnd[0], pv.rnd[1]);
                    V3<F> toPrev = cs.isect.pos - r.org;
                    const F d2 = LengthSquared(toPrev), id2 = F(1) / d2, id = std::sqrt(id2);
                    toPrev = toPrev * id;
                    cs.ssJ *= std::fabs(Dot(toPrev, cs.isect.sn) * id2) * (cs.lcJ * (F(1) / s.lights[li].totalArea));
                }
                HandleHitLight(s, d, li, r, p.screen, cs, out);
                return;
            }
        }
        if (!hit || (maxDepth != -1 && d + 1 >= maxDepth)) break;
        if (d == 1) {
            p.lensPos = cs.isect.pos;
            const F dd = DistanceSquared(cs.isect.pos, r.org);
            if (dd <= F(0)) { out.clear(); return; }
            cs.lensC = cs.lensC * (g_plant == 20 ? F(1) : F(1) / dd);  // 20: planted mistake (tests)
        }
        v.disc = g.uni();
        if (d + 2 >= minDepth) {
            F dpick = 1;
            DirectLightingInit(s, v, dpick, g);
            DirectLighting(s, o, true, d, cs, p.screen, dpick, v, out);
        }
        int maxL = maxDepth == -1 ? (int)lsv.size() - 1 : std::min(maxDepth - d - 3, (int)lsv.size() - 1);
        for (int ld = 0; ld <= maxL; ld++)
            if (d + ld + 3 >= minDepth) ConnectVertex(s, o, true, d, ld, lsv[ld], p.lgt[ld], cs, v, p.screen, out);
        v.rnd[0] = g.uni(); v.rnd[1] = g.uni();
        V3<F> bc;
        if (!BSDFSampling(s, o, false, false, d, cs, v, cs, r.dir, bc)) break;
        if (!RussianRoulette(d, bc, v.rr, cs.thr, g)) break;
        r.org = cs.isect.pos;
        r.minT = F(kIsectEps);
        r.maxT = std::numeric_limits<F>::infinity();
    }
}

// path.h:87-98 (Path::Initialize): truncate to the (c,l) strategy.
template <class F> inline void Truncate(PathS<F> &p, int c, int l) {
    p.cam.resize(std::max(c - 1, 0));
    p.lgt.resize(std::max(l - 1, 0));
    p.c = c;
    p.l = l;
}

// ---- unidirectional path for the direct-lighting pre-pass (integrator.h:33-74)
struct UniDummy {};
template <class F> struct UState {
    Ray<F> ray;
    V3<F> thr;
    F lastPdf = 1;
    Isect<F> isect;
    V3<F> wi;
    F ssJ = 1, lcJ = 0;
};
// path.cpp:124-176 (uni HandleHitLight)
template <class F> inline void UniHandleHitLight(const SceneT<F> &s, int camDepth, int li, const UState<F> &st, const F screen[2], std::vector<Contrib<F>> &out) {
    const LightT<F> &L = s.lights[li];
    V3<F> em;
    F dPdf, emPdf;
    LightEmission(L, st.ray.dir, st.isect.sn, em, dPdf, emPdf);
    if (em.sum() > F(0)) {
        F d2 = DistanceSquared(st.ray.org, st.isect.pos);
        F cosT = -Dot(st.ray.dir, st.isect.sn);
        dPdf *= (d2 / cosT);
        V3<F> c = st.thr.mul(em);
        F mis = 1;
        if (camDepth > 0) {
            F pick = PickLightProb(s, li);
            F ratio = (dPdf * pick) / st.lastPdf;
            mis = F(1) / (F(1) + ratio * ratio);  // path.cpp:26-31 MISWeight
            c = c * mis;
        }
        F score = Luminance(c);
        if (score > F(0)) {
            Contrib<F> k;
            k.c = 2 + camDepth; k.l = 0; k.screen[0] = screen[0]; k.screen[1] = screen[1];
            k.contrib = c; k.lsScore = score; k.ssScore = score * st.ssJ; k.mis = mis;
            out.push_back(k);
        }
    }
}
// path.cpp:188-268 (uni DirectLighting; no shading-normal check here)
template <class F> inline void UniDirectLighting(const SceneT<F> &s, int camDepth, const UState<F> &st, const F screen[2], F pick, const SurfV<F> &v, std::vector<Contrib<F>> &out) {
    const Material<F> &m = MatOf(s, v.prim);
    const LightT<F> &L = s.lights[v.dlLight];
    V3<F> toL, lc;
    F dist, cosL, dPdf, emPdf;
    if (!LightSampleDirect(s, L, v.dlPrim, st.isect.pos, v.dlRnd[0], v.dlRnd[1], toL, dist, lc, cosL, dPdf, emPdf)) return;
    if (Occluded(s, st.isect.pos, toL, dist)) return;
    BsdfOut<F> b;
    BsdfEval(m, false, st.wi, st.isect.sn, toL, b);
    if (b.contrib.isZero()) return;
    V3<F> c = st.thr.mul(b.contrib).mul(lc) * (F(1) / pick);
    F ratio = b.pdf / (dPdf * pick);
    F mis = F(1) / (F(1) + ratio * ratio);
    c = c * mis;
    F score = Luminance(c);
    if (score > F(0)) {
        Contrib<F> k;
        k.c = 2 + camDepth; k.l = 1; k.screen[0] = screen[0]; k.screen[1] = screen[1];
        k.contrib = c; k.lsScore = score; k.ssScore = score * st.ssJ; k.mis = mis;
        out.push_back(k);
    }
}
// path.cpp:270-338 (uni BSDFSampling, perturb = false)
template <class F> inline bool UniBSDFSampling(const SceneT<F> &s, const Opts<F> &o, UState<F> &st, SurfV<F> &v, V3<F> &bc) {
    const Material<F> &m = MatOf(s, v.prim);
    BsdfOut<F> b;
    v.useAbs = Roughness(m, v.disc) > o.roughThr ? F(1) : F(0);
    V3<F> dir;
    if (!BsdfSample(m, false, st.wi, st.isect.sn, v.rnd[0], v.rnd[1], v.disc, dir, b)) return false;
    st.lastPdf = b.pdf;
    bc = b.contrib;
    if (v.useAbs == F(1)) {
        F jac;
        ToSphericalCoord(dir, v.rnd[0], v.rnd[1], jac);
        st.lcJ = F(1) / jac;
        st.ssJ *= jac * b.pdf;
    } else {
        st.lcJ = b.pdf;
    }
    st.thr = st.thr.mul(bc);
    st.ray.dir = dir;
    st.ray.org = st.isect.pos;
    return true;
}
// path.cpp:343-431 (GeneratePath for a pixel; uniform order: screen jitter
// (2), then per vertex disc, [light pick, rnd(2), tri pick], rnd(2), rr)
template <class F> inline void GeneratePathUni(const SceneT<F> &s, const Opts<F> &o, int px, int py, int minDepth, int maxDepth,
                                               std::vector<Contrib<F>> &out, Src<F> &g) {
    F screen[2];
    screen[0] = (F(px) + g.uni()) / F(s.cam.W);
    screen[1] = (F(py) + g.uni()) / F(s.cam.H);
    UState<F> st;
    st.thr = V3<F>(1, 1, 1);
    SamplePrimary(s.cam, screen[0], screen[1], st.ray);
    std::vector<SurfV<F>> verts;
    for (int d = 0;; d++) {
        verts.push_back(SurfV<F>());
        SurfV<F> &v = verts.back();
        bool hit = Intersect(s, st.ray, st.isect);
        if (hit) v.prim = st.isect.prim;
        const int li = hit ? s.triLight[v.prim] : -1;
        if (li >= 0 && d + 1 >= minDepth) {
            if (o.lightCoord && d > 1) {
                // path length >= 3 only; unreachable with maxDepth <= 2
            }
            UniHandleHitLight(s, d, li, st, screen, out);
            return;
        }
        if (!hit || (maxDepth != -1 && d + 1 >= maxDepth)) break;
        v.disc = g.uni();
        if (d == 1 && DistanceSquared(st.isect.pos, st.ray.org) <= F(0)) { out.clear(); return; }
        st.wi = -st.ray.dir;
        if (d + 2 >= minDepth) {
            F pick = 1;
            DirectLightingInit(s, v, pick, g);
            UniDirectLighting(s, d, st, screen, pick, v, out);
        }
        v.rnd[0] = g.uni(); v.rnd[1] = g.uni();
        V3<F> bc;
        if (!UniBSDFSampling(s, o, st, v, bc)) break;
        if (!RussianRoulette(d, bc, v.rr, st.thr, g)) break;
        st.ray.minT = F(kIsectEps);
        st.ray.maxT = std::numeric_limits<F>::infinity();
    }
}

}  // namespace orc

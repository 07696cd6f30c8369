// Device path space: fixed-topology re-evaluation of a bidirectional subpath
// (PerturbPathBidir, path.cpp:1480-1628), subpath generation for large steps
// (GenerateSubpath, path.cpp:1123-1269), full BDPT generation for the
// pre-pass (GeneratePathBidir, path.cpp:978-1121) and the unidirectional
// direct-lighting path (GeneratePath, path.cpp:343-431), with the MIS
// bookkeeping of path.cpp:433-976.
//
// A path state lives in memory behind an SRef (pointers + strides), so the
// same code reads the chain-state slot records in HBM (vertex fields
// interleaved, vertex stride 3) and the packed records of the C ABI (one
// array per field, stride 1). The evaluator is templated on
// the scalar type: float for the primal, jets for derivatives, in which case
// the triangles of the state are re-intersected instead of traced and no
// occlusion tests are made (the function the reference's AD twin
// differentiates, path.cpp:2987-3175).
//
// With the template flag LENS the evaluation also keeps the reference's lens
// bookkeeping (lensContrib per subpath state, the contribution's lensScore
// and lensVertexPos, path.cpp:433-976, 1430-1478): the h2mc integrator's lens
// mutation reads it. Without it (the DR loop's kernels) none of that code is
// compiled in.
#pragma once
#include "dev_scene.h"
#include "dev_bsdf.h"

namespace drmlt {

constexpr int kVMax = 16;   // surface vertices per subpath
constexpr int kDMax = 32;   // primary-sample dimensions

struct SRef {
    int *c, *l, *light, *lprim;
    float *screen;
    float *u;
    int *cprim, *lgprim;
    float *cdisc, *ldisc, *crr, *lrr;
    int s;   // stride of u and screen
    int vs;  // stride of the vertex fields (cprim .. lrr)
};

struct Opts {
    float roughThr;      // dptoptions.h:19
    int lightCoord;      // dptoptions.h:28
    int chadMisQuirk;    // path.cpp:2598-2604 (derivatives only)
};

template <class T> struct POut {
    bool valid;
    V3<T> contrib;
    T lsScore, ssScore;
    float screen[2];
    float mis;
    int c, l;
    T lensScore;         // LENS evaluations only
    V3<float> lensPos;   // LENS evaluations only: the second camera vertex (lensVertexPos)
};

// path.cpp:433-444
template <class T> struct BState {
    Isect<T> isect;
    V3<T> wi;
    T prev, thisw;
    V3<T> thr;
    T ssJ;
    float lcJ;
    V3<T> lensC;  // lensContrib (LENS evaluations only)
};

template <class T> DEV T MIS(const T &p) { return p * p; }

// path.cpp:39-58
template <class T> DEV T snc(bool adjoint, const V3<T> &wi, const Isect<T> &is, const V3<T> &wo) {
    const T cwi = dot(is.sn, wi), cwo = dot(is.sn, wo), wg = dot(is.gn, wi), og = dot(is.gn, wo);
    if (val(wg) * val(cwi) <= 0 || val(og) * val(cwo) <= 0) return T(0);
    if (adjoint) return s_fabs((og * cwi) / (wg * cwo));
    return T(1);
}

DEV Mat mat_of(const DevScene &S, int prim) { return load_mat(S.mats + 16 * S.tmat[prim]); }
DEV float modulo1(float a) { float r = fmodf(a, 1.0f); return r < 0.0f ? r + 1.0f : r; }  // utils.h:383-386

// A visibility test the caller takes over: the segment from o along the unit
// direction d, occluded as occluded() decides it for distance dist. The walk
// of dev_walk.h passes one so that the test is traced at its shared trace
// site; with none the test is made on the spot.
struct ShadowQ {
    V3<float> o, d;
    float dist;
};
template <class T> DEV bool occluded_or_defer(const DevScene &S, ShadowQ *sh, const V3<T> &org, const V3<T> &dir, const T &dist) {
    if (sh) {
        sh->o = lower(org);
        sh->d = lower(dir);
        sh->dist = float(val(dist));
        return false;
    }
    return occluded(S, org, dir, dist);
}
// path.cpp:455-474
template <class T> DEV void emit_from_camera(const DevScene &S, const T &sx, const T &sy, RayT<T> &r, BState<T> &st) {
    RayT<T> cr;
    sample_primary(S.cam, T(0.5f), T(0.5f), cr);
    sample_primary(S.cam, sx, sy, r);
    const T cosAtCam = dot(cr.dir, r.dir);
    const T ipcd = T(S.cam.dist) / cosAtCam;
    const T pdfW = ipcd * ipcd / cosAtCam;
    const float npix = float(S.cam.W * S.cam.H);
    st.thr = vconst<T>(1, 1, 1);
    st.prev = MIS(T(npix) / pdfW);
    st.thisw = T(0);
    st.ssJ = T(1);
    st.lensC = vconst<T>(1, 1, 1);
}
// path.cpp:489-505 + arealight.cpp:77-104
template <class T> DEV void emit_from_light(const DevScene &S, const Opts &o, int li, int prim, const T &p0, const T &p1, const T &d0,
                                            const T &d1, RayT<T> &r, BState<T> &st) {
    const float pick = pick_light_prob(S, li);
    V3<T> nrm;
    float shapePdf;
    sample_tri(S, li, prim, p0, p1, r.org, nrm, shapePdf);
    V3<T> d = sample_cos_hemisphere(d0, d1);
    V3<T> b0, b1;
    coord_system(nrm, b0, b1);
    r.dir = d.x * b0 + d.y * b1 + d.z * nrm;
    const float *e = S.lightEmit + 4 * li;
    const float k = float(3.14159265358979323846) / shapePdf;
    st.thr = vconst<T>(e[0] * k, e[1] * k, e[2] * k);
    T cosL = d.z;
    T emPdf = d.z * T(kINVPI) * T(shapePdf);
    T dPdf = T(shapePdf);
    emPdf = emPdf * T(pick);
    dPdf = dPdf * T(pick);
    st.thr = st.thr * T(1.0f / pick);
    st.prev = MIS(dPdf / emPdf);
    st.thisw = o.chadMisQuirk ? T(0) : MIS(cosL / emPdf);
    st.ssJ = T(1);
    st.lensC = vconst<T>(0, 0, 0);
}
// path.cpp:507-517
template <class T> DEV void convert_mis(const RayT<T> &r, BState<T> &st) {
    st.prev = st.prev * MIS(length2(r.org - st.isect.pos));
    T invC = T(1) / MIS(s_fabs(dot(r.dir, st.isect.sn)));
    st.prev = st.prev * invC;
    st.thisw = st.thisw * invC;
}

template <class T> DEV void emit(POut<T> &out, int c, int l, const V3<T> &contrib, const T &score, const T &ssJ, float sx, float sy, float mis) {
    out.valid = true;
    out.c = c; out.l = l;
    out.contrib = contrib;
    out.lsScore = score;
    out.ssScore = score * ssJ;
    out.screen[0] = sx; out.screen[1] = sy;
    out.mis = mis;
}

// path.cpp:519-613. Returns false when the reference clears the contribution
// list. LENS: prevLensC is the lensContrib of the light-subpath state before
// this vertex.
template <class T, bool FIXED, bool LENS = false>
DEV bool connect_to_camera(const DevScene &S, const Opts &o, int lgtDepth, const BState<T> &st, int prim, float disc, const V3<T> &prevPos,
                           POut<T> &out, ShadowQ *sh = nullptr, const V3<T> *prevLensC = nullptr) {
    RayT<T> cr;
    sample_primary(S.cam, T(0.5f), T(0.5f), cr);
    V3<T> toCam = cr.org - st.isect.pos;
    if (-val(dot(cr.dir, toCam)) <= 0) return true;
    T sx, sy;
    if (!project_point(S.cam, st.isect.pos, sx, sy)) return true;
    const T d2 = length2(toCam), d = s_sqrt(d2);
    toCam = toCam * (T(1) / d);
    if (!FIXED && occluded_or_defer(S, sh, st.isect.pos, toCam, d)) return true;
    const Mat m = mat_of(S, prim);
    BOut<T> b;
    bsdf_eval(m, true, st.wi, st.isect.sn, toCam, b);
    if (is_zero(b.contrib)) return true;
    T fac = snc(true, st.wi, st.isect, toCam);
    if (val(fac) <= 0) return true;
    b.contrib = b.contrib * fac;
    V3<T> lensC;
    if constexpr (LENS) lensC = vconst<T>(0, 0, 0);
    if (roughness(m, disc) > o.roughThr && lgtDepth >= 1) {
        const T dd = length2(st.isect.pos - prevPos);
        if (val(dd) <= 0) { out.valid = false; return false; }
        if constexpr (LENS) {
            BOut<T> e;
            bsdf_eval(m, false, toCam, st.isect.sn, st.wi, e);
            lensC = mul(e.contrib, *prevLensC) * (T(1) / dd);
        }
    }
    const T cosAtCam = -dot(cr.dir, toCam);
    const T ipcd = T(S.cam.dist) / cosAtCam;
    const T i2sa = ipcd * ipcd / cosAtCam;
    const T i2surf = i2sa * s_fabs(b.cosWo) / d2;
    const float npix = float(S.cam.W * S.cam.H);
    const T wLight = MIS(i2surf / T(npix)) * (st.prev + st.thisw * MIS(b.revPdf));
    const T mis = T(1) / (wLight + T(1));
    const T surf2img = b.cosWo / i2surf;
    V3<T> c = mul((b.contrib * mis) * (T(1) / (T(npix) * surf2img)), st.thr);
    const T score = luminance(c);
    if (val(score) > 0) {
        emit(out, 1, 2 + lgtDepth, c, score, st.ssJ, float(val(sx)), float(val(sy)), float(val(mis)));
        if constexpr (LENS) out.lensScore = luminance(lensC);
    }
    return true;
}

// path.cpp:615-724 (BSDFSampling<adjoint, perturb>); cur and next may alias.
// rnd holds bsdfRndParam; for a generated absolute-parameterised vertex it is
// replaced by the spherical coordinates of the sampled direction. LENS:
// lensContrib at subpath vertex `depth` (path.cpp:664-711).
template <class T, bool LENS = false>
DEV bool bsdf_sampling(const DevScene &S, const Opts &o, bool adjoint, bool perturb, const BState<T> &cur, int prim, float disc, T &rnd0, T &rnd1,
                       BState<T> &next, V3<T> &dir, V3<T> &bcOut, int depth = 0) {
    const Mat m = mat_of(S, prim);
    BOut<T> b;
    const bool useAbs = roughness(m, disc) > o.roughThr;
    V3<T> bc;
    T ssJ = cur.ssJ;
    float lcJ;
    bool setSS = false;
    if (!perturb || !useAbs) {
        if (!bsdf_sample(m, adjoint, cur.wi, cur.isect.sn, rnd0, rnd1, disc, dir, b)) return false;
        bc = b.contrib;
        if (useAbs) {
            float u0, u1, jac;
            to_spherical(lower(dir), u0, u1, jac);
            rnd0 = T(u0); rnd1 = T(u1);
            lcJ = 1.0f / jac;
            ssJ = cur.ssJ * (T(jac) * b.pdf);
            setSS = true;
        } else {
            lcJ = float(val(b.pdf));
        }
    } else {
        T jac;
        dir = sample_sphere(rnd0, rnd1, jac);
        bsdf_eval(m, adjoint, cur.wi, cur.isect.sn, dir, b);
        if (is_zero(b.contrib) || val(b.pdf) <= 0) return false;
        bc = b.contrib * (T(1) / b.pdf);
        lcJ = 1.0f / float(val(jac));
        ssJ = cur.ssJ * (jac * b.pdf);
        setSS = true;
    }
    T fac = snc(adjoint, cur.wi, cur.isect, dir);
    if (val(fac) <= 0) return false;
    if constexpr (LENS) {
        // bc is still contrib / pdf here
        const bool grazing = fabsf(float(val(b.cosWo))) < 1e-2f || fabsf(float(val(dot(dir, cur.isect.gn)))) < 1e-2f;
        V3<T> lc;
        if (useAbs) {
            if (adjoint) {
                lc = grazing ? vconst<T>(0, 0, 0) : bc * b.pdf;
            } else if (depth == 0) {
                lc = grazing ? vconst<T>(0, 0, 0) : mul(cur.lensC, bc * b.pdf);
            } else if (depth == 1) {
                BOut<T> imp;
                bsdf_eval(m, true, dir, cur.isect.sn, cur.wi, imp);
                if (fabsf(float(val(imp.cosWo))) < 1e-2f || fabsf(float(val(dot(cur.wi, cur.isect.gn)))) < 1e-2f)
                    lc = vconst<T>(0, 0, 0);
                else
                    lc = mul(cur.lensC, imp.contrib * snc(true, dir, cur.isect, cur.wi));
            } else {
                lc = cur.lensC;
            }
        } else {
            lc = (adjoint || depth <= 1) ? vconst<T>(0, 0, 0) : cur.lensC;
        }
        next.lensC = adjoint ? lc * fac : lc;
    }
    bc = bc * fac;
    bcOut = bc;
    const T newThis = MIS(b.cosWo / b.pdf) * (cur.thisw * MIS(b.revPdf) + cur.prev);
    next.lcJ = lcJ;
    if (setSS) next.ssJ = ssJ;
    next.thisw = newThis;
    next.prev = MIS(T(1) / b.pdf);
    next.thr = mul(cur.thr, bc);
    return true;
}

// path.cpp:726-810 (bidirectional DirectLighting)
// LENS: disc is the vertex's bsdfDiscrete.
template <class T, bool FIXED, bool LENS = false>
DEV void direct_lighting(const DevScene &S, const Opts &o, int camDepth, const BState<T> &st, float sx, float sy, int li, int lprim, const T &r0,
                         const T &r1, int prim, POut<T> &out, ShadowQ *sh = nullptr, float disc = 0.0f) {
    const float pick = pick_light_prob(S, li);
    const Mat m = mat_of(S, prim);
    V3<T> toL, lc;
    T dist, cosL, dPdf, emPdf;
    if (!light_sample_direct(S, li, lprim, st.isect.pos, r0, r1, toL, dist, lc, cosL, dPdf, emPdf)) return;
    if (!FIXED && occluded_or_defer(S, sh, st.isect.pos, toL, dist)) return;
    BOut<T> b;
    bsdf_eval(m, false, st.wi, st.isect.sn, toL, b);
    if (is_zero(b.contrib)) return;
    T fac = snc(false, st.wi, st.isect, toL);
    if (val(fac) <= 0) return;
    b.contrib = b.contrib * fac;
    V3<T> c = mul(mul(st.thr, b.contrib), lc) * T(1.0f / pick);
    const T wL = MIS(b.pdf / (T(pick) * dPdf));
    const T wC = MIS(emPdf * b.cosWo / (dPdf * cosL)) * (st.prev + st.thisw * MIS(b.revPdf));
    const T mis = T(1) / (wL + T(1) + wC);
    c = c * mis;
    const T score = luminance(c);
    if (val(score) > 0) {
        emit(out, 2 + camDepth, 1, c, score, st.ssJ, sx, sy, float(val(mis)));
        if constexpr (LENS) {
            V3<T> lensC = st.lensC;
            if (camDepth == 1) {
                if (roughness(m, disc) > o.roughThr) {
                    BOut<T> imp;
                    bsdf_eval(m, true, toL, st.isect.sn, st.wi, imp);
                    lensC = mul(lensC, imp.contrib * snc(true, toL, st.isect, st.wi));
                } else {
                    lensC = vconst<T>(0, 0, 0);
                }
            }
            out.lensScore = camDepth >= 1 ? luminance(lensC) : T(0);
        }
    }
}
// path.cpp:812-860 (bidirMIS = true)
template <class T, bool LENS = false>
DEV void handle_hit_light(const DevScene &S, int camDepth, int li, const RayT<T> &r, float sx, float sy, const BState<T> &st, POut<T> &out) {
    V3<T> em;
    T dPdf, emPdf;
    light_emission(S, li, r.dir, st.isect.sn, em, dPdf, emPdf);
    if (val(em.x + em.y + em.z) > 0) {
        V3<T> c = mul(st.thr, em);
        float misf = 1.0f;
        if (camDepth > 0) {
            const float pick = pick_light_prob(S, li);
            dPdf = dPdf * T(pick);
            emPdf = emPdf * T(pick);
            T wC = MIS(dPdf) * st.prev + MIS(emPdf) * st.thisw;
            T mis = T(1) / (T(1) + wC);
            c = c * mis;
            misf = float(val(mis));
        }
        T score = luminance(c);
        if (val(score) > 0) {
            emit(out, 2 + camDepth, 0, c, score, st.ssJ, sx, sy, misf);
            if constexpr (LENS) out.lensScore = camDepth >= 2 ? luminance(st.lensC) : T(0);
        }
    }
}
// path.cpp:862-976
// LENS: o, cdisc and ldisc (the two vertices' bsdfDiscrete) decide their
// parameterisations.
template <class T, bool FIXED, bool LENS = false>
DEV void connect_vertex(const DevScene &S, int camDepth, int lgtDepth, const BState<T> &ls, int lprimV, const BState<T> &cs, int cprimV, float sx,
                        float sy, POut<T> &out, ShadowQ *sh = nullptr, const Opts *o = nullptr, float cdisc = 0.0f, float ldisc = 0.0f) {
    V3<T> toL = ls.isect.pos - cs.isect.pos;
    const T d2 = length2(toL), d = s_sqrt(d2);
    toL = toL * (T(1) / d);
    if (!FIXED && occluded_or_defer(S, sh, cs.isect.pos, toL, d)) return;
    const Mat cm = mat_of(S, cprimV);
    BOut<T> cb;
    bsdf_eval(cm, false, cs.wi, cs.isect.sn, toL, cb);
    if (is_zero(cb.contrib)) return;
    T cf = snc(false, cs.wi, cs.isect, toL);
    if (val(cf) <= 0) return;
    cb.contrib = cb.contrib * cf;
    const Mat lm = mat_of(S, lprimV);
    BOut<T> lb;
    bsdf_eval(lm, true, ls.wi, ls.isect.sn, -toL, lb);
    if (is_zero(lb.contrib)) return;
    T lf = snc(true, ls.wi, ls.isect, -toL);
    if (val(lf) <= 0) return;
    lb.contrib = lb.contrib * lf;
    const T G = T(1) / d2;
    const T camPdfA = cb.pdf * lb.cosWo * G;
    const T lgtPdfA = lb.pdf * cb.cosWo * G;
    const T wL = MIS(camPdfA) * (ls.prev + ls.thisw * MIS(lb.revPdf));
    const T wC = MIS(lgtPdfA) * (cs.prev + cs.thisw * MIS(cb.revPdf));
    const T mis = T(1) / (wL + T(1) + wC);
    V3<T> c = mul(mul(mul(ls.thr, cs.thr), cb.contrib), lb.contrib) * G;
    c = c * mis;
    const T ssJ = ls.ssJ * cs.ssJ;
    const T score = luminance(c);
    if (val(score) > 0) {
        emit(out, 2 + camDepth, 2 + lgtDepth, c, score, ssJ, sx, sy, float(val(mis)));
        if constexpr (LENS) {
            V3<T> lensC = cs.lensC;
            if (camDepth == 0) {
                const bool camAbs = roughness(cm, cdisc) > o->roughThr, lgtAbs = roughness(lm, ldisc) > o->roughThr;
                lensC = (camAbs && lgtAbs) ? mul(lb.contrib, cb.contrib) * G : vconst<T>(0, 0, 0);
            } else if (camDepth == 1) {
                if (roughness(cm, cdisc) > o->roughThr) {
                    BOut<T> imp;
                    bsdf_eval(cm, true, toL, cs.isect.sn, cs.wi, imp);
                    lensC = mul(lensC, imp.contrib * snc(true, toL, cs.isect, cs.wi));
                } else {
                    lensC = vconst<T>(0, 0, 0);
                }
            }
            out.lensScore = luminance(lensC);
        }
    }
}
// path.cpp:1430-1478 (bidirectional LightCoordinateSampling)
// LENS: lensContrib at camera vertex camDepth (path.cpp:1462-1476); o and
// disc (the vertex's bsdfDiscrete) decide its parameterisation.
template <class T, bool FIXED, bool LENS = false>
DEV bool light_coordinate_sampling(const DevScene &S, int prim, const T &r0, const T &r1, int nprim, BState<T> &st, V3<T> &dir,
                                   ShadowQ *sh = nullptr, int camDepth = -1, const Opts *o = nullptr, float disc = 0.0f) {
    const Mat m = mat_of(S, prim);
    const int li = S.tlight[nprim];
    V3<T> np, nn;
    float shapePdf;
    sample_tri(S, li, nprim, r0, r1, np, nn, shapePdf);
    dir = np - st.isect.pos;
    const T d2 = length2(dir), d = s_sqrt(d2);
    dir = dir * (T(1) / d);
    if (!FIXED && occluded_or_defer(S, sh, st.isect.pos, dir, d)) return false;
    BOut<T> b;
    bsdf_eval(m, false, st.wi, st.isect.sn, dir, b);
    if (is_zero(b.contrib)) return false;
    V3<T> bc = b.contrib * (T(1) / b.pdf);
    st.thr = mul(st.thr, bc);
    st.ssJ = st.ssJ * (s_fabs(dot(dir, nn) * (T(1) / d2)) * b.pdf);
    st.thisw = MIS(b.cosWo / b.pdf) * (st.thisw * MIS(b.revPdf) + st.prev);
    st.prev = MIS(T(1) / b.pdf);
    if constexpr (LENS) {
        if (camDepth == 1) {
            if (roughness(m, disc) > o->roughThr) {
                BOut<T> imp;
                bsdf_eval(m, true, dir, st.isect.sn, st.wi, imp);
                st.lensC = mul(st.lensC, imp.contrib * snc(true, dir, st.isect, st.wi));
            } else {
                st.lensC = vconst<T>(0, 0, 0);
            }
        }
    }
    return true;
}

// ------------------------------------------------------------------ u providers
// Primal: u_k <- (u_k + off_k) mod 1, written back in place (path.cpp:1271-1274).
template <class OFF> struct PerturbU {
    float *u;
    int s, k;
    OFF off;
    DEV float next() {
        float v = modulo1(u[k * s] + off(k));
        u[k * s] = v;
        k++;
        return v;
    }
};
struct OffArray {
    const float *o;
    int s;
    DEV float operator()(int k) const { return o[k * s]; }
};
struct OffZero {
    DEV float operator()(int) const { return 0.0f; }
};
// Green's y* offset: x_k - y_k (delayed.cpp:476-482)
struct OffDiff {
    const float *x, *y;
    int sx, sy;
    DEV float operator()(int k) const { return x[k * sx] - y[k * sy]; }
};
// Jets: u_k seeded with t = [k == row] and g[j] = [k == col0 + j].
template <class S, int K> struct JetU {
    const float *u;
    int s, k, row, col0;
    DEV Jet<S, K> next() {
        Jet<S, K> r = Jet<S, K>(S(u[k * s]));
        r.t = k == row ? S(1) : S(0);
#pragma unroll
        for (int j = 0; j < K; j++) r.g[j] = (k == col0 + j) ? S(1) : S(0);
        k++;
        return r;
    }
};

// The state's own u, unperturbed, in T: the primal of the fixed-topology
// function (evaluate_path_bidir_{c}_{l}_static, src/chad.cpp:183-238).
template <class T> struct PlainU {
    const float *u;
    int s, k;
    DEV T next() { return T(u[k++ * s]); }
};

// Discrete-dimension perturbation (path.cpp:1482 etc.): bsdfDiscrete <- (d + n) mod 1
struct DiscArray {
    const float *d;
    int s;
    DEV float operator()(int slot) const { return d[slot * s]; }
};
struct DiscNone {
    DEV float operator()(int) const { return 0.0f; }
};

// PerturbPathBidir (path.cpp:1480-1628) on the state behind `ref`, in place.
// U yields the perturbed primary samples in GetPathPos order; DISC(slot) gives
// the discrete offset of camera vertex slot d (slot d) or light vertex d (slot
// kVMax + d). FIXED: re-intersect the state's triangles, no occlusion, no writes.
// LENS: also out.lensScore and out.lensPos.
template <class T, bool FIXED, class U, class DISC, bool LENS = false>
DEV void perturb_eval(const DevScene &S, const Opts &o, const SRef &ref, U &uprov, const DISC &disc, POut<T> &out) {
    out.valid = false;
    const int c = *ref.c, l = *ref.l, s = ref.s, vs = ref.vs;
    const int nl = l - 1, nc = c - 1;
    BState<T> ls;
    int lastLightPrim = -1;
    if (l > 1) {
        const int li = *ref.light, lp = *ref.lprim;
        T p0 = uprov.next(), p1 = uprov.next(), d0 = uprov.next(), d1 = uprov.next();
        RayT<T> r;
        emit_from_light(S, o, li, lp, p0, p1, d0, d1, r, ls);
        r.minT = kIsectEps;
        r.maxT = __builtin_inff();
        V3<T> prevLensC;
        if constexpr (LENS) prevLensC = vconst<T>(0, 0, 0);
        for (int d = 0; d < nl; d++) {
            int prim;
            if (FIXED) {
                prim = ref.lgprim[d * vs];
                if (!re_intersect(S, prim, r, ls.isect)) return;
            } else {
                if (!intersect(S, r, ls.isect)) return;
                prim = ls.isect.prim;
                ref.lgprim[d * vs] = prim;
            }
            ls.wi = -r.dir;
            float dv = ref.ldisc[d * vs];
            if (!FIXED) { dv = modulo1(dv + disc(kVMax + d)); ref.ldisc[d * vs] = dv; }
            convert_mis(r, ls);
            lastLightPrim = prim;
            if (d == nl - 1 && c == 1) {
                connect_to_camera<T, FIXED, LENS>(S, o, d, ls, prim, dv, r.org, out, nullptr, &prevLensC);
                if (out.valid && !FIXED) { ref.screen[0] = out.screen[0]; ref.screen[s] = out.screen[1]; }
                return;
            }
            if (d == nl - 1) break;
            T r0 = uprov.next(), r1 = uprov.next();
            V3<T> dir, bc;
            if (!bsdf_sampling<T, LENS>(S, o, true, true, ls, prim, dv, r0, r1, ls, dir, bc, d)) return;
            ls.thr = ls.thr * T(ref.lrr[d * vs]);
            if constexpr (LENS) prevLensC = ls.lensC;
            r.org = ls.isect.pos;
            r.dir = dir;
        }
    }
    T sx = uprov.next(), sy = uprov.next();
    const float sxf = float(val(sx)), syf = float(val(sy));
    RayT<T> r;
    BState<T> cs;
    emit_from_camera(S, sx, sy, r, cs);
    bool lcp = false;
    for (int d = 0; d < nc; d++) {
        int prim = -1;
        bool hit;
        if (FIXED) {
            prim = ref.cprim[d * vs];
            hit = re_intersect(S, prim, r, cs.isect);
        } else {
            hit = intersect(S, r, cs.isect);
            if (hit) { prim = cs.isect.prim; ref.cprim[d * vs] = prim; }
        }
        cs.wi = -r.dir;
        if (hit) convert_mis(r, cs);
        if (d == nc - 1 && l == 0) {
            if (hit) {
                const int li = S.tlight[prim];
                if (li >= 0) handle_hit_light<T, LENS>(S, d, li, r, sxf, syf, cs, out);
            }
            return;
        }
        if (!hit) return;
        float dv = ref.cdisc[d * vs];
        if (!FIXED) { dv = modulo1(dv + disc(d)); ref.cdisc[d * vs] = dv; }
        if (d == 1) {
            const T dd = length2(cs.isect.pos - r.org);
            if (val(dd) <= 0) { out.valid = false; return; }
            if constexpr (LENS) {
                out.lensPos = lower(cs.isect.pos);
                cs.lensC = cs.lensC * (T(1) / dd);
            }
        }
        if (d == nc - 1) {
            if (l == 1) {
                T r0 = uprov.next(), r1 = uprov.next();
                direct_lighting<T, FIXED, LENS>(S, o, d, cs, sxf, syf, *ref.light, *ref.lprim, r0, r1, prim, out, nullptr, dv);
            } else {
                connect_vertex<T, FIXED, LENS>(S, d, nl - 1, ls, lastLightPrim, cs, prim, sxf, syf, out, nullptr, &o, dv,
                                               nl > 0 ? ref.ldisc[(nl - 1) * vs] : 0.0f);
            }
            return;
        }
        T r0 = uprov.next(), r1 = uprov.next();
        if (o.lightCoord && d == nc - 2 && l == 0) {
            const int np = ref.cprim[(nc - 1) * vs];
            if (np >= 0 && S.tlight[np] >= 0) lcp = true;
        }
        V3<T> dir, bc;
        if (lcp) {
            if (!light_coordinate_sampling<T, FIXED, LENS>(S, prim, r0, r1, ref.cprim[(nc - 1) * vs], cs, dir, nullptr, d, &o, dv)) return;
        } else {
            if (!bsdf_sampling<T, LENS>(S, o, false, true, cs, prim, dv, r0, r1, cs, dir, bc, d)) return;
        }
        cs.thr = cs.thr * T(ref.crr[d * vs]);
        r.org = cs.isect.pos;
        r.dir = dir;
        r.minT = kIsectEps;
        r.maxT = __builtin_inff();
    }
}

// ------------------------------------------------------------------ lens mutation
// |sin(acos(z))| of a unit direction's z, in a form jets carry.
template <class T> DEV T sin_of_acos(const T &z) { return s_sqrt(s_max(T(1) - z * z, typename Base<T>::type(0))); }

// path.cpp:1879-1912 (PerturbLensFirstVertex): the first camera vertex aims
// at lensPos, the second vertex's position of the state's last evaluation.
// Not FIXED: occlusion tested, the direction's spherical coordinates written
// to u[k], u[k + 1] (the vertex's bsdfRndParam).
template <class T, bool FIXED>
DEV bool lens_first_vertex(const DevScene &S, const Opts &o, const V3<float> &lensPos, BState<T> &st, int prim, float disc, const SRef &ref,
                           int k, V3<T> &dir) {
    const Mat m = mat_of(S, prim);
    if (!(roughness(m, disc) > o.roughThr)) return false;
    dir = lift<T>(lensPos) - st.isect.pos;
    const T dist = s_sqrt(length2(dir));
    dir = dir * (T(1) / dist);
    if constexpr (!FIXED) {
        if (occluded(S, st.isect.pos, dir, dist)) return false;
        float u0, u1, jac;
        to_spherical(lower(dir), u0, u1, jac);
        ref.u[k * ref.s] = u0;
        ref.u[(k + 1) * ref.s] = u1;
    }
    BOut<T> b;
    bsdf_eval(m, false, st.wi, st.isect.sn, dir, b);
    if (is_zero(b.contrib) || val(b.pdf) <= 0 || fabsf(float(val(b.cosWo))) < 1e-2f || fabsf(float(val(dot(dir, st.isect.gn)))) < 1e-2f)
        return false;
    st.lensC = mul(st.lensC, b.contrib);
    const V3<T> bc = b.contrib * (T(1) / b.pdf);
    st.ssJ = st.ssJ * (b.pdf * sin_of_acos(dir.z));
    st.thisw = MIS(b.cosWo / b.pdf) * (st.thisw * MIS(b.revPdf) + st.prev);
    st.prev = MIS(T(1) / b.pdf);
    st.thr = mul(st.thr, bc);
    return true;
}
// path.cpp:1835-1877 (PerturbLensLastVertex): the light subpath's
// next-to-last vertex aims at the point the camera ray through scr meets on
// the last vertex (light vertex slot nxt): a ray query, which names that
// vertex's new triangle (not FIXED), or its own triangle (FIXED).
template <class T, bool FIXED>
DEV bool lens_last_vertex(const DevScene &S, const Opts &o, const T scr[2], BState<T> &st, int prim, float disc, const SRef &ref, int nxt,
                          int k, V3<T> &dir) {
    const Mat m = mat_of(S, prim);
    if (!(roughness(m, disc) > o.roughThr)) return false;
    RayT<T> r;
    sample_primary(S.cam, scr[0], scr[1], r);
    Isect<T> ni;
    if constexpr (FIXED) {
        if (!re_intersect(S, ref.lgprim[nxt * ref.vs], r, ni)) return false;
    } else {
        if (!intersect(S, r, ni)) return false;
        ref.lgprim[nxt * ref.vs] = ni.prim;
    }
    dir = ni.pos - st.isect.pos;
    const T dist = s_sqrt(length2(dir));
    dir = dir * (T(1) / dist);
    if constexpr (!FIXED) {
        if (occluded(S, st.isect.pos, dir, dist)) return false;
        float u0, u1, jac;
        to_spherical(lower(dir), u0, u1, jac);
        ref.u[k * ref.s] = u0;
        ref.u[(k + 1) * ref.s] = u1;
    }
    BOut<T> b;
    bsdf_eval(m, true, st.wi, st.isect.sn, dir, b);
    if (is_zero(b.contrib) || val(b.pdf) <= 0) return false;
    const T fac = snc(true, dir, st.isect, st.wi);
    if (val(fac) <= 0) return false;
    b.contrib = b.contrib * fac;
    st.lensC = b.contrib;
    const V3<T> bc = b.contrib * (T(1) / b.pdf);
    st.ssJ = st.ssJ * (b.pdf * sin_of_acos(dir.z));
    st.thisw = MIS(b.cosWo / b.pdf) * (st.thisw * MIS(b.revPdf) + st.prev);
    st.prev = MIS(T(1) / b.pdf);
    st.thr = mul(st.thr, bc);
    return true;
}

// PerturbLensBidir (path.cpp:1914-2067) of the state behind `ref` with its
// screen position moved to scr. Every vertex keeps its triangle (the state's
// own primary samples, no offsets, no discrete perturbation) except the one
// the new camera ray meets: the first camera vertex, or for c = 1 the last
// light vertex, found by a ray query. Not FIXED: the lens mutation's
// proposal; writes the new triangle, the re-aimed vertex's primary samples,
// the screen position and the contribution's screen position into the state.
// FIXED: re-intersection only, no occlusion tests, no writes -- the lens
// function whose derivatives in scr give the lens Gaussian (PathFuncMode::
// Lens, path.cpp:2987-3175). Tracks the lens bookkeeping (LENS).
template <class T, bool FIXED>
DEV void perturb_lens_eval(const DevScene &S, const Opts &o, const SRef &ref, const T scr[2], const V3<float> &lensPos, POut<T> &out) {
    out.valid = false;
    const int c = *ref.c, l = *ref.l, s = ref.s, vs = ref.vs;
    const int nl = l - 1, nc = c - 1;
    BState<T> ls;
    int lastLightPrim = -1;
    int k = 0;  // GetPathPos index of the next primary sample
    if (l > 1) {
        const int li = *ref.light, lp = *ref.lprim;
        RayT<T> r;
        emit_from_light(S, o, li, lp, T(ref.u[0]), T(ref.u[s]), T(ref.u[2 * s]), T(ref.u[3 * s]), r, ls);
        k = 4;
        r.minT = kIsectEps;
        r.maxT = __builtin_inff();
        V3<T> prevLensC = vconst<T>(0, 0, 0);
        for (int d = 0; d < nl; d++) {
            const int prim = ref.lgprim[d * vs];
            if (!re_intersect(S, prim, r, ls.isect)) return;
            ls.wi = -r.dir;
            const float dv = ref.ldisc[d * vs];
            convert_mis(r, ls);
            lastLightPrim = prim;
            if (d == nl - 1 && c == 1) {
                connect_to_camera<T, FIXED, true>(S, o, d, ls, prim, dv, r.org, out, nullptr, &prevLensC);
                if (out.valid && !FIXED) { ref.screen[0] = out.screen[0]; ref.screen[s] = out.screen[1]; }
                return;
            }
            if (d == nl - 1) break;
            V3<T> dir, bc;
            if (d == nl - 2 && c == 1) {
                if (!lens_last_vertex<T, FIXED>(S, o, scr, ls, prim, dv, ref, nl - 1, k, dir)) return;
            } else {
                T r0 = T(ref.u[k * s]), r1 = T(ref.u[(k + 1) * s]);
                if (!bsdf_sampling<T, true>(S, o, true, true, ls, prim, dv, r0, r1, ls, dir, bc, d)) return;
            }
            k += 2;
            ls.thr = ls.thr * T(ref.lrr[d * vs]);
            prevLensC = ls.lensC;
            r.org = ls.isect.pos;
            r.dir = dir;
        }
    }
    if (c < 2) return;
    const float sxf = float(val(scr[0])), syf = float(val(scr[1]));
    if constexpr (!FIXED) {
        ref.u[k * s] = sxf;
        ref.u[(k + 1) * s] = syf;
        ref.screen[0] = sxf;
        ref.screen[s] = syf;
    }
    k += 2;
    RayT<T> r;
    BState<T> cs;
    emit_from_camera(S, scr[0], scr[1], r, cs);
    bool lcp = false;
    for (int d = 0; d < nc; d++) {
        int prim = ref.cprim[d * vs];
        bool hit = true;
        if (d == 0 && !FIXED) {
            hit = intersect(S, r, cs.isect);
            prim = hit ? cs.isect.prim : -1;
            ref.cprim[0] = prim;
        } else if (prim < 0) {
            hit = false;
        } else if (!re_intersect(S, prim, r, cs.isect)) {
            return;
        }
        cs.wi = -r.dir;
        if (hit) convert_mis(r, cs);
        if (d == nc - 1 && l == 0) {
            if (hit) {
                const int li = S.tlight[prim];
                if (li >= 0) handle_hit_light<T, true>(S, d, li, r, sxf, syf, cs, out);
            }
            return;
        }
        if (!hit) return;
        const float dv = ref.cdisc[d * vs];
        if (d == 1) {
            const T dd = length2(cs.isect.pos - r.org);
            if (val(dd) <= 0) { out.valid = false; return; }
            out.lensPos = lower(cs.isect.pos);
            cs.lensC = cs.lensC * (T(1) / dd);
        }
        if (d == nc - 1) {
            if (l == 1) {
                const T r0 = T(ref.u[k * s]), r1 = T(ref.u[(k + 1) * s]);
                direct_lighting<T, true, true>(S, o, d, cs, sxf, syf, *ref.light, *ref.lprim, r0, r1, prim, out, nullptr, dv);
            } else if (!FIXED && d == 0) {
                connect_vertex<T, false, true>(S, d, nl - 1, ls, lastLightPrim, cs, prim, sxf, syf, out, nullptr, &o, dv,
                                               ref.ldisc[(nl - 1) * vs]);
            } else {
                connect_vertex<T, true, true>(S, d, nl - 1, ls, lastLightPrim, cs, prim, sxf, syf, out, nullptr, &o, dv,
                                              ref.ldisc[(nl - 1) * vs]);
            }
            return;
        }
        if (d == nc - 2 && l == 0) {
            const int np = ref.cprim[(nc - 1) * vs];
            if (np >= 0 && S.tlight[np] >= 0) lcp = true;
        }
        V3<T> dir, bc;
        if (d == 0) {
            if (!lens_first_vertex<T, FIXED>(S, o, lensPos, cs, prim, dv, ref, k, dir)) return;
        } else {
            T r0 = T(ref.u[k * s]), r1 = T(ref.u[(k + 1) * s]);
            if (lcp) {
                if (!light_coordinate_sampling<T, true, true>(S, prim, r0, r1, ref.cprim[(nc - 1) * vs], cs, dir, nullptr, d, &o, dv)) return;
            } else {
                if (!bsdf_sampling<T, true>(S, o, false, true, cs, prim, dv, r0, r1, cs, dir, bc, d)) return;
            }
        }
        k += 2;
        cs.thr = cs.thr * T(ref.crr[d * vs]);
        r.org = cs.isect.pos;
        r.dir = dir;
        r.minT = kIsectEps;
        r.maxT = __builtin_inff();
    }
}

// Gradient and Hessian of log lensScore in the screen position at scr (the
// lens function, FIXED perturb_lens_eval) by two forward-over-forward passes:
// g = (d/ds0, d/ds1), h = (H00, H01, H11). False without a contribution.
DEV bool lens_jets(const DevScene &S, const Opts &o, const SRef &ref, const float scr[2], const V3<float> &lensPos, float g[2], float h[3]) {
    {
        // row x: the gradient and the Hessian's first row
        Jet<float, 2> sj[2];
        for (int k = 0; k < 2; k++) {
            sj[k] = Jet<float, 2>(scr[k]);
            sj[k].t = k == 0 ? 1.0f : 0.0f;
            sj[k].g[0] = k == 0 ? 1.0f : 0.0f;
            sj[k].g[1] = k == 1 ? 1.0f : 0.0f;
        }
        POut<Jet<float, 2>> out;
        perturb_lens_eval<Jet<float, 2>, true>(S, o, ref, sj, lensPos, out);
        if (!out.valid || !(out.lensScore.v > 0.0f)) return false;
        const Jet<float, 2> lf = s_log(out.lensScore);
        g[0] = lf.g[0]; g[1] = lf.g[1];
        h[0] = lf.h[0]; h[1] = lf.h[1];
    }
    {
        // row y, column y: the second diagonal entry (one column carries it;
        // the jet's columns are computed independently of each other)
        Jet<float, 1> sj[2];
        for (int k = 0; k < 2; k++) {
            sj[k] = Jet<float, 1>(scr[k]);
            sj[k].t = k == 1 ? 1.0f : 0.0f;
            sj[k].g[0] = k == 1 ? 1.0f : 0.0f;
        }
        POut<Jet<float, 1>> out;
        perturb_lens_eval<Jet<float, 1>, true>(S, o, ref, sj, lensPos, out);
        if (!out.valid || !(out.lensScore.v > 0.0f)) return false;
        h[2] = s_log(out.lensScore).h[0];
    }
    return true;
}

// The lens score and lensVertexPos of the state behind ref: its
// fixed-topology re-evaluation with the lens bookkeeping (the value the
// state's own evaluation tracked in the reference). 0 without a contribution.
DEV float lens_state_eval(const DevScene &S, const Opts &o, const SRef &ref, V3<float> &lensPos) {
    PlainU<float> up{ref.u, ref.s, 0};
    POut<float> out;
    out.lensPos = V3<float>(0.0f, 0.0f, 0.0f);
    perturb_eval<float, true, PlainU<float>, DiscNone, true>(S, o, ref, up, DiscNone{}, out);
    lensPos = out.lensPos;
    return out.valid ? out.lensScore : 0.0f;
}

// Copy a state between two references (header + D dims + vertex fields).
DEV void copy_state(const SRef &a, const SRef &b) {
    *b.c = *a.c; *b.l = *a.l; *b.light = *a.light; *b.lprim = *a.lprim;
    b.screen[0] = a.screen[0]; b.screen[b.s] = a.screen[a.s];
    const int c = *a.c, l = *a.l;
    const int P = c + l - 1, D = (P > 2 ? P : 2) * 2;
    for (int k = 0; k < D; k++) b.u[k * b.s] = a.u[k * a.s];
    for (int k = 0; k < c - 1; k++) { b.cprim[k * b.vs] = a.cprim[k * a.vs]; b.cdisc[k * b.vs] = a.cdisc[k * a.vs]; b.crr[k * b.vs] = a.crr[k * a.vs]; }
    for (int k = 0; k < l - 1; k++) { b.lgprim[k * b.vs] = a.lgprim[k * a.vs]; b.ldisc[k * b.vs] = a.ldisc[k * a.vs]; b.lrr[k * b.vs] = a.lrr[k * a.vs]; }
}
DEV int dim_of(int c, int l) { const int P = c + l - 1; return (P > 2 ? P : 2) * 2; }  // path.h:132-139, path.cpp:2286-2289

// ------------------------------------------------------------------ generation
// Writes u in GetPathPos order (path.cpp:2393-2439) as the subpath is built.
struct UWriter {
    float *u;
    int s, k;
    DEV void put(float v) { u[k * s] = v; k++; }
};

// path.cpp:1123-1269, uniform order as the CPU restatement: light pos(2),
// dir(2), light pick, triangle pick; per light vertex disc, rnd(2); screen(2);
// per camera vertex disc, then [light pick, rnd(2), triangle pick] or rnd(2).
template <class G> DEV void generate_subpath(const DevScene &S, const Opts &o, int camLen, int lgtLen, const SRef &ref, G &g, POut<float> &out) {
    out.valid = false;
    const int s = ref.s, vs = ref.vs;
    *ref.c = camLen; *ref.l = lgtLen;
    *ref.light = -1; *ref.lprim = -1;
    for (int k = 0; k < kVMax; k++) {
        ref.cprim[k * vs] = -1; ref.lgprim[k * vs] = -1; ref.crr[k * vs] = 1.0f; ref.lrr[k * vs] = 1.0f;
        ref.cdisc[k * vs] = 0.0f; ref.ldisc[k * vs] = 0.0f;
    }
    for (int k = 0; k < kDMax; k++) ref.u[k * s] = 0.0f;
    // the camera subpath needs two vertices unless a light subpath of two or
    // more ends in the camera connection (the reference asserts camLength > 1
    // there, path.cpp:1186); anything else is no path
    if (camLen < 2 && lgtLen < 2) return;
    UWriter uw{ref.u, s, 0};
    RayT<float> r;
    BState<float> ls;
    int lastLightPrim = -1;
    if (lgtLen > 1) {
        float p0 = g.uni(), p1 = g.uni(), d0 = g.uni(), d1 = g.uni();
        const int li = pick_light(S, g.uni());
        const int lp = pick_light_tri(S, li, g.uni());
        *ref.light = li; *ref.lprim = lp;
        uw.put(p0); uw.put(p1); uw.put(d0); uw.put(d1);
        emit_from_light(S, o, li, lp, p0, p1, d0, d1, r, ls);
        r.minT = kIsectEps;
        r.maxT = __builtin_inff();
        for (int d = 0;; d++) {
            if (!intersect(S, r, ls.isect)) return;
            const int prim = ls.isect.prim;
            ref.lgprim[d * vs] = prim;
            const float dv = g.uni();
            ref.ldisc[d * vs] = dv;
            ls.wi = -r.dir;
            convert_mis(r, ls);
            lastLightPrim = prim;
            if (d + 2 == lgtLen) {
                if (camLen == 1) {
                    connect_to_camera<float, false>(S, o, d, ls, prim, dv, r.org, out);
                    if (out.valid) { ref.screen[0] = out.screen[0]; ref.screen[s] = out.screen[1]; }
                    return;
                }
                break;
            }
            float r0 = g.uni(), r1 = g.uni();
            V3<float> dir, bc;
            if (!bsdf_sampling(S, o, true, false, ls, prim, dv, r0, r1, ls, dir, bc)) return;
            uw.put(r0); uw.put(r1);
            r.org = ls.isect.pos;
            r.dir = dir;
        }
    }
    BState<float> cs;
    const float sx = g.uni(), sy = g.uni();
    uw.put(sx); uw.put(sy);
    ref.screen[0] = sx; ref.screen[s] = sy;
    emit_from_camera(S, sx, sy, r, cs);
    for (int d = 0;; d++) {
        const bool hit = intersect(S, r, cs.isect);
        const int prim = hit ? cs.isect.prim : -1;
        if (hit) ref.cprim[d * vs] = prim;
        cs.wi = -r.dir;
        if (hit) convert_mis(r, cs);
        if (d + 2 >= camLen && lgtLen == 0) {
            if (hit && S.tlight[prim] >= 0) {
                const int li = S.tlight[prim];
                if (d > 1) {
                    float u0, u1;
                    get_sample_param(S, prim, cs.isect.pos, u0, u1);
                    ref.u[(uw.k - 2) * s] = u0;
                    ref.u[(uw.k - 1) * s] = u1;
                    V3<float> toPrev = cs.isect.pos - r.org;
                    const float d2 = length2(toPrev), id2 = 1.0f / d2, id = sqrtf(id2);
                    toPrev = toPrev * id;
                    cs.ssJ *= fabsf(dot(toPrev, cs.isect.sn) * id2) * (cs.lcJ * (1.0f / S.lightArea[li]));
                }
                handle_hit_light(S, d, li, r, sx, sy, cs, out);
            }
            return;
        }
        if (!hit) return;
        if (d == 1 && length2(cs.isect.pos - r.org) <= 0.0f) return;
        const float dv = g.uni();
        ref.cdisc[d * vs] = dv;
        if (d + 2 == camLen) {
            if (lgtLen == 1) {
                const int li = pick_light(S, g.uni());
                const float r0 = g.uni(), r1 = g.uni();
                const int lp = pick_light_tri(S, li, g.uni());
                *ref.light = li; *ref.lprim = lp;
                uw.put(r0); uw.put(r1);
                direct_lighting<float, false>(S, o, d, cs, sx, sy, li, lp, r0, r1, prim, out);
            } else {
                connect_vertex<float, false>(S, d, lgtLen - 2, ls, lastLightPrim, cs, prim, sx, sy, out);
            }
            return;
        }
        float r0 = g.uni(), r1 = g.uni();
        V3<float> dir, bc;
        if (!bsdf_sampling(S, o, false, false, cs, prim, dv, r0, r1, cs, dir, bc)) return;
        uw.put(r0); uw.put(r1);
        r.org = cs.isect.pos;
        r.dir = dir;
        r.minT = kIsectEps;
        r.maxT = __builtin_inff();
    }
}

// path.cpp:400-415: one uniform per call
template <class G> DEV bool russian_roulette(int depth, const V3<float> &bc, float &rrW, V3<float> &thr, G &g) {
    float p = 1.0f;
    if (depth >= 3) p = fminf(fmaxf(bc.x, fmaxf(bc.y, bc.z)), 0.95f);
    if (g.uni() > p) return false;
    rrW = 1.0f / p;
    thr = thr * rrW;
    return true;
}

// One contribution of GeneratePathBidir: strategy, scores and where to find it.
struct BidirContrib {
    int c, l;
    float rgb[3], lsScore, ssScore, sx, sy, mis;
};

// path.cpp:978-1121 (all strategies). Calls sink(contrib) per contribution.
// When `keep` names a strategy (c, l) > 0, the truncated (c, l) subpath state
// is written to `ref` (the pre-pass seed replay, delayed.cpp:125-167). The
// light-subpath states after the first are value-initialised in the reference,
// so a relative-parameterised light vertex leaves ssJacobian == 0.
template <int MAXL, class G, class SINK>
DEV void generate_path_bidir(const DevScene &S, const Opts &o, int minDepth, int maxDepth, G &g, SINK &sink, int keepC, int keepL,
                             const SRef *ref) {
    int keepDone = 0;
    BState<float> lsv[MAXL + 1];
    int lprims[MAXL + 1];
    float ldiscs[MAXL + 1], lrrs[MAXL + 1], lrnd[2 * (MAXL + 1)];
    float lp0 = g.uni(), lp1 = g.uni(), ld0 = g.uni(), ld1 = g.uni();
    const int li = pick_light(S, g.uni());
    const int ltri = pick_light_tri(S, li, g.uni());
    RayT<float> r;
    emit_from_light(S, o, li, ltri, lp0, lp1, ld0, ld1, r, lsv[0]);
    r.minT = kIsectEps;
    r.maxT = __builtin_inff();
    int nls = 0;  // light states with an intersection
    for (int d = 0; d <= MAXL; d++) {
        if (!intersect(S, r, lsv[d].isect)) break;
        nls = d + 1;
        lprims[d] = lsv[d].isect.prim;
        ldiscs[d] = g.uni();
        lrrs[d] = 1.0f;
        lsv[d].wi = -r.dir;
        convert_mis(r, lsv[d]);
        if (d + 2 >= minDepth) {
            POut<float> out;
            out.valid = false;
            connect_to_camera<float, false>(S, o, d, lsv[d], lprims[d], ldiscs[d], r.org, out);
            if (out.valid) {
                sink(out, 1, 2 + d);
                if (keepC == 1 && keepL == 2 + d) keepDone = 4;
            }
        }
        if (keepDone) break;
        if (maxDepth != -1 && d + 2 >= maxDepth) break;
        if (d + 1 > MAXL) break;
        BState<float> nx;
        nx.ssJ = 0.0f;
        nx.thr = V3<float>(0, 0, 0);
        lrnd[2 * d] = g.uni(); lrnd[2 * d + 1] = g.uni();
        V3<float> dir, bc;
        if (!bsdf_sampling(S, o, true, false, lsv[d], lprims[d], ldiscs[d], lrnd[2 * d], lrnd[2 * d + 1], nx, dir, bc)) break;
        if (!russian_roulette(d, bc, lrrs[d], nx.thr, g)) break;
        lsv[d + 1] = nx;
        r.org = lsv[d].isect.pos;
        r.dir = dir;
    }
    BState<float> cs;
    const float sx = g.uni(), sy = g.uni();
    emit_from_camera(S, sx, sy, r, cs);
    int cprims[kVMax];
    float cdiscs[kVMax], crrs[kVMax], crnd[2 * kVMax];
    int dlLight = -1, dlPrim = -1;
    float dlR0 = 0, dlR1 = 0;
    for (int d = 0; d < kVMax && !keepDone; d++) {
        const bool hit = intersect(S, r, cs.isect);
        cprims[d] = hit ? cs.isect.prim : -1;
        crrs[d] = 1.0f;
        cdiscs[d] = 0.0f;
        crnd[2 * d] = 0.0f;
        crnd[2 * d + 1] = 0.0f;
        cs.wi = -r.dir;
        if (hit) convert_mis(r, cs);
        if (d + 1 >= minDepth) {
            const int lli = hit ? S.tlight[cprims[d]] : -1;
            if (lli >= 0) {
                if (o.lightCoord && d > 1) {
                    get_sample_param(S, cprims[d], cs.isect.pos, crnd[2 * (d - 1)], crnd[2 * (d - 1) + 1]);
                    V3<float> toPrev = cs.isect.pos - r.org;
                    const float d2 = length2(toPrev), id2 = 1.0f / d2, id = sqrtf(id2);
                    toPrev = toPrev * id;
                    cs.ssJ *= fabsf(dot(toPrev, cs.isect.sn) * id2) * (cs.lcJ * (1.0f / S.lightArea[lli]));
                }
                POut<float> out;
                out.valid = false;
                handle_hit_light(S, d, lli, r, sx, sy, cs, out);
                if (out.valid) {
                    sink(out, 2 + d, 0);
                    if (keepC == 2 + d && keepL == 0) keepDone = 1;
                }
                break;
            }
        }
        if (!hit || (maxDepth != -1 && d + 1 >= maxDepth)) break;
        if (d == 1 && length2(cs.isect.pos - r.org) <= 0.0f) break;
        cdiscs[d] = g.uni();
        if (d + 2 >= minDepth) {
            const int dli = pick_light(S, g.uni());
            const float r0 = g.uni(), r1 = g.uni();
            const int dlp = pick_light_tri(S, dli, g.uni());
            POut<float> out;
            out.valid = false;
            direct_lighting<float, false>(S, o, d, cs, sx, sy, dli, dlp, r0, r1, cprims[d], out);
            if (out.valid) {
                sink(out, 2 + d, 1);
                if (keepC == 2 + d && keepL == 1) { keepDone = 2; dlLight = dli; dlPrim = dlp; dlR0 = r0; dlR1 = r1; }
            }
        }
        int maxL = maxDepth == -1 ? nls - 1 : min(maxDepth - d - 3, nls - 1);
        for (int ld = 0; ld <= maxL; ld++) {
            if (d + ld + 3 >= minDepth) {
                POut<float> out;
                out.valid = false;
                connect_vertex<float, false>(S, d, ld, lsv[ld], lprims[ld], cs, cprims[d], sx, sy, out);
                if (out.valid) {
                    sink(out, 2 + d, 2 + ld);
                    if (keepC == 2 + d && keepL == 2 + ld) keepDone = 3;
                }
            }
        }
        if (keepDone) break;
        crnd[2 * d] = g.uni(); crnd[2 * d + 1] = g.uni();
        V3<float> dir, bc;
        if (!bsdf_sampling(S, o, false, false, cs, cprims[d], cdiscs[d], crnd[2 * d], crnd[2 * d + 1], cs, dir, bc)) break;
        if (!russian_roulette(d, bc, crrs[d], cs.thr, g)) break;
        r.org = cs.isect.pos;
        r.dir = dir;
        r.minT = kIsectEps;
        r.maxT = __builtin_inff();
    }
    if (ref) *ref->c = 0;  // marks "strategy not found" unless written below
    if (ref && keepDone) {
        const int s = ref->s, vs = ref->vs;
        const int c = keepC, l = keepL;
        *ref->c = c; *ref->l = l;
        *ref->light = -1; *ref->lprim = -1;
        for (int k = 0; k < kVMax; k++) {
            ref->cprim[k * vs] = k < c - 1 ? cprims[k] : -1;
            ref->cdisc[k * vs] = k < c - 1 ? cdiscs[k] : 0.0f;
            ref->crr[k * vs] = k < c - 1 ? crrs[k] : 1.0f;
            ref->lgprim[k * vs] = k < l - 1 ? lprims[k] : -1;
            ref->ldisc[k * vs] = k < l - 1 ? ldiscs[k] : 0.0f;
            ref->lrr[k * vs] = k < l - 1 ? lrrs[k] : 1.0f;
        }
        for (int k = 0; k < kDMax; k++) ref->u[k * s] = 0.0f;
        UWriter uw{ref->u, s, 0};
        if (l > 1) {
            *ref->light = li; *ref->lprim = ltri;
            uw.put(lp0); uw.put(lp1); uw.put(ld0); uw.put(ld1);
            for (int k = 0; k < l - 2; k++) { uw.put(lrnd[2 * k]); uw.put(lrnd[2 * k + 1]); }
        }
        if (c > 1) {
            uw.put(sx); uw.put(sy);
            for (int k = 0; k < c - 2; k++) { uw.put(crnd[2 * k]); uw.put(crnd[2 * k + 1]); }
            if (l == 1) { *ref->light = dlLight; *ref->lprim = dlPrim; uw.put(dlR0); uw.put(dlR1); }
        }
    }
}

// ------------------------------------------------------------- direct pass
// path.cpp:124-268, 270-338, 343-431 (unidirectional, path length <= 2)
template <class G> DEV void generate_path_uni_direct(const DevScene &S, const Opts &o, int px, int py, G &g, float *film) {
    const float sx = (float(px) + g.uni()) / float(S.cam.W);
    const float sy = (float(py) + g.uni()) / float(S.cam.H);
    RayT<float> r;
    sample_primary(S.cam, sx, sy, r);
    V3<float> thr(1, 1, 1);
    float lastPdf = 1.0f;
    auto splat = [&](const V3<float> &c) {
        int ix = min(max(int(sx * S.cam.W), 0), S.cam.W - 1), iy = min(max(int(sy * S.cam.H), 0), S.cam.H - 1);
        if (isfinite(c.x) && isfinite(c.y) && isfinite(c.z)) {
            float *p = film + 3 * (size_t(iy) * S.cam.W + ix);
            atomicAdd(p, c.x); atomicAdd(p + 1, c.y); atomicAdd(p + 2, c.z);
        }
    };
    for (int d = 0; d < 2; d++) {
        Isect<float> is;
        const bool hit = intersect(S, r, is);
        const int li = hit ? S.tlight[is.prim] : -1;
        if (li >= 0) {
            V3<float> em;
            float dPdf, emPdf;
            light_emission(S, li, r.dir, is.sn, em, dPdf, emPdf);
            if (em.x + em.y + em.z > 0.0f) {
                const float d2 = length2(r.org - is.pos), cosT = -dot(r.dir, is.sn);
                dPdf *= (d2 / cosT);
                V3<float> c = mul(thr, em);
                if (d > 0) {
                    const float ratio = (dPdf * pick_light_prob(S, li)) / lastPdf;
                    c = c * (1.0f / (1.0f + ratio * ratio));
                }
                if (luminance(c) > 0.0f) splat(c);
            }
            return;
        }
        if (!hit || d + 1 >= 2) return;
        const float dv = g.uni();
        const Mat m = mat_of(S, is.prim);
        const V3<float> wi = -r.dir;
        {
            const int dli = pick_light(S, g.uni());
            const float pick = pick_light_prob(S, dli);
            const float r0 = g.uni(), r1 = g.uni();
            const int dlp = pick_light_tri(S, dli, g.uni());
            V3<float> toL, lc;
            float dist, cosL, dPdf, emPdf;
            if (light_sample_direct(S, dli, dlp, is.pos, r0, r1, toL, dist, lc, cosL, dPdf, emPdf) && !occluded(S, is.pos, toL, dist)) {
                BOut<float> b;
                bsdf_eval(m, false, wi, is.sn, toL, b);
                if (!is_zero(b.contrib)) {
                    V3<float> c = mul(mul(thr, b.contrib), lc) * (1.0f / pick);
                    const float ratio = b.pdf / (dPdf * pick);
                    c = c * (1.0f / (1.0f + ratio * ratio));
                    if (luminance(c) > 0.0f) splat(c);
                }
            }
        }
        float r0 = g.uni(), r1 = g.uni();
        BOut<float> b;
        V3<float> dir;
        if (!bsdf_sample(m, false, wi, is.sn, r0, r1, dv, dir, b)) return;
        lastPdf = b.pdf;
        thr = mul(thr, b.contrib);
        (void)g.uni();  // Russian roulette at depth < 3 always continues (path.cpp:400-415)
        r.org = is.pos;
        r.dir = dir;
        r.minT = kIsectEps;
        r.maxT = __builtin_inff();
    }
}

}  // namespace drmlt

// ORACLE (test infrastructure only). C entry points over the CPU restatement
// in orc_*.h, loaded with ctypes by tests/, by bench.py's cpu_baseline leg and
// by the smoke check. The product never links or loads this library.
#include "orc_dr.h"
#include <atomic>
#include <chrono>
#include <mutex>
#include <thread>

using namespace orc;

// Flat path record shared with the product's parity surface (same layout as
// drmlt_path_rec in include/drmlt_amd.h).
struct PathRec {
    int32_t c, l;
    int32_t light, lprim;
    float screen[2];
    float u[32];
    int32_t cprim[16], lgprim[16];
    float cdisc[16], ldisc[16];
    float crr[16], lrr[16];
};
static_assert(sizeof(PathRec) == 536, "record layout");

struct OrcScene {
    SceneT<float> f;
    SceneT<double> d;
};

template <class F> static PathS<F> FromRec(const PathRec &r) {
    PathS<F> p;
    p.c = r.c; p.l = r.l;
    p.cam.resize(std::max(r.c - 1, 0));
    p.lgt.resize(std::max(r.l - 1, 0));
    for (int i = 0; i < (int)p.cam.size(); i++) { p.cam[i].prim = r.cprim[i]; p.cam[i].disc = r.cdisc[i]; p.cam[i].rr = r.crr[i]; }
    for (int i = 0; i < (int)p.lgt.size(); i++) { p.lgt[i].prim = r.lgprim[i]; p.lgt[i].disc = r.ldisc[i]; p.lgt[i].rr = r.lrr[i]; }
    if (r.l >= 2) { p.lLight = r.light; p.lPrim = r.lprim; }
    if (r.l == 1 && !p.cam.empty()) { p.cam.back().dlLight = r.light; p.cam.back().dlPrim = r.lprim; }
    F u[32];
    for (int i = 0; i < 32; i++) u[i] = F(r.u[i]);
    SetPathPos(p, u);
    if (r.c == 1) { p.screen[0] = r.screen[0]; p.screen[1] = r.screen[1]; }
    return p;
}
template <class F> static void ToRec(const PathS<F> &p, PathRec &r) {
    std::memset(&r, 0, sizeof(r));
    r.c = p.c; r.l = p.l;
    for (int i = 0; i < 16; i++) { r.cprim[i] = -1; r.lgprim[i] = -1; r.crr[i] = 1; r.lrr[i] = 1; }
    for (int i = 0; i < (int)p.cam.size() && i < 16; i++) { r.cprim[i] = p.cam[i].prim; r.cdisc[i] = float(p.cam[i].disc); r.crr[i] = float(p.cam[i].rr); }
    for (int i = 0; i < (int)p.lgt.size() && i < 16; i++) { r.lgprim[i] = p.lgt[i].prim; r.ldisc[i] = float(p.lgt[i].disc); r.lrr[i] = float(p.lgt[i].rr); }
    r.light = -1; r.lprim = -1;
    if (p.l >= 2) { r.light = p.lLight; r.lprim = p.lPrim; }
    if (p.l == 1 && !p.cam.empty()) { r.light = p.cam.back().dlLight; r.lprim = p.cam.back().dlPrim; }
    F u[DMAX] = {0};
    GetPathPos(p, u);
    const int D = Dim<F>(p.c, p.l);
    for (int i = 0; i < D && i < 32; i++) r.u[i] = float(u[i]);
    r.screen[0] = float(p.screen[0]);
    r.screen[1] = float(p.screen[1]);
}
template <class F> static void ContribOut(const std::vector<Contrib<F>> &out, float *c8) {
    std::memset(c8, 0, 8 * sizeof(float));
    if (out.empty()) return;
    const Contrib<F> &k = out[0];
    c8[0] = 1; c8[1] = float(k.contrib.x); c8[2] = float(k.contrib.y); c8[3] = float(k.contrib.z);
    c8[4] = float(k.lsScore); c8[5] = float(k.ssScore); c8[6] = float(k.screen[0]); c8[7] = float(k.screen[1]);
}
static Material<double> MatD(const float *m) {
    Material<double> M;
    M.type = int(m[0]);
    M.a = V3<double>(m[1], m[2], m[3]);
    M.b = V3<double>(m[4], m[5], m[6]);
    M.expo = m[7]; M.alpha = m[8]; M.eta = m[9]; M.invEta = m[10]; M.ksWeight = m[11]; M.twoSided = m[12] != 0;
    return M;
}

static void ParallelFor(int n, int nthreads, const std::function<void(int, int)> &fn) {
    if (nthreads <= 1) { for (int i = 0; i < n; i++) fn(i, 0); return; }
    std::atomic<int> next(0);
    std::vector<std::thread> th;
    for (int t = 0; t < nthreads; t++)
        th.emplace_back([&, t]() {
            for (;;) {
                int i = next.fetch_add(1);
                if (i >= n) break;
                fn(i, t);
            }
        });
    for (auto &x : th) x.join();
}

extern "C" int orc_path_rec_size() { return int(sizeof(PathRec)); }

extern "C" void *orc_scene_new(int nt, const float *triP, const float *triN, const int *triMat, const int *triLight, int nm, const float *mats,
                    int nl, const int *lightTris, const float *lightEmit, const float *cam, int useBvh) {
    OrcScene *s = new OrcScene();
    LoadScene(s->f, nt, triP, triN, triMat, triLight, nm, mats, nl, lightTris, lightEmit, cam);
    LoadScene(s->d, nt, triP, triN, triMat, triLight, nm, mats, nl, lightTris, lightEmit, cam);
    s->f.useBvh = s->d.useBvh = useBvh != 0;
    return s;
}
extern "C" void orc_scene_free(void *h) { delete (OrcScene *)h; }

extern "C" void orc_trace(void *h, int n, const float *org, const float *dir, const float *tmin, const float *tmax, int *prim, float *t) {
    OrcScene *s = (OrcScene *)h;
    for (int i = 0; i < n; i++) prim[i] = TraceClosest(s->f, org + 3 * i, dir + 3 * i, tmin[i], tmax[i], t[i]);
}
extern "C" void orc_occluded(void *h, int n, const float *org, const float *dir, const float *tmin, const float *tmax, int *occ) {
    OrcScene *s = (OrcScene *)h;
    for (int i = 0; i < n; i++) occ[i] = TraceAny(s->f, org + 3 * i, dir + 3 * i, tmin[i], tmax[i]) ? 1 : 0;
}

extern "C" void orc_bsdf_eval(const float *mat, int adjoint, const double *wi, const double *nn, const double *wo, double *out6) {
    Material<double> M = MatD(mat);
    BsdfOut<double> o;
    BsdfEval(M, adjoint != 0, V3<double>(wi[0], wi[1], wi[2]), V3<double>(nn[0], nn[1], nn[2]), V3<double>(wo[0], wo[1], wo[2]), o);
    out6[0] = o.contrib.x; out6[1] = o.contrib.y; out6[2] = o.contrib.z;
    out6[3] = o.cosWo; out6[4] = o.pdf; out6[5] = o.revPdf;
}
extern "C" int orc_bsdf_sample(const float *mat, int adjoint, const double *wi, const double *nn, double r0, double r1, double disc, double *wo3, double *out6) {
    Material<double> M = MatD(mat);
    BsdfOut<double> o;
    V3<double> wo;
    bool ok = BsdfSample(M, adjoint != 0, V3<double>(wi[0], wi[1], wi[2]), V3<double>(nn[0], nn[1], nn[2]), r0, r1, disc, wo, o);
    wo3[0] = wo.x; wo3[1] = wo.y; wo3[2] = wo.z;
    out6[0] = o.contrib.x; out6[1] = o.contrib.y; out6[2] = o.contrib.z;
    out6[3] = o.cosWo; out6[4] = o.pdf; out6[5] = o.revPdf;
    return ok ? 1 : 0;
}
extern "C" double orc_roughness(const float *mat, double disc) { return Roughness(MatD(mat), disc); }
// orc_bsdf_eval over n direction pairs (wi, wo: n x 3; one normal) into n x 6.
extern "C" void orc_bsdf_eval_batch(const float *mat, int adjoint, int n, const double *wi, const double *nn, const double *wo, double *out6) {
    const Material<double> M = MatD(mat);
    const V3<double> N(nn[0], nn[1], nn[2]);
    for (int i = 0; i < n; i++) {
        BsdfOut<double> o;
        BsdfEval(M, adjoint != 0, V3<double>(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]), N, V3<double>(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), o);
        double *q = out6 + 6 * size_t(i);
        q[0] = o.contrib.x; q[1] = o.contrib.y; q[2] = o.contrib.z;
        q[3] = o.cosWo; q[4] = o.pdf; q[5] = o.revPdf;
    }
}
// Plants mistake `which` (orc_core.h, g_plant: the BSDF code; 20 and 21 the
// lens bookkeeping of orc_path.h); 0 removes it.
extern "C" void orc_plant(int which) { g_plant = which; }

template <class F>
static void PerturbT(const SceneT<F> &s, int n, const PathRec *in, const float *off, const float *disc, int occl, int fixed, int chadQuirk,
                     PathRec *out, float *c8) {
    Opts<F> o;
    o.chadMisQuirk = chadQuirk != 0;
    for (int i = 0; i < n; i++) {
        PathS<F> p = FromRec<F>(in[i]);
        F of[DMAX] = {0}, di[2 * VMAX] = {0};
        for (int k = 0; k < 32; k++) of[k] = F(off[32 * i + k]);
        for (int k = 0; k < 2 * VMAX; k++) di[k] = F(disc[32 * i + k]);
        std::vector<Contrib<F>> res;
        PerturbPathBidir(s, o, of, di, p, res, occl != 0, fixed != 0);
        ToRec(p, out[i]);
        ContribOut(res, c8 + 8 * i);
        if (!res.empty()) { out[i].screen[0] = c8[8 * i + 6]; out[i].screen[1] = c8[8 * i + 7]; }
    }
}
extern "C" void orc_perturb(void *h, int dbl, int n, const PathRec *in, const float *off, const float *disc, int occl, int fixed, int chadQuirk,
                 PathRec *out, float *c8) {
    OrcScene *s = (OrcScene *)h;
    if (dbl) PerturbT(s->d, n, in, off, disc, occl, fixed, chadQuirk, out, c8);
    else PerturbT(s->f, n, in, off, disc, occl, fixed, chadQuirk, out, c8);
}

template <class F>
static void GenSubT(const SceneT<F> &s, int n, const int *cl, const float *uni, int nuni, PathRec *out, float *c8) {
    Opts<F> o;
    for (int i = 0; i < n; i++) {
        ArraySrc<F> g(uni + size_t(i) * nuni, nuni);
        PathS<F> p;
        std::vector<Contrib<F>> res;
        GenerateSubpath(s, o, cl[2 * i], cl[2 * i + 1], p, res, g);
        Truncate(p, cl[2 * i], cl[2 * i + 1]);
        ToRec(p, out[i]);
        ContribOut(res, c8 + 8 * i);
        if (!res.empty()) { out[i].screen[0] = c8[8 * i + 6]; out[i].screen[1] = c8[8 * i + 7]; }
    }
}
extern "C" void orc_generate_subpath(void *h, int dbl, int n, const int *cl, const float *uni, int nuni, PathRec *out, float *c8) {
    OrcScene *s = (OrcScene *)h;
    if (dbl) GenSubT(s->d, n, cl, uni, nuni, out, c8);
    else GenSubT(s->f, n, cl, uni, nuni, out, c8);
}

// The lens score and lensVertexPos of a state from its fixed-topology
// re-evaluation (the state's own triangles, unperturbed, no occlusion).
template <class F> static F LensPrepare(const SceneT<F> &s, const Opts<F> &o, PathS<F> &p) {
    PathS<F> q = p;
    F off[DMAX] = {0}, di[2 * VMAX] = {0};
    std::vector<Contrib<F>> res;
    PerturbPathBidir(s, o, off, di, q, res, false, true);
    p.lensPos = q.lensPos;
    return res.empty() ? F(0) : res[0].lensScore;
}
// Per state: lensScore, lensVertexPos (4 floats) of the fixed re-evaluation.
extern "C" void orc_lens_score(void *h, int dbl, int n, const PathRec *in, float *out4) {
    OrcScene *s = (OrcScene *)h;
    for (int i = 0; i < n; i++) {
        if (dbl) {
            PathS<double> p = FromRec<double>(in[i]);
            out4[4 * i] = float(LensPrepare(s->d, Opts<double>(), p));
            out4[4 * i + 1] = float(p.lensPos.x); out4[4 * i + 2] = float(p.lensPos.y); out4[4 * i + 3] = float(p.lensPos.z);
        } else {
            PathS<float> p = FromRec<float>(in[i]);
            out4[4 * i] = LensPrepare(s->f, Opts<float>(), p);
            out4[4 * i + 1] = p.lensPos.x; out4[4 * i + 2] = p.lensPos.y; out4[4 * i + 3] = p.lensPos.z;
        }
    }
}
// GenerateSubpath's tracked lensScore per sample (0 without a contribution).
extern "C" void orc_generate_subpath_lens(void *h, int dbl, int n, const int *cl, const float *uni, int nuni, float *lens) {
    OrcScene *s = (OrcScene *)h;
    for (int i = 0; i < n; i++) {
        std::vector<Contrib<double>> rd;
        std::vector<Contrib<float>> rf;
        if (dbl) {
            ArraySrc<double> g(uni + size_t(i) * nuni, nuni);
            PathS<double> p;
            GenerateSubpath(s->d, Opts<double>(), cl[2 * i], cl[2 * i + 1], p, rd, g);
            lens[i] = rd.empty() ? 0.0f : float(rd[0].lensScore);
        } else {
            ArraySrc<float> g(uni + size_t(i) * nuni, nuni);
            PathS<float> p;
            GenerateSubpath(s->f, Opts<float>(), cl[2 * i], cl[2 * i + 1], p, rf, g);
            lens[i] = rf.empty() ? 0.0f : rf[0].lensScore;
        }
    }
}
template <class F>
static void LensT(const SceneT<F> &s, int n, const PathRec *in, const float *screen, int occl, int fixed, PathRec *out, float *c9) {
    Opts<F> o;
    for (int i = 0; i < n; i++) {
        PathS<F> p = FromRec<F>(in[i]);
        LensPrepare(s, o, p);
        const F sc[2] = {F(screen[2 * i]), F(screen[2 * i + 1])};
        std::vector<Contrib<F>> res;
        PerturbLensBidir(s, o, sc, p, res, occl != 0, fixed != 0);
        ToRec(p, out[i]);
        ContribOut(res, c9 + 9 * i);
        c9[9 * i + 8] = res.empty() ? 0.0f : float(res[0].lensScore);
        if (!res.empty()) { out[i].screen[0] = c9[9 * i + 6]; out[i].screen[1] = c9[9 * i + 7]; }
    }
}
// PerturbLensBidir (path.cpp:1914-2067) of each state to screen position
// screen[2i..2i+1]: the perturbed state, contribution (8 floats) and lensScore.
extern "C" void orc_perturb_lens(void *h, int dbl, int n, const PathRec *in, const float *screen, int occl, int fixed, PathRec *out, float *c9) {
    OrcScene *s = (OrcScene *)h;
    if (dbl) LensT(s->d, n, in, screen, occl, fixed, out, c9);
    else LensT(s->f, n, in, screen, occl, fixed, out, c9);
}
// Central differences (step hstep) of the lens function (LogLensFixed) at each
// state's own screen position (the contribution's): log lensScore, gradient 2,
// Hessian 2 x 2.
extern "C" void orc_lens_grad_hess_fd(void *h, int n, const PathRec *in, const float *screen, double hstep, double *logl, double *grad,
                                      double *hess) {
    OrcScene *s = (OrcScene *)h;
    Opts<double> o;
    for (int i = 0; i < n; i++) {
        PathS<double> p = FromRec<double>(in[i]);
        LensPrepare(s->d, o, p);
        const double s0[2] = {screen[2 * i], screen[2 * i + 1]};
        logl[i] = LogLensFixed(s->d, o, p, s0);
        LensGradHessFD(s->d, o, p, s0, hstep, grad + 2 * i, hess + 4 * i);
    }
}

// The h2mc lens mutation's acceptance (LensMutateAt, h2mc.cpp:463-530) in
// double for given screen offsets off[2i..2i+1]: the current state is the
// record with its contribution row c8 (ssScore, screen position) and the lens
// score of its fixed re-evaluation; lens Gaussians with sigma lensSigma from
// central differences of step fdStep, derivatives up to path length
// maxDerv. a[i] = 0 where the perturbed path has no contribution.
extern "C" void orc_lens_accept(void *h, int n, const PathRec *in, const float *c8, const float *off, double lensSigma, double fdStep,
                                int maxDerv, double *a) {
    OrcScene *s = (OrcScene *)h;
    DrChainCtx<double> cx;
    cx.s = &s->d;
    cx.sd = &s->d;
    cx.dro.lensStdDev = lensSigma;
    cx.dro.fdStep = fdStep;
    cx.dro.maxDervDepth = maxDerv;
    cx.b = 1.0;
    for (int i = 0; i < n; i++) {
        MState<double> cur, prop;
        cur.valid = true;
        cur.path = FromRec<double>(in[i]);
        const float *r = c8 + 8 * size_t(i);
        cur.sp.c = in[i].c;
        cur.sp.l = in[i].l;
        cur.sp.contrib = V3<double>(r[1], r[2], r[3]);
        cur.sp.lsScore = r[4];
        cur.sp.ssScore = r[5];
        cur.sp.screen[0] = r[6];
        cur.sp.screen[1] = r[7];
        cur.sp.lensScore = LensPrepare(s->d, cx.o, cur.path);
        const double o2[2] = {double(off[2 * i]), double(off[2 * i + 1])};
        a[i] = LensMutateAt(cx, cur, prop, o2);
    }
}

// Per sample up to maxc contributions, each 10 floats: c, l, r, g, b, lsScore, ssScore, sx, sy, mis.
template <class F>
static void GenBidirT(const SceneT<F> &s, int n, int minDepth, int maxDepth, const float *uni, int nuni, int maxc, float *cs, int *nc) {
    Opts<F> o;
    for (int i = 0; i < n; i++) {
        ArraySrc<F> g(uni + size_t(i) * nuni, nuni);
        PathS<F> p;
        std::vector<Contrib<F>> res;
        GeneratePathBidir(s, o, minDepth, maxDepth, p, res, g);
        nc[i] = int(res.size());
        for (int k = 0; k < (int)res.size() && k < maxc; k++) {
            float *q = cs + (size_t(i) * maxc + k) * 10;
            q[0] = float(res[k].c); q[1] = float(res[k].l);
            q[2] = float(res[k].contrib.x); q[3] = float(res[k].contrib.y); q[4] = float(res[k].contrib.z);
            q[5] = float(res[k].lsScore); q[6] = float(res[k].ssScore); q[7] = float(res[k].screen[0]); q[8] = float(res[k].screen[1]);
            q[9] = float(res[k].mis);
        }
    }
}
extern "C" void orc_generate_bidir(void *h, int dbl, int n, int minDepth, int maxDepth, const float *uni, int nuni, int maxc, float *cs, int *nc) {
    OrcScene *s = (OrcScene *)h;
    if (dbl) GenBidirT(s->d, n, minDepth, maxDepth, uni, nuni, maxc, cs, nc);
    else GenBidirT(s->f, n, minDepth, maxDepth, uni, nuni, maxc, cs, nc);
}

extern "C" void orc_grad_hess_fd(void *h, int n, const PathRec *in, double hstep, int chadQuirk, double *logf, double *grad, double *hess) {
    OrcScene *s = (OrcScene *)h;
    Opts<double> o;
    o.chadMisQuirk = chadQuirk != 0;
    for (int i = 0; i < n; i++) {
        PathS<double> p = FromRec<double>(in[i]);
        const int D = Dim<double>(p.c, p.l);
        double zero[DMAX] = {0};
        logf[i] = LogFFixed(s->d, o, p, zero);
        std::vector<double> g(D), H(D * D);
        GradHessFD(s->d, o, p, D, hstep, g.data(), H.data());
        for (int a = 0; a < 32; a++) grad[32 * i + a] = a < D ? g[a] : 0.0;
        for (int a = 0; a < 32 * 32; a++) hess[1024 * size_t(i) + a] = 0.0;
        for (int a = 0; a < D; a++)
            for (int b2 = 0; b2 < D; b2++) hess[1024 * size_t(i) + a * 32 + b2] = H[a * D + b2];
    }
}
extern "C" void orc_logf_fixed(void *h, int n, const PathRec *in, const double *off, int chadQuirk, double *logf) {
    OrcScene *s = (OrcScene *)h;
    Opts<double> o;
    o.chadMisQuirk = chadQuirk != 0;
    for (int i = 0; i < n; i++) {
        PathS<double> p = FromRec<double>(in[i]);
        logf[i] = LogFFixed(s->d, o, p, off + 32 * size_t(i));
    }
}

// grad: n x 32, hess: n x 32 x 32 (row-major, leading dimension 32)
extern "C" void orc_compute_gaussian(int n, const int *Ds, const double *sc, double sigma, const double *grad, const double *hess, double *mean,
                          double *covL, double *invCov, double *logDet) {
    H2Param pr(sigma);
    for (int i = 0; i < n; i++) {
        const int D = Ds[i];
        std::vector<double> g(D), H(D * D);
        for (int a = 0; a < D; a++) g[a] = grad[32 * size_t(i) + a];
        for (int a = 0; a < D; a++)
            for (int b = 0; b < D; b++) H[a * D + b] = hess[1024 * size_t(i) + a * 32 + b];
        Gauss G;
        ComputeGaussian(pr, sc[i], D, g.data(), H.data(), G);
        for (int a = 0; a < 32; a++) mean[32 * size_t(i) + a] = a < D ? G.mean[a] : 0.0;
        for (int a = 0; a < 1024; a++) { covL[1024 * size_t(i) + a] = 0; invCov[1024 * size_t(i) + a] = 0; }
        for (int a = 0; a < D; a++)
            for (int b = 0; b < D; b++) {
                covL[1024 * size_t(i) + a * 32 + b] = G.covL[a * D + b];
                invCov[1024 * size_t(i) + a * 32 + b] = G.invCov[a * D + b];
            }
        logDet[i] = G.logDet;
    }
}
extern "C" double orc_gaussian_logpdf(int D, const double *x, const double *mean, const double *invCov32, double logDet) {
    Gauss G;
    G.D = D;
    G.mean.assign(mean, mean + D);
    G.invCov.resize(D * D);
    for (int a = 0; a < D; a++)
        for (int b = 0; b < D; b++) G.invCov[a * D + b] = invCov32[a * 32 + b];
    G.logDet = logDet;
    return GaussianLogPdf(x, G);
}
// GenerateSample (gaussian.cpp:36-47) with the caller's standard normals.
extern "C" void orc_gaussian_sample(int D, const double *nrm, const double *mean, const double *covL32, double *x) {
    Gauss G;
    G.D = D;
    G.mean.assign(mean, mean + D);
    G.covL.resize(D * D);
    for (int a = 0; a < D; a++)
        for (int b = 0; b < D; b++) G.covL[a * D + b] = covL32[a * 32 + b];
    GenerateSample(G, nrm, x);
}
extern "C" void orc_jacobi_eigen(int D, const double *A, double *lam, double *V) {
    std::vector<double> a(A, A + D * D), v(D * D);
    JacobiEigen(D, a.data(), v.data());
    for (int i = 0; i < D; i++) lam[i] = a[i * D + i];
    for (int i = 0; i < D * D; i++) V[i] = v[i];
}

// Direct-lighting pre-pass (integrator.h:33-74): film += contrib of the
// unidirectional path of length <= 2, directSpp samples per pixel; the film is
// returned un-normalised (divide by spp when merging).
extern "C" void orc_direct_pass(void *h, int spp, uint64_t seed, int nthreads, float *film) {
    OrcScene *s = (OrcScene *)h;
    const int W = s->f.cam.W, H = s->f.cam.H;
    std::vector<std::vector<float>> films(std::max(nthreads, 1), std::vector<float>(size_t(W) * H * 3, 0.f));
    Opts<float> o;
    ParallelFor(H, nthreads, [&](int y, int t) {
        Philox rng(seed, uint64_t(y));
        PhiloxSrc<float> g(&rng);
        for (int x = 0; x < W; x++)
            for (int k = 0; k < spp; k++) {
                std::vector<Contrib<float>> out;
                GeneratePathUni(s->f, o, x, y, -1, 2, out, g);
                for (auto &c : out) {
                    int ix = std::min(std::max(int(c.screen[0] * W), 0), W - 1);
                    int iy = std::min(std::max(int(c.screen[1] * H), 0), H - 1);
                    float *px = &films[t][(size_t(iy) * W + ix) * 3];
                    if (std::isfinite(c.contrib.x) && std::isfinite(c.contrib.y) && std::isfinite(c.contrib.z)) {
                        px[0] += c.contrib.x; px[1] += c.contrib.y; px[2] += c.contrib.z;
                    }
                }
            }
    });
    for (size_t i = 0; i < size_t(W) * H * 3; i++) {
        float a = 0;
        for (auto &f : films) a += f[i];
        film[i] = a;
    }
}

// Pre-pass (delayed.cpp:14-173): numInit bidirectional samples with min
// length 3, b = sum(lsScore)/numInit, lengthFunc[len] = sum of lsScore per
// length, numChains equal-spaced seeds. Returns the number of candidates.
extern "C" int orc_prepass(void *h, int numInit, int maxDepth, int numChains, uint64_t seed, int nthreads, double *bOut, float *lengthFunc, int nLen,
                PathRec *seeds, float *seedC8) {
    OrcScene *s = (OrcScene *)h;
    Opts<float> o;
    struct Cand { int sample, c, l; float ls; };
    std::vector<std::vector<Cand>> per(numInit);
    ParallelFor(numInit, nthreads, [&](int i, int) {
        Philox rng(seed, uint64_t(i));
        PhiloxSrc<float> g(&rng);
        PathS<float> p;
        std::vector<Contrib<float>> out;
        GeneratePathBidir(s->f, o, 3, maxDepth, p, out, g);
        for (auto &c : out) per[i].push_back({i, c.c, c.l, c.lsScore});
    });
    std::vector<Cand> all;
    double total = 0;
    std::vector<float> lf(nLen, 0.f);
    for (auto &v : per)
        for (auto &c : v) {
            all.push_back(c);
            total += c.ls;
            int len = c.c + c.l - 1;
            if (len < nLen) lf[len] += c.ls;
        }
    for (int i = 0; i < nLen; i++) lengthFunc[i] = lf[i];
    *bOut = total / double(numInit);
    if (all.empty()) return 0;
    std::vector<double> cdf(all.size() + 1, 0.0);
    for (size_t i = 0; i < all.size(); i++) cdf[i + 1] = cdf[i] + all[i].ls;
    const double interval = cdf.back() / double(numChains);
    Philox prng(seed ^ 0x5eed5eedull, 0);
    double pos = prng.uni() * interval;
    int cdfPos = 0;
    for (int i = 0; i < numChains; i++) {
        // delayed.cpp:138-140 caps the cursor at size - 1 and never leaves the
        // loop when pos lies in the last candidate's interval; the walk stops
        // at the cap instead
        while (cdfPos < int(all.size()) - 1 && pos > cdf[cdfPos]) cdfPos++;
        const Cand &cd = all[std::max(cdfPos - 1, 0)];
        Philox rng(seed, uint64_t(cd.sample));
        PhiloxSrc<float> g(&rng);
        PathS<float> p;
        std::vector<Contrib<float>> out;
        GeneratePathBidir(s->f, o, 3, maxDepth, p, out, g);
        Contrib<float> sel;
        for (auto &c : out)
            if (c.c == cd.c && c.l == cd.l) sel = c;
        Truncate(p, sel.c, sel.l);
        ToRec(p, seeds[i]);
        std::vector<Contrib<float>> one(1, sel);
        ContribOut(one, seedC8 + 8 * i);
        seeds[i].screen[0] = sel.screen[0];
        seeds[i].screen[1] = sel.screen[1];
        pos += interval;
    }
    return int(all.size());
}

// CPU DRMLT chain loop (delayed.cpp:334-673) over numChains chains, itersPerChain
// iterations each, nthreads std::threads (chunk size 1, as ParallelFor). The
// film returned is the indirect buffer (un-normalised). opts: type, largeStepProb,
// perturbStdDev, anisoPerturbProb, anisoStdDev, discreteStdDev, roughThr,
// maxDervDepth, acceptanceMap, chadMisQuirk, fdStep, secondStage, h2mc,
// uniformMixing, lensPerturbProb, lensStdDev. stats: the eight DR counters,
// then lens mutations and their acceptances.
extern "C" void orc_render(void *h, const double *opts, const PathRec *seeds, const float *seedC8, int numChains, const float *lengthFunc, int nLen,
                float b, int64_t itersPerChain, int nthreads, uint64_t seed, float *film, uint64_t *stats10, double *seconds) {
    OrcScene *s = (OrcScene *)h;
    DrOptions dro;
    dro.type = int(opts[0]);
    dro.largeStepProb = opts[1];
    dro.perturbStdDev = opts[2];
    dro.anisoPerturbProb = opts[3];
    dro.anisoStdDev = opts[4];
    dro.discreteStdDev = opts[5];
    dro.roughThr = opts[6];
    dro.maxDervDepth = int(opts[7]);
    dro.acceptanceMap = opts[8] != 0;
    dro.chadMisQuirk = opts[9] != 0;
    dro.fdStep = opts[10];
    dro.secondStage = opts[11] != 0;
    dro.h2mc = opts[12] != 0;
    dro.uniformMixing = opts[13];
    dro.lensPerturbProb = opts[14];
    dro.lensStdDev = opts[15];
    if (dro.h2mc) dro.secondStage = false;
    if (!dro.secondStage) dro.acceptanceMap = false;
    Dist1D<float> ld;
    ld.build(lengthFunc, nLen);
    const int W = s->f.cam.W, H = s->f.cam.H;
    const int nt = std::max(nthreads, 1);
    std::vector<std::vector<float>> films(nt, std::vector<float>(size_t(W) * H * 3, 0.f));
    std::vector<DrStats> st(nt);
    auto t0 = std::chrono::steady_clock::now();
    ParallelFor(numChains, nt, [&](int chain, int t) {
        DrChainCtx<float> cx;
        cx.s = &s->f;
        cx.sd = &s->d;
        cx.o.roughThr = float(dro.roughThr);
        cx.od.roughThr = dro.roughThr;
        cx.dro = dro;
        cx.lengthDist = &ld;
        cx.b = b;
        cx.W = W;
        cx.H = H;
        float *fl = films[t].data();
        cx.splat = [&](const float *sp, const V3<float> &v) {
            int ix = std::min(std::max(int(sp[0] * W), 0), W - 1);
            int iy = std::min(std::max(int(sp[1] * H), 0), H - 1);
            if (std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z)) {
                float *px = fl + (size_t(iy) * W + ix) * 3;
                px[0] += v.x; px[1] += v.y; px[2] += v.z;
            }
        };
        Philox rng(seed, uint64_t(chain) + 1000000007ull);
        MState<float> cur, first, second, star;
        cur.valid = true;
        cur.path = FromRec<float>(seeds[chain]);
        const float *c8 = seedC8 + 8 * chain;
        cur.sp.c = seeds[chain].c; cur.sp.l = seeds[chain].l;
        cur.sp.contrib = V3<float>(c8[1], c8[2], c8[3]);
        cur.sp.lsScore = c8[4]; cur.sp.ssScore = c8[5];
        cur.sp.screen[0] = c8[6]; cur.sp.screen[1] = c8[7];
        // a seed's lens score and lensVertexPos: the fixed re-evaluation (the
        // seed records carry neither)
        if (dro.h2mc && dro.lensPerturbProb > 0) cur.sp.lensScore = LensPrepare(s->f, cx.o, cur.path);
        for (int64_t it = 0; it < itersPerChain; it++) DrIteration(cx, cur, first, second, star, rng, st[t]);
    });
    auto t1 = std::chrono::steady_clock::now();
    *seconds = std::chrono::duration<double>(t1 - t0).count();
    for (size_t i = 0; i < size_t(W) * H * 3; i++) {
        float a = 0;
        for (auto &f : films) a += f[i];
        film[i] = a;
    }
    DrStats tot;
    for (auto &x : st) {
        tot.largeStep += x.largeStep; tot.largeStepAcc += x.largeStepAcc; tot.stage1 += x.stage1; tot.stage1Acc += x.stage1Acc;
        tot.iso += x.iso; tot.isoAcc += x.isoAcc; tot.h2mc += x.h2mc; tot.h2mcAcc += x.h2mcAcc;
        tot.lens += x.lens; tot.lensAcc += x.lensAcc;
    }
    stats10[0] = tot.largeStep; stats10[1] = tot.largeStepAcc; stats10[2] = tot.stage1; stats10[3] = tot.stage1Acc;
    stats10[4] = tot.iso; stats10[5] = tot.isoAcc; stats10[6] = tot.h2mc; stats10[7] = tot.h2mcAcc;
    stats10[8] = tot.lens; stats10[9] = tot.lensAcc;
}


// Camera and area-sampling probes for the oracle's own tests.
extern "C" void orc_camera_ray(void *h, double sx, double sy, double *org3, double *dir3, double *minmax) {
    OrcScene *s = (OrcScene *)h;
    Ray<double> r;
    SamplePrimary(s->d.cam, sx, sy, r);
    org3[0] = r.org.x; org3[1] = r.org.y; org3[2] = r.org.z;
    dir3[0] = r.dir.x; dir3[1] = r.dir.y; dir3[2] = r.dir.z;
    minmax[0] = r.minT; minmax[1] = r.maxT;
}
extern "C" int orc_project(void *h, const double *p3, double *s2) {
    OrcScene *s = (OrcScene *)h;
    return ProjectPoint(s->d.cam, V3<double>(p3[0], p3[1], p3[2]), s2[0], s2[1]) ? 1 : 0;
}
extern "C" void orc_sample_tri(void *h, int light, int prim, double u0, double u1, double *pos3, double *nrm3, double *pdf) {
    OrcScene *s = (OrcScene *)h;
    V3<double> p, n;
    SampleTri(s->d, s->d.lights[light], prim, u0, u1, p, n, *pdf);
    pos3[0] = p.x; pos3[1] = p.y; pos3[2] = p.z;
    nrm3[0] = n.x; nrm3[1] = n.y; nrm3[2] = n.z;
}
extern "C" void orc_sample_param(void *h, int prim, const double *pos3, double *u2) {
    OrcScene *s = (OrcScene *)h;
    GetSampleParam(s->d, prim, V3<double>(pos3[0], pos3[1], pos3[2]), u2[0], u2[1]);
}

// ---------------------------------------------------------------- serialized form
// The reference's serialized subpath (Serialize(scene, path, subPath),
// path.cpp:2303-2391) and scene (Serialize(scene), scene.cpp:126-131 with
// camera.cpp:30-36, animatedtransform.h:28-34, bounds.h:30-33 and 48-51),
// written from path records of this scene: the inputs of the reference's
// derivative functions. Record sizes: triangle-mesh shape 46
// (trianglemesh.cpp:3-11), BSDF 10 (bsdf.cpp:7-11), light 56 (light.cpp:7-10).
namespace {
constexpr int kOShape = 46, kOBsdf = 10, kOLight = 56;
struct OWriter {
    float *b;
    int pos = 0;
    void put(float v) { b[pos++] = v; }
    void put3(const V3<float> &v) { put(v.x); put(v.y); put(v.z); }
    void skip(int n) { pos += n; }
};
// TriangleMesh::Serialize (trianglemesh.cpp:143-185), static mesh (the time-1
// copies repeat time 0), no uvs; invArea: the inverse area of the emitter's
// triangle range, else of the triangle
void ser_shape(const SceneT<float> &s, int prim, OWriter &w) {
    const int start = w.pos;
    w.put(0.0f);  // ShapeType::TriangleMesh
    w.put(0.0f);  // isMoving
    const V3<float> p0(s.fp0[3 * prim], s.fp0[3 * prim + 1], s.fp0[3 * prim + 2]);
    const V3<float> e1(s.fe1[3 * prim], s.fe1[3 * prim + 1], s.fe1[3 * prim + 2]);
    const V3<float> e2(s.fe2[3 * prim], s.fe2[3 * prim + 1], s.fe2[3 * prim + 2]);
    for (int t = 0; t < 2; t++) {
        w.put3(p0); w.put3(e1); w.put3(e2);
        w.put3(s.n0[prim]); w.put3(s.n1[prim]); w.put3(s.n2[prim]);
    }
    w.put(1.0f);  // BoolToFloat(st.size() == 0)
    w.skip(6);    // uvs undefined
    const int li = s.triLight[prim];
    const float area = li >= 0 ? s.lights[li].totalArea : 0.5f * Length(Cross(e1, e2));
    w.put(1.0f / area);
    w.pos = start + kOShape;
}
// Lambertian / Phong / RoughDielectric ::Serialize (lambertian.cpp:10-13,
// phong.cpp:13-19, roughdielectric.cpp:13-20)
void ser_bsdf(const SceneT<float> &s, int prim, OWriter &w) {
    const int start = w.pos;
    const Material<float> &M = s.mats[s.triMat[prim]];
    w.put(float(M.type));
    if (M.type == 0) {
        w.put3(M.a);
    } else if (M.type == 1) {
        w.put3(M.a); w.put3(M.b); w.put(M.expo); w.put(M.ksWeight);
    } else {
        w.put3(M.a); w.put3(M.b); w.put(M.eta); w.put(M.invEta); w.put(M.alpha);
    }
    w.pos = start + kOBsdf;
}
// AreaLight::Serialize (arealight.cpp:17-22)
void ser_light(const SceneT<float> &s, int prim, OWriter &w) {
    const int start = w.pos;
    w.put(1.0f);  // LightType::AreaLight
    ser_shape(s, prim, w);
    w.put3(s.lights[s.triLight[prim]].emission);
    w.pos = start + kOLight;
}
// MakeQuaternion (quaternion.cpp:4-36) of a rotation matrix
void quat_of(const float m[4][4], float q[4]) {
    const double tr = double(m[0][0]) + m[1][1] + m[2][2];
    if (tr > 1e-7) {
        double sq = std::sqrt(tr + 1.0);
        q[3] = float(sq / 2.0);
        sq = 0.5 / sq;
        q[0] = float((m[2][1] - m[1][2]) * sq);
        q[1] = float((m[0][2] - m[2][0]) * sq);
        q[2] = float((m[1][0] - m[0][1]) * sq);
    } else {
        const int nxt[3] = {1, 2, 0};
        int i = 0;
        if (m[1][1] > m[0][0]) i = 1;
        if (m[2][2] > m[i][i]) i = 2;
        const int j = nxt[i], k = nxt[j];
        double sq = std::sqrt((m[i][i] - (m[j][j] + m[k][k])) + 1.0);
        double qq[3];
        qq[i] = sq * 0.5;
        if (sq != 0.0) sq = 0.5 / sq;
        q[3] = float((m[k][j] - m[j][k]) * sq);
        qq[j] = (m[j][i] + m[i][j]) * sq;
        qq[k] = (m[k][i] + m[i][k]) * sq;
        q[0] = float(qq[0]); q[1] = float(qq[1]); q[2] = float(qq[2]);
    }
}
}  // namespace

extern "C" void orc_serialize_scene(void *h, int useLightCoord, float *sp) {
    const SceneT<float> &s = ((OrcScene *)h)->f;
    OWriter w{sp};
    w.put(useLightCoord ? 1.0f : 0.0f);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) w.put(s.cam.s2c[j][i]);  // column-major (utils.h:348-355)
    float q[4];
    quat_of(s.cam.c2w, q);
    const V3<float> pos(s.cam.c2w[0][3], s.cam.c2w[1][3], s.cam.c2w[2][3]);
    w.put(0.0f);  // isMoving
    w.put3(pos); w.put3(pos);
    for (int t = 0; t < 2; t++) for (int k = 0; k < 4; k++) w.put(q[k]);
    w.put(float(s.cam.W) * float(s.cam.H));
    w.put(s.cam.dist);
    // BSphere of the scene bounds: center, radius (bounds.h:30-33)
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (int t = 0; t < s.nt; t++)
        for (int v = 0; v < 3; v++)
            for (int a = 0; a < 3; a++) {
                const float x = s.fp0[3 * t + a] + (v == 1 ? s.fe1[3 * t + a] : v == 2 ? s.fe2[3 * t + a] : 0.0f);
                lo[a] = std::min(lo[a], x); hi[a] = std::max(hi[a], x);
            }
    const V3<float> c(0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2]));
    w.put3(c);
    w.put(0.5f * Length(V3<float>(hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2])));
}

// primary: n x pstride (time 0, then u in GetPathPos order); vert: n x
// vstride, zero-filled past what the path writes.
extern "C" void orc_serialize(void *h, int n, const PathRec *recs, int pstride, float *primary, int vstride, float *vert) {
    const SceneT<float> &s = ((OrcScene *)h)->f;
    for (int i = 0; i < n; i++) {
        const PathRec &R = recs[i];
        float *pr = primary + size_t(pstride) * i;
        std::fill(pr, pr + pstride, 0.0f);
        const int P = R.c + R.l - 1, D = 2 * std::max(P, 2);
        pr[0] = 0.0f;  // path.time
        for (int k = 0; k < D; k++) pr[1 + k] = R.u[k];
        float *v = vert + size_t(vstride) * i;
        std::fill(v, v + vstride, 0.0f);
        OWriter w{v};
        w.put3(V3<float>(s.cam.c2w[0][3], s.cam.c2w[1][3], s.cam.c2w[2][3]));  // lensVertexPos
        const int c = R.c, l = R.l;
        if (l > 1) {
            w.put(PickLightProb(s, R.light));
            ser_light(s, R.lprim, w);
            bool done = false;
            for (int d = 0; d < l - 1; d++) {
                ser_shape(s, R.lgprim[d], w);
                w.put(R.ldisc[d]);
                // useAbsoluteParam (path.cpp:273) at the default roughness threshold 0.03
                w.put(Roughness(s.mats[s.triMat[R.lgprim[d]]], R.ldisc[d]) > 0.03f ? 1.0f : 0.0f);
                ser_bsdf(s, R.lgprim[d], w);
                if (d == l - 2 && c == 1) { done = true; break; }
                if (d == l - 2) break;
                w.put(R.lrr[d]);
            }
            if (done) continue;
        }
        for (int d = 0; d < c - 1; d++) {
            ser_shape(s, R.cprim[d], w);
            if (d == c - 2) {
                if (l == 0) {
                    ser_light(s, R.cprim[d], w);
                    w.put(PickLightProb(s, s.triLight[R.cprim[d]]));
                } else if (l == 1) {
                    ser_light(s, R.lprim, w);
                    ser_bsdf(s, R.cprim[d], w);
                    w.put(PickLightProb(s, R.light));
                } else {
                    ser_bsdf(s, R.cprim[d], w);
                }
                break;
            }
            w.put(R.cdisc[d]);
            w.put(Roughness(s.mats[s.triMat[R.cprim[d]]], R.cdisc[d]) > 0.03f ? 1.0f : 0.0f);
            ser_bsdf(s, R.cprim[d], w);
            w.put(R.crr[d]);
        }
    }
}

// ORACLE (test infrastructure only): the DR chain loop of
// DRIntegrator::Render (delayed.cpp:396-606), the mutations it calls
// (mutation.cpp:7-237, drsmallstep.cpp:113-212) and the pre-pass
// (delayed.cpp:14-173), restated for the CPU baseline and for statistical
// tests. Derivatives come from central finite differences in double of the
// fixed-topology log f (the function chad differentiates, path.cpp:3159).
#pragma once
#include "orc_path.h"
#include "orc_gauss.h"
#include <functional>

namespace orc {

// Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy as 1, 2, 3", SC'11)
struct Philox {
    uint32_t key[2];
    uint32_t ctr[4];
    uint32_t buf[4];
    int have = 0;
    Philox(uint64_t k = 0, uint64_t stream = 0) {
        key[0] = uint32_t(k); key[1] = uint32_t(k >> 32);
        ctr[0] = 0; ctr[1] = 0; ctr[2] = uint32_t(stream); ctr[3] = uint32_t(stream >> 32);
    }
    static inline void mulhilo(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
        uint64_t p = uint64_t(a) * uint64_t(b);
        hi = uint32_t(p >> 32); lo = uint32_t(p);
    }
    void block(uint32_t out[4]) {
        uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]}, k0 = key[0], k1 = key[1];
        for (int r = 0; r < 10; r++) {
            uint32_t hi0, lo0, hi1, lo1;
            mulhilo(0xD2511F53u, c[0], hi0, lo0);
            mulhilo(0xCD9E8D57u, c[2], hi1, lo1);
            uint32_t n0 = hi1 ^ c[1] ^ k0, n1 = lo1, n2 = hi0 ^ c[3] ^ k1, n3 = lo0;
            c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
            k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
        }
        for (int i = 0; i < 4; i++) out[i] = c[i];
        if (++ctr[0] == 0) ++ctr[1];
    }
    uint32_t next() {
        if (have == 0) { block(buf); have = 4; }
        return buf[--have];
    }
    double uni() { return double(next() >> 8) * (1.0 / 16777216.0); }
    double normal() {
        double u1 = (double(next() >> 8) + 1.0) * (1.0 / 16777216.0);
        double u2 = uni();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};
template <class F> struct PhiloxSrc : Src<F> {
    Philox *p;
    explicit PhiloxSrc(Philox *p_) : p(p_) {}
    F uni() override { return F(p->uni()); }
};

// std::min / std::max semantics of the reference's Clamp (utils.h:245-248): NaN propagates
template <class F> inline F ClampRef(F v, F lb, F ub) {
    F m = (v < lb) ? lb : v;
    return (ub < m) ? ub : m;
}
template <class F> inline bool IsInvalid(F x) { return std::isnan(double(x)) || std::isinf(double(x)) || x < F(0); }
template <class F> inline bool FlipCoin(F x, F d) { return (x == F(1)) || (d < x); }

struct DrOptions {
    int type = 0;  // 0 green, 1 mira
    double largeStepProb = 0.3, perturbStdDev = 0.01, anisoPerturbProb = 0.5, anisoStdDev = 0.0025;
    double discreteStdDev = 0.01, roughThr = 0.03;
    int maxDervDepth = 9;
    bool lightCoord = true, acceptanceMap = false, chadMisQuirk = true;
    // derivatives of log f for the H2MC Gaussians (the reference takes them
    // from chad's exact AD): central differences of this step in primary-
    // sample units, in FP64. At 1e-4 the Hessians of glossy (Phong,
    // rough dielectric) vertices are off enough to lower the H2MC
    // acceptance visibly; from 3e-6 down the acceptance no longer moves and
    // equals the one exact derivatives give.
    double fdStep = 2e-6;
    // false: integrator=mcmc, the first stage alone (src/integrators/mcmc.cpp:
    // 125-160: large or isotropic small step, splats (1 - a) and a, accept with
    // probability a); the acceptance map is a DR option and stays off there
    bool secondStage = true;
    // integrator=h2mc (src/integrators/h2mc.cpp:196-250 with
    // H2MCSmallStep::Mutate 448-600): a small step is isotropic with
    // probability uniformMixing; otherwise, when the current state has a
    // lens score, a lens mutation with probability lensPerturbProb, else one
    // H2MC move; each accepted with min(1, ratio); no second stage follows
    bool h2mc = false;
    double uniformMixing = 0.2;
    double lensPerturbProb = 0.0, lensStdDev = 0.01;  // dptoptions.h:20-21
};

struct DrStats {
    uint64_t largeStep = 0, largeStepAcc = 0, stage1 = 0, stage1Acc = 0, iso = 0, isoAcc = 0, h2mc = 0, h2mcAcc = 0;
    uint64_t lens = 0, lensAcc = 0;
};

template <class F> struct MState {
    bool valid = false;
    PathS<F> path;
    Contrib<F> sp;
    F scoreSum = 0;
    bool gInit = false;
    Gauss g;
    bool lgInit = false;  // lensGaussianInitialized
    Gauss lg;             // the lens Gaussian (2-D, in screen coordinates)
    bool hasSplat = false;
    F spl[2] = {0, 0};
    V3<F> splVal;
};

// Fixed-topology log ssScore in double (triangles from the state, no occlusion).
inline double LogFFixed(const SceneT<double> &s, const Opts<double> &o, const PathS<double> &p0, const double *off) {
    PathS<double> p = p0;
    double zero[2 * VMAX] = {0};
    std::vector<Contrib<double>> out;
    PerturbPathBidir(s, o, off, zero, p, out, false, true);
    if (out.empty() || !(out[0].ssScore > 0)) return -std::numeric_limits<double>::infinity();
    return std::log(out[0].ssScore);
}
// Central differences, h in primary-sample units: gradient (f(u+h)-f(u-h))/2h,
// Hessian from the four-point stencil (diagonal from f(u+2h), f(u), f(u-2h)).
inline void GradHessFD(const SceneT<double> &s, const Opts<double> &o, const PathS<double> &p, int D, double h, double *grad, double *hess) {
    std::vector<double> off(DMAX, 0.0);
    const double f0 = LogFFixed(s, o, p, off.data());
    auto F = [&](int i, double di, int j, double dj) {
        std::fill(off.begin(), off.end(), 0.0);
        if (i >= 0) off[i] += di;
        if (j >= 0) off[j] += dj;
        return LogFFixed(s, o, p, off.data());
    };
    for (int i = 0; i < D; i++) grad[i] = (F(i, h, -1, 0) - F(i, -h, -1, 0)) / (2 * h);
    for (int i = 0; i < D; i++) {
        hess[i * D + i] = (F(i, 2 * h, -1, 0) - 2 * f0 + F(i, -2 * h, -1, 0)) / (4 * h * h);
        for (int j = i + 1; j < D; j++) {
            double v = (F(i, h, j, h) - F(i, h, j, -h) - F(i, -h, j, h) + F(i, -h, j, -h)) / (4 * h * h);
            hess[i * D + j] = hess[j * D + i] = v;
        }
    }
}

// The function the reference's lens derivatives differentiate
// (PathFuncMode::Lens, path.cpp:2987-3175): log lensScore of the state with
// its screen position moved to `screen`, triangles fixed, no occlusion.
inline double LogLensFixed(const SceneT<double> &s, const Opts<double> &o, const PathS<double> &p0, const double screen[2]) {
    PathS<double> p = p0;
    std::vector<Contrib<double>> out;
    PerturbLensBidir(s, o, screen, p, out, false, true);
    if (out.empty() || !(out[0].lensScore > 0)) return -std::numeric_limits<double>::infinity();
    return std::log(out[0].lensScore);
}
// Central differences of LogLensFixed at the screen position s0 (the stencils of GradHessFD).
inline void LensGradHessFD(const SceneT<double> &s, const Opts<double> &o, const PathS<double> &p, const double s0[2], double h, double *grad,
                           double *hess) {
    auto F = [&](int i, double di, int j, double dj) {
        double sc[2] = {s0[0], s0[1]};
        if (i >= 0) sc[i] += di;
        if (j >= 0) sc[j] += dj;
        return LogLensFixed(s, o, p, sc);
    };
    const double f0 = F(-1, 0, -1, 0);
    for (int i = 0; i < 2; i++) grad[i] = (F(i, h, -1, 0) - F(i, -h, -1, 0)) / (2 * h);
    for (int i = 0; i < 2; i++) hess[i * 2 + i] = (F(i, 2 * h, -1, 0) - 2 * f0 + F(i, -2 * h, -1, 0)) / (4 * h * h);
    hess[1] = hess[2] = (F(0, h, 1, h) - F(0, h, 1, -h) - F(0, -h, 1, h) + F(0, -h, 1, -h)) / (4 * h * h);
}

template <class F> inline PathS<double> ToDouble(const PathS<F> &p) {
    PathS<double> q;
    q.c = p.c; q.l = p.l;
    q.lensPos = V3<double>(p.lensPos.x, p.lensPos.y, p.lensPos.z);
    for (int i = 0; i < 2; i++) { q.screen[i] = p.screen[i]; q.lPos[i] = p.lPos[i]; q.lDir[i] = p.lDir[i]; }
    q.lLight = p.lLight; q.lPrim = p.lPrim;
    auto cv = [](const SurfV<F> &v) {
        SurfV<double> w;
        w.prim = v.prim; w.rnd[0] = v.rnd[0]; w.rnd[1] = v.rnd[1]; w.disc = v.disc; w.useAbs = v.useAbs;
        w.dlLight = v.dlLight; w.dlPrim = v.dlPrim; w.dlRnd[0] = v.dlRnd[0]; w.dlRnd[1] = v.dlRnd[1]; w.rr = v.rr;
        return w;
    };
    for (auto &v : p.cam) q.cam.push_back(cv(v));
    for (auto &v : p.lgt) q.lgt.push_back(cv(v));
    return q;
}

template <class F> struct DrChainCtx {
    const SceneT<F> *s;
    const SceneT<double> *sd;
    Opts<F> o;
    Opts<double> od;
    DrOptions dro;
    const Dist1D<F> *lengthDist;
    F b;
    int W, H;
    std::function<void(const F *, const V3<F> &)> splat;
    double derivSeconds = 0;
};

template <class F> inline void InitGaussian(DrChainCtx<F> &cx, MState<F> &st) {
    const int c = st.sp.c, l = st.sp.l, P = c + l - 1, D = Dim<F>(st.path.c, st.path.l);
    const bool hasFunc = c >= 1 && l >= 0 && c + l >= 3 && P <= cx.dro.maxDervDepth;
    H2Param pr(cx.dro.perturbStdDev);
    if (hasFunc) {
        std::vector<double> g(D, 0.0), h(D * D, 0.0);
        if (double(st.sp.ssScore) > 1e-15) {
            PathS<double> pd = ToDouble(st.path);
            Opts<double> od = cx.od;
            od.chadMisQuirk = cx.dro.chadMisQuirk;
            GradHessFD(*cx.sd, od, pd, D, cx.dro.fdStep, g.data(), h.data());
            bool fin = true;
            for (double x : g) fin = fin && std::isfinite(x);
            for (double x : h) fin = fin && std::isfinite(x);
            if (!fin) { std::fill(g.begin(), g.end(), 0.0); std::fill(h.begin(), h.end(), 0.0); }
        }
        ComputeGaussian(pr, double(st.sp.ssScore), D, g.data(), h.data(), st.g);
    } else {
        IsotropicGaussian(D, pr.sigma, st.g);
    }
    st.gInit = true;
}

// initLensGaussian (h2mc.cpp:466-499): lens derivatives exist for c + l > 3
// within the derivative depth (path.cpp:3195-3197, 3268-3283); zero
// derivatives when lensScore <= 1e-15 or non-finite; ComputeGaussian with the
// lens sigma and the state's ssScore (h2mc.cpp:607-644).
template <class F> inline void InitLensGaussian(DrChainCtx<F> &cx, MState<F> &st) {
    const int c = st.sp.c, l = st.sp.l, P = c + l - 1;
    H2Param pr(cx.dro.lensStdDev);
    if (c >= 1 && l >= 0 && c + l > 3 && P <= cx.dro.maxDervDepth) {
        double g[2] = {0, 0}, h[4] = {0, 0, 0, 0};
        if (double(st.sp.lensScore) > 1e-15) {
            const PathS<double> pd = ToDouble(st.path);
            const double s0[2] = {double(st.sp.screen[0]), double(st.sp.screen[1])};
            LensGradHessFD(*cx.sd, cx.od, pd, s0, cx.dro.fdStep, g, h);
            bool fin = true;
            for (double x : g) fin = fin && std::isfinite(x);
            for (double x : h) fin = fin && std::isfinite(x);
            if (!fin) { g[0] = g[1] = 0; h[0] = h[1] = h[2] = h[3] = 0; }
        }
        ComputeGaussian(pr, double(st.sp.ssScore), 2, g, h, st.lg);
    } else {
        IsotropicGaussian(2, pr.sigma, st.lg);
    }
    st.lgInit = true;
}

template <class F> inline void DiscOffsets(Philox &rng, double sd, F *d) {
    for (int i = 0; i < 2 * VMAX; i++) d[i] = F(sd * rng.normal());
}

// mutation.cpp:126-208 (SmallStep::Mutate with sigma)
template <class F> inline F SmallStepMutate(DrChainCtx<F> &cx, MState<F> &cur, MState<F> &prop, double sigma, Philox &rng) {
    prop.path = cur.path;
    const int D = Dim<F>(cur.path.c, cur.path.l);
    F off[DMAX] = {0}, disc[2 * VMAX];
    for (int i = 0; i < D; i++) off[i] = F(sigma * rng.normal());
    DiscOffsets(rng, cx.dro.discreteStdDev, disc);
    std::vector<Contrib<F>> out;
    PerturbPathBidir(*cx.s, cx.o, off, disc, prop.path, out);
    prop.gInit = false;
    prop.lgInit = false;  // the reference leaves the proposal object's flag as it was (see DESIGN)
    F a = 0;
    if (!out.empty()) {
        prop.sp = out[0];
        a = ClampRef(prop.sp.ssScore / cur.sp.ssScore, F(0), F(1));
        prop.hasSplat = true;
        prop.spl[0] = out[0].screen[0]; prop.spl[1] = out[0].screen[1];
        prop.splVal = out[0].contrib * (cx.b / out[0].lsScore);
    }
    return a;
}
// mutation.cpp:210-237 (SmallStep::Trace)
template <class F> inline F Trace(DrChainCtx<F> &cx, const F *off, MState<F> &prop, MState<F> &star, Philox &rng) {
    if (!prop.valid) return F(0);
    star.path = prop.path;
    F disc[2 * VMAX];
    DiscOffsets(rng, cx.dro.discreteStdDev, disc);
    std::vector<Contrib<F>> out;
    PerturbPathBidir(*cx.s, cx.o, off, disc, star.path, out);
    star.gInit = false;
    if (out.empty()) return F(0);
    star.sp = out[0];
    return ClampRef(star.sp.ssScore / prop.sp.ssScore, F(0), F(1));
}
// mutation.cpp:7-124 (LargeStep::Mutate, multiplexed)
template <class F> inline F LargeStepMutate(DrChainCtx<F> &cx, MState<F> &cur, MState<F> &prop, Philox &rng) {
    PhiloxSrc<F> g(&rng);
    const int len = cx.lengthDist->SampleDiscrete(g.uni());
    const int lgt = std::min(std::max(int(g.uni() * F(len + 1)), 0), len);
    const int cam = len - lgt + 1;
    std::vector<Contrib<F>> out;
    GenerateSubpath(*cx.s, cx.o, cam, lgt, prop.path, out, g);
    prop.gInit = false;
    prop.lgInit = false;
    F a = 0;
    if (!out.empty()) {
        (void)g.uni();  // contribution pick (one candidate when multiplexed)
        prop.sp = out[0];
        prop.scoreSum = out[0].lsScore;
        if (cur.valid) {
            const int lc = cur.sp.c + cur.sp.l - 1, lp = prop.sp.c + prop.sp.l - 1;
            a = ClampRef((F(lp + 1) * prop.sp.lsScore / cx.lengthDist->Pmf(lp)) /
                             (F(lc + 1) * cur.sp.lsScore / cx.lengthDist->Pmf(lc)),
                         F(0), F(1));
        }
        prop.hasSplat = true;
        prop.spl[0] = out[0].screen[0]; prop.spl[1] = out[0].screen[1];
        prop.splVal = out[0].contrib * (cx.b / prop.scoreSum);
    }
    Truncate(prop.path, cam, lgt);
    return a;
}
// drsmallstep.cpp:113-212 (PH2MCSmallStep::Mutate)
template <class F> inline double H2MCMutate(DrChainCtx<F> &cx, MState<F> &cur, MState<F> &prop, Philox &rng) {
    const int D = Dim<F>(cur.path.c, cur.path.l);
    if (!cur.gInit) InitGaussian(cx, cur);
    double n[DMAX], offd[DMAX];
    for (int i = 0; i < D; i++) n[i] = rng.normal();
    GenerateSample(cur.g, n, offd);
    F off[DMAX] = {0}, disc[2 * VMAX];
    for (int i = 0; i < D; i++) off[i] = F(offd[i]);
    DiscOffsets(rng, cx.dro.discreteStdDev, disc);
    prop.path = cur.path;
    std::vector<Contrib<F>> out;
    PerturbPathBidir(*cx.s, cx.o, off, disc, prop.path, out);
    prop.lgInit = false;  // h2mc.cpp:583
    if (out.empty()) return 0.0;
    prop.sp = out[0];
    InitGaussian(cx, prop);
    double neg[DMAX];
    for (int i = 0; i < D; i++) neg[i] = -double(off[i]);
    double offr[DMAX];
    for (int i = 0; i < D; i++) offr[i] = double(off[i]);
    double py = GaussianLogPdf(offr, cur.g), px = GaussianLogPdf(neg, prop.g);
    double a = std::exp(px - py) * double(prop.sp.ssScore) / double(cur.sp.ssScore);
    prop.hasSplat = false;
    if (out[0].lsScore > F(0) && std::isfinite(double(out[0].lsScore))) {
        prop.hasSplat = true;
        prop.spl[0] = out[0].screen[0]; prop.spl[1] = out[0].screen[1];
        prop.splVal = out[0].contrib * (cx.b / out[0].lsScore);
    }
    return a;
}

// h2mc.cpp:463-530 (the lens branch of H2MCSmallStep::Mutate): offset ~ the
// current state's lens Gaussian, screen position moved by it (mod 1),
// PerturbLensBidir, acceptance min(1, N_z(-off) / N_x(off) lensScore_z /
// lensScore_x).
template <class F> inline F LensMutateAt(DrChainCtx<F> &cx, MState<F> &cur, MState<F> &prop, const F *off);
template <class F> inline F LensMutate(DrChainCtx<F> &cx, MState<F> &cur, MState<F> &prop, Philox &rng) {
    if (!cur.lgInit) InitLensGaussian(cx, cur);
    double n[2], offd[2];
    n[0] = rng.normal();
    n[1] = rng.normal();
    GenerateSample(cur.lg, n, offd);
    const F off[2] = {F(offd[0]), F(offd[1])};
    return LensMutateAt(cx, cur, prop, off);
}
// The lens mutation by a given screen offset (the current state's lens
// Gaussian is initialised when missing).
template <class F> inline F LensMutateAt(DrChainCtx<F> &cx, MState<F> &cur, MState<F> &prop, const F *off) {
    if (!cur.lgInit) InitLensGaussian(cx, cur);
    const F screen[2] = {Modulo(cur.sp.screen[0] + off[0], F(1)), Modulo(cur.sp.screen[1] + off[1], F(1))};
    prop.path = cur.path;
    std::vector<Contrib<F>> out;
    PerturbLensBidir(*cx.s, cx.o, screen, prop.path, out);
    prop.gInit = false;
    prop.lgInit = false;
    if (out.empty()) return F(0);
    prop.sp = out[0];
    InitLensGaussian(cx, prop);
    const double x[2] = {double(off[0]), double(off[1])}, neg[2] = {-double(off[0]), -double(off[1])};
    const double py = GaussianLogPdf(x, cur.lg), px = GaussianLogPdf(neg, prop.lg);
    const F a = ClampRef(F(std::exp(px - py) * double(prop.sp.lensScore) / double(cur.sp.lensScore)), F(0), F(1));
    prop.hasSplat = true;
    prop.spl[0] = out[0].screen[0]; prop.spl[1] = out[0].screen[1];
    prop.splVal = out[0].contrib * (cx.b / out[0].lsScore);
    return a;
}

template <class F> inline void WeightedSplat(DrChainCtx<F> &cx, const MState<F> &st, F w) {
    if (cx.dro.acceptanceMap || !st.hasSplat) return;
    V3<F> v = st.splVal * w;
    if (!v.isZero()) cx.splat(st.spl, v);
}

// One iteration of delayed.cpp:396-606 for one chain.
template <class F> inline void DrIteration(DrChainCtx<F> &cx, MState<F> &cur, MState<F> &first, MState<F> &second, MState<F> &star,
                                           Philox &rng, DrStats &st) {
    F a1 = 0, aStar = 0;
    double a2 = 0;
    bool acc1 = false, acc2 = false, isAniso = false;
    const bool isLarge = FlipCoin(F(cx.dro.largeStepProb), F(rng.uni()));
    // h2mc small steps (h2mc.cpp:451-466): the uniform-mixing coin, then,
    // when the current state has a lens score, the lens coin
    bool h2mcMove = false, lensMove = false;
    if (cur.valid && !isLarge && cx.dro.h2mc && !(double(rng.uni()) < cx.dro.uniformMixing)) {
        lensMove = cur.sp.lensScore > F(0) && double(rng.uni()) < cx.dro.lensPerturbProb;
        h2mcMove = !lensMove;
    }
    if (!cur.valid || isLarge) {
        st.largeStep++;
        a1 = LargeStepMutate(cx, cur, first, rng);
        if (IsInvalid(a1)) { a1 = 0; acc1 = false; }
        else { acc1 = FlipCoin(a1, F(rng.uni())); if (acc1) st.largeStepAcc++; }
    } else if (lensMove) {
        // the lens mutation: a first-stage move accepted with probability a
        st.lens++;
        a1 = LensMutate(cx, cur, first, rng);
        if (IsInvalid(a1)) { a1 = 0; acc1 = false; }
        else { acc1 = FlipCoin(a1, F(rng.uni())); if (acc1) st.lensAcc++; }
    } else if (h2mcMove) {
        // the H2MC move as the only stage: a1 = 0 (so w1 = 0 below), a2 = min(1, ratio)
        isAniso = true;
        st.h2mc++;
        a2 = H2MCMutate(cx, cur, second, rng);
        if (IsInvalid(a2)) a2 = 0;
        a2 = std::min(1.0, a2);
        acc2 = FlipCoin(F(a2), F(rng.uni()));
        if (acc2) st.h2mcAcc++;
    } else {
        st.stage1++;
        a1 = SmallStepMutate(cx, cur, first, cx.dro.perturbStdDev, rng);
        if (IsInvalid(a1)) { a1 = 0; acc1 = false; }
        else { acc1 = FlipCoin(a1, F(rng.uni())); if (acc1) st.stage1Acc++; }
    }
    const bool doSecond = cx.dro.secondStage && !acc1 && !isLarge && cur.valid;
    if (doSecond) {
        isAniso = FlipCoin(F(cx.dro.anisoPerturbProb), F(rng.uni()));
        if (isAniso) { a2 = H2MCMutate(cx, cur, second, rng); st.h2mc++; }
        else { a2 = double(SmallStepMutate(cx, cur, second, cx.dro.anisoStdDev, rng)); st.iso++; }
        if (IsInvalid(a2)) {
            a2 = 0;
        } else if (cx.dro.type == 0) {
            const int D = Dim<F>(cur.path.c, cur.path.l);
            F x[DMAX], y[DMAX], off[DMAX] = {0};
            GetPathPos(cur.path, x);
            GetPathPos(first.path, y);
            for (int i = 0; i < D; i++) off[i] = x[i] - y[i];
            aStar = Trace(cx, off, second, star, rng);
            if (IsInvalid(aStar)) aStar = 0;
            if (aStar == F(1)) {
                a2 = 0;
            } else {
                a2 = std::min(1.0, a2 * double(1.0f - float(aStar)) / double(1.0f - float(a1)));
                acc2 = FlipCoin(F(a2), F(rng.uni()));
                if (acc2) { if (isAniso) st.h2mcAcc++; else st.isoAcc++; }
            }
        } else {
            aStar = std::min(F(1), first.sp.ssScore / second.sp.ssScore);
            if (IsInvalid(aStar)) aStar = 1;
            if (aStar == F(1)) {
                a2 = 0;
            } else {
                const int D = Dim<F>(cur.path.c, cur.path.l);
                F x[DMAX], y[DMAX], z[DMAX];
                GetPathPos(cur.path, x);
                GetPathPos(first.path, y);
                GetPathPos(second.path, z);
                Gauss gi;
                IsotropicGaussian(D, cx.dro.perturbStdDev, gi);
                double zy[DMAX], xy[DMAX];
                for (int i = 0; i < D; i++) { zy[i] = double(z[i]) - double(y[i]); xy[i] = double(x[i]) - double(y[i]); }
                double tr = std::exp(GaussianLogPdf(zy, gi) - GaussianLogPdf(xy, gi));
                if (IsInvalid(tr)) {
                    a2 = 0;
                } else {
                    a2 = std::min(1.0, a2 * tr * double(1.0f - float(aStar)) / double(1.0f - float(a1)));
                    acc2 = FlipCoin(F(a2), F(rng.uni()));
                    if (acc2) { if (isAniso) st.h2mcAcc++; else st.isoAcc++; }
                }
            }
        }
    }
    const F w1 = a1, w2 = (F(1) - w1) * F(a2), wx = F(1) - w1 - w2;
    if (cur.valid && wx > F(0)) WeightedSplat(cx, cur, wx);
    if (w1 > F(0)) WeightedSplat(cx, first, w1);
    if (w2 > F(0)) WeightedSplat(cx, second, w2);
    if (acc1 || acc2) {
        if (acc1) {
            Truncate(first.path, first.sp.c, first.sp.l);
            std::swap(cur, first);
            cur.gInit = false;
            if (!isLarge && cx.dro.acceptanceMap && first.hasSplat) cx.splat(first.spl, V3<F>(1, 0, 0));
        } else {
            Truncate(second.path, second.sp.c, second.sp.l);
            std::swap(cur, second);
            if (!isAniso) cur.gInit = false;
            if (cx.dro.acceptanceMap && second.hasSplat) cx.splat(second.spl, V3<F>(0, 1, 0));
        }
        cur.valid = true;
    }
}

}  // namespace orc

// gfx950 kernels of the DRMLT mutation loop (DRIntegrator::Render,
// delayed.cpp:334-673) and its parity surface.
//
// Plain form: one thread runs one whole DR iteration of one chain
// (k_dr_plain), one launch per step. Batched form: the iteration is split at
// the expensive boundaries -- stage 1, derivatives of the current state,
// Gaussian, stage-2 proposal, derivatives of the proposal, Gaussian, stage-2
// acceptance -- into persistent kernels over work lists binned by path
// length. Both forms draw from the same Philox streams (key = seed, counter =
// iteration, chain, stage tag), so for a given seed they produce the same
// chains; only the order of the floating-point film additions differs.
#pragma once
#include "drmlt_kernels.h"
#include "dev_walk.h"

namespace drmlt {

enum : uint32_t { TAG_S1 = 1, TAG_S2 = 2, TAG_S3 = 3, TAG_LENS = 4, TAG_PRE = 16, TAG_DIRECT = 17 };

// ------------------------------------------------------------------ helpers
struct GenArray {  // uniform stream from an array (parity tests)
    const float *u;
    int n, i;
    DEV float uni() { return i < n ? u[i++] : 0.5f; }
};
struct DiscNormal {  // N(0, sigma) draws for the discrete dimensions
    Philox *g;
    float sigma;
    DEV float operator()(int) const { return float(sigma * g->normal()); }
};
struct OffNormal {   // isotropic proposal offsets N(0, sigma)
    Philox *g;
    float sigma;
    DEV float operator()(int) const { return float(sigma * g->normal()); }
};

DEV bool is_invalid(double x) { return isnan(x) || isinf(x) || x < 0; }         // delayed.cpp:252
DEV bool flip_coin(double x, double d) { return (x == 1.0) || (d < x); }        // delayed.cpp:250
DEV float clamp_ref(float v) { float m = (v < 0.0f) ? 0.0f : v; return (1.0f < m) ? 1.0f : m; }  // utils.h:245-248 (NaN passes)

DEV void store_contrib(const Chains &ch, int p, const POut<float> &o) {
    float *rgb = &ch.rgb[p], *scr = &ch.screen[p];
    rgb[0] = o.contrib.x;
    rgb[1] = o.contrib.y;
    rgb[2] = o.contrib.z;
    ch.ls[p] = o.lsScore;
    ch.ss[p] = o.ssScore;
    scr[0] = o.screen[0];
    scr[1] = o.screen[1];
}

DEV void splat(const Chains &ch, const DrParams &P, int W, int H, float *film, int p, float w) {
    if (P.acceptanceMap || !ch.hasSplat[p]) return;
    const float *rgb = &ch.rgb[p], *scr = &ch.screen[p];
    const float k = w * (P.b / ch.ls[p]);
    const float r = rgb[0] * k, g = rgb[1] * k, b = rgb[2] * k;
    if (r == 0.0f && g == 0.0f && b == 0.0f) return;
    if (!(isfinite(r) && isfinite(g) && isfinite(b))) return;
    const int ix = min(max(int(scr[0] * W), 0), W - 1), iy = min(max(int(scr[1] * H), 0), H - 1);
    float *px = film + 3 * (size_t(iy) * W + ix);
    atomicAdd(px, r);
    atomicAdd(px + 1, g);
    atomicAdd(px + 2, b);
}
DEV void splat_const(const Chains &ch, int W, int H, float *film, int p, float r, float g, float b) {
    if (!ch.hasSplat[p]) return;
    const float *scr = &ch.screen[p];
    const int ix = min(max(int(scr[0] * W), 0), W - 1), iy = min(max(int(scr[1] * H), 0), H - 1);
    float *px = film + 3 * (size_t(iy) * W + ix);
    atomicAdd(px, r);
    atomicAdd(px + 1, g);
    atomicAdd(px + 2, b);
}

// Algorithmic chain-state bytes of one path state: 4 D (u) + 4 (P + 1)
// (bsdfDiscrete) + 4 (P + 1) (triangle ids) + 16 (c, l, scores). A path
// evaluation reads one state and writes one. Counted in the counting kernels
// only.
DEV int state_bytes(int c, int l) {
    const int P = c + l - 1, D = (P > 2 ? P : 2) * 2;
    return 4 * D + 8 * (P + 1) + 16;
}
DEV void count_bytes(const DevScene &S, unsigned long long b) {
#ifdef DRMLT_COUNT_WORK
    atomicAdd(S.counters + 3, b);
#else
    (void)S; (void)b;
#endif
}
DEV void count_state(const DevScene &S, int c, int l) { count_bytes(S, (unsigned long long)(2 * state_bytes(c, l))); }

// wave-aggregated counter increment; the lowest active lane adds for the
// wave, so the call may sit in divergent code
DEV void count_if(unsigned long long *ctr, bool pred) {
    const unsigned long long m = __ballot(pred);
    const int leader = __ffsll((long long)__ballot(1)) - 1;
    if (int(threadIdx.x & 63) == leader && m) atomicAdd(ctr, (unsigned long long)__popcll(m));
}
enum { ST_LARGE = 0, ST_LARGE_ACC, ST_S1, ST_S1_ACC, ST_ISO, ST_ISO_ACC, ST_H2MC, ST_H2MC_ACC, ST_LENS, ST_LENS_ACC, ST_COUNT };
static_assert(ST_COUNT == kStatCounters, "loop counters");

DEV int perm_cur(int perm) { return perm & 3; }
DEV int perm_first(int perm) { return (perm >> 2) & 3; }
DEV int perm_second(int perm) { return (perm >> 4) & 3; }
DEV int make_perm(int c, int f, int s) { return c | (f << 2) | (s << 4); }

// ------------------------------------------------------------ derivatives
DEV int jet_passes(int D) {
    const int nch = (D + kJetK - 1) / kJetK;
    int n = 0;
    for (int i = 0; i < D; i++) n += nch - i / kJetK;
    return n;
}
DEV void jet_pass_index(int D, int p, int &row, int &col0) {
    const int nch = (D + kJetK - 1) / kJetK;
    for (int i = 0; i < D; i++) {
        const int k = nch - i / kJetK;
        if (p < k) { row = i; col0 = (i / kJetK + p) * kJetK; return; }
        p -= k;
    }
    row = 0; col0 = 0;
}
// One forward-over-forward pass: gradient block and Hessian row `row`,
// columns [col0, col0 + K), of log f at the state behind ref.
// Writes G(j) and Hrow(j) through the callbacks. Returns false if the
// fixed-topology evaluation produced no contribution.
template <class S, class GW, class HW>
DEV bool jet_pass(const DevScene &Sc, const Opts &o, const SRef &ref, int D, int row, int col0, GW gw, HW hw) {
    JetU<S, kJetK> up{ref.u, ref.s, 0, row, col0};
    POut<Jet<S, kJetK>> out;
    perturb_eval<Jet<S, kJetK>, true>(Sc, o, ref, up, DiscNone{}, out);
    if (!out.valid || !(out.ssScore.v > S(0))) return false;
    const Jet<S, kJetK> lf = s_log(out.ssScore);
    for (int j = 0; j < kJetK; j++) {
        const int c = col0 + j;
        if (c < D) {
            gw(c, lf.g[j]);
            hw(c, lf.h[j]);
        }
    }
    return true;
}

// Derivative function availability (path.cpp:3186-3206 with maxDervDepth)
DEV bool has_derivative(int c, int l, int maxD) { return c >= 1 && l >= 0 && c + l >= 3 && c + l - 1 <= maxD; }

// Derivative passes of state slot p into the derivative buffers of `chain`,
// all in one thread (the plain form; k_b_deriv spreads the same passes over
// threads). A failed fixed-topology evaluation marks the gradient non-finite.
DEV void derivs_thread(const DevScene &Sc, const Chains &ch, const DrParams &P, int chain, int p) {
    const int c = ch.c[p], l = ch.l[p], D = dim_of(c, l);
    if (!has_derivative(c, l, P.maxDervDepth) || D > kGMax || !(double(ch.ss[p]) > 1e-15)) return;
    float *gr = ch.grad + size_t(chain) * kGMax, *hs = ch.hess + size_t(chain) * kGMax * kGMax;
    const SRef ref = slot_ref(ch, p);
    const int np = jet_passes(D);
    for (int q = 0; q < np; q++) {
        int row, col0;
        jet_pass_index(D, q, row, col0);
        const bool ok = jet_pass<float>(Sc, P.oDeriv, ref, D, row, col0,
                                        [&](int cc, float gv) { if (row == 0) gr[cc] = gv; },
                                        [&](int cc, float hv) { if (cc >= row) hs[row * kGMax + cc] = hv; });
        if (!ok) { gr[0] = __builtin_nanf(""); return; }
    }
}

// initGaussian (drsmallstep.cpp:129-166) of state slot p from the derivative
// buffers of `chain`: isotropic without derivatives (path too long, f <= 1e-15
// or a non-finite gradient / Hessian, drsmallstep.cpp:135-163), else
// ComputeGaussian. Marks the slot's Gaussian as initialised.
template <class X, class W> DEV void gaussian_from_derivs(const X &x, const Chains &ch, const DrParams &P, int chain, int p, W &w) {
    const int c = ch.c[p], l = ch.l[p], D = dim_of(c, l);
    const GSlot G = g_ref(ch, p);
    const double ssv = ch.ss[p];
    const float *gr = ch.grad + size_t(chain) * kGMax, *hs = ch.hess + size_t(chain) * kGMax * kGMax;
    bool derivs = has_derivative(c, l, P.maxDervDepth) && D <= kGMax && D <= W::kMaxD;
    if (derivs) {
        if (x.lane == 0) w.flag = ssv > 1e-15 ? 1 : 0;
        x.sync();
        if (w.flag) {
            for (int t = x.lane; t < D + D * D; t += x.n) {
                const int i = (t - D) / D, j = (t - D) % D;
                const float v = t < D ? gr[t] : (i <= j ? hs[i * kGMax + j] : 0.0f);
                if (!isfinite(v)) w.flag = 0;
            }
        }
        x.sync();
        derivs = w.flag != 0;
        x.sync();
    }
    if (derivs) {
        gauss_compute(x, P.h2, ssv, D, [&](int i) { return double(gr[i]); }, [&](int i, int j) { return double(hs[i * kGMax + j]); },
                      w, G);
    } else {
        gauss_iso(x, D, P.h2.sigma, G);
        x.sync();
    }
    if (x.lane == 0) ch.gInit[p] = 1;
}

// H2MC proposal offset, drawn from the current state's Gaussian with the
// stage-2 stream's first D normals (drsmallstep.cpp:171-179,
// gaussian.cpp:36-47); dr_stage2_propose skips those normals.
template <class X, class W> DEV void h2mc_offset(const X &x, const Chains &ch, const DrParams &P, int chain, uint32_t iter, W &w) {
    const int C = ch.C;
    const int pc = chain + C * perm_cur(ch.perm[chain]);
    const int D = dim_of(ch.c[pc], ch.l[pc]);
    const Philox g(P.seed, iter, chain, TAG_S2);
    gauss_sample(x, D, g_ref(ch, pc), [&](int j) { return g.normal_at(j); }, w,
                 [&](int i, double v) { chain_off(ch, chain)[i] = float(v); });
}

// H2MC acceptance before the delayed-rejection correction
// (drsmallstep.cpp:181-196): N_z(-off) / N_x(off) * f_z / f_x, unclamped.
template <class X, class W> DEV void h2mc_ratio(const X &x, const Chains &ch, int chain, W &w) {
    const int C = ch.C;
    const int perm = ch.perm[chain];
    const int pc = chain + C * perm_cur(perm), ps = chain + C * perm_second(perm);
    const int D = dim_of(ch.c[pc], ch.l[pc]);
    const float *o = chain_off(ch, chain);
    auto off = [&](int i) { return double(o[i]); };
    const double py = gauss_logpdf(x, D, off, 1.0, g_ref(ch, pc), w);
    const double px = gauss_logpdf(x, D, off, -1.0, g_ref(ch, ps), w);
    if (x.lane == 0) ch.a2d[chain] = exp(px - py) * double(ch.ss[ps]) / double(ch.ss[pc]);
}

// ------------------------------------------------------------ lens mutation
// The h2mc integrator's lens mutation (h2mc.cpp:463-530). A state's lens
// score and lensVertexPos come from its fixed-topology re-evaluation with the
// lens bookkeeping (lens_state_eval), computed once per state and kept in the
// slot header (lens flags bit 0) and its lens slot.
DEV void lens_score_ensure(const DevScene &Sc, const Chains &ch, const DrParams &P, int chain) {
    const int pc = chain + ch.C * perm_cur(ch.perm[chain]);
    if (ch.lflags[pc] & 1) return;
    float ls = 0.0f;
    V3<float> lp(0.0f, 0.0f, 0.0f);
    const int c = ch.c[pc], l = ch.l[pc];
    const bool wellFormed = c >= 1 && l >= 0 && c <= kVMax + 1 && l <= kVMax + 1 && c + l - 1 >= 2 && 2 * (c + l - 1) <= kDMax;
    if (ch.valid[pc] && wellFormed) ls = lens_state_eval(Sc, P.o, slot_ref(ch, pc), lp);
    ch.lscore[pc] = isfinite(ls) ? ls : 0.0f;
    set_lens_pos(ch, pc, lp);
    ch.lflags[pc] = (ch.lflags[pc] & ~1) | 1;
}
// initLensGaussian (h2mc.cpp:466-499) of state slot p: lens derivatives exist
// for c + l > 3 within the derivative depth (path.cpp:3195-3197); they are
// zero when lensScore <= 1e-15 or not finite; ComputeGaussian with the lens
// sigma and the state's ssScore (h2mc.cpp:607-644), else isotropic.
DEV void lens_gaussian(const DevScene &Sc, const Chains &ch, const DrParams &P, int p, GWorkT<2> &w) {
    const SeqExec x;
    const int c = ch.c[p], l = ch.l[p];
    const LSlot G = lens_ref(ch, p);
    if (c >= 1 && l >= 0 && c + l > 3 && c + l - 1 <= P.maxDervDepth) {
        float g[2] = {0.0f, 0.0f}, h[3] = {0.0f, 0.0f, 0.0f};
        if (double(ch.lscore[p]) > 1e-15) {
            const float scr[2] = {(&ch.screen[p])[0], (&ch.screen[p])[1]};
            const bool ok = lens_jets(Sc, P.o, slot_ref(ch, p), scr, lens_pos(ch, p), g, h);
            if (!ok || !(isfinite(g[0]) && isfinite(g[1]) && isfinite(h[0]) && isfinite(h[1]) && isfinite(h[2]))) {
                g[0] = g[1] = 0.0f;
                h[0] = h[1] = h[2] = 0.0f;
            }
        }
        gauss_compute(x, P.h2lens, double(ch.ss[p]), 2, [&](int i) { return double(g[i]); },
                      [&](int i, int j) { return double(i == 0 ? (j == 0 ? h[0] : h[1]) : h[2]); }, w, G);
    } else {
        gauss_iso(x, 2, P.h2lens.sigma, G);
    }
    ch.lflags[p] = ch.lflags[p] | 2;
}
// One lens mutation of `chain` into its first-proposal slot, accepted with
// min(1, N_z(-off) / N_x(off) lensScore_z / lensScore_x); splats (1 - a, a),
// state swap and counts as h2mc.cpp:206-243 do them. Draws from the chain's
// lens stream. The plain form calls it from dr_stage1, the batched form from
// k_b_lens over list 5.
DEV void lens_move(const DevScene &Sc, const Chains &ch, const DrParams &P, int chain, uint32_t iter, float *film, unsigned long long *stats) {
    Philox g(P.seed, iter, chain, TAG_LENS);
    const SeqExec x;
    GWorkT<2> w;
    const int perm = ch.perm[chain], C = ch.C;
    const int pc = chain + C * perm_cur(perm), pf = chain + C * perm_first(perm);
    if (!(ch.lflags[pc] & 2)) lens_gaussian(Sc, ch, P, pc, w);
    float off[2];
    gauss_sample(x, 2, lens_ref(ch, pc), [&](int) { return g.normal(); }, w, [&](int i, double v) { off[i] = float(v); });
    const float scr[2] = {modulo1((&ch.screen[pc])[0] + off[0]), modulo1((&ch.screen[pc])[1] + off[1])};
    const V3<float> lp = lens_pos(ch, pc);
    copy_state(slot_ref(ch, pc), slot_ref(ch, pf));
    POut<float> out;
    out.lensPos = lp;
    perturb_lens_eval<float, false>(Sc, P.o, slot_ref(ch, pf), scr, lp, out);
    ch.gInit[pf] = 0;
    ch.lflags[pf] = 0;
    count_state(Sc, ch.c[pc], ch.l[pc]);
    float a = 0.0f;
    if (out.valid) {
        store_contrib(ch, pf, out);
        ch.hasSplat[pf] = 1;
        ch.lscore[pf] = out.lensScore;
        set_lens_pos(ch, pf, out.lensPos);
        ch.lflags[pf] = 1;
        lens_gaussian(Sc, ch, P, pf, w);
        const double py = gauss_logpdf(x, 2, [&](int i) { return double(off[i]); }, 1.0, lens_ref(ch, pc), w);
        const double px = gauss_logpdf(x, 2, [&](int i) { return double(off[i]); }, -1.0, lens_ref(ch, pf), w);
        a = clamp_ref(float(exp(px - py) * double(out.lensScore) / double(ch.lscore[pc])));
    }
    bool acc = false;
    if (is_invalid(a)) a = 0.0f;
    else acc = flip_coin(a, g.uni());
    count_if(stats + ST_LENS, true);
    count_if(stats + ST_LENS_ACC, acc);
    ch.a1[chain] = a;
    ch.flags[chain] = 128 | (acc ? 1 : 0) | (out.valid ? 32 : 0);
    chain_off(ch, chain)[0] = off[0];  // the screen offset, for drmlt_read_step
    chain_off(ch, chain)[1] = off[1];
    const int W = Sc.cam.W, H = Sc.cam.H;
    const float wx = 1.0f - a;
    if (wx > 0.0f) splat(ch, P, W, H, film, pc, wx);
    if (a > 0.0f) splat(ch, P, W, H, film, pf, a);
    if (acc) {
        ch.perm[chain] = make_perm(perm_first(perm), perm_cur(perm), perm_second(perm));
        ch.valid[pf] = 1;
    }
}

// ------------------------------------------------------------ DR stages
// Each stage is split into a beginning (random draws that choose the move,
// the proposal slot), the path evaluation, and an end (acceptance, splats,
// state swap). The plain form evaluates the path with perturb_eval /
// generate_subpath on the spot; the batched form runs the same evaluation as
// a walk (dev_walk.h) whose rays the lanes of a wave trace together.

// Stage 1 (delayed.cpp:404-445): a large step (LargeStep::Mutate,
// mutation.cpp:7-124, multiplexed) or an isotropic small step
// (SmallStep::Mutate, mutation.cpp:126-165) into the first-proposal slot.
struct Stage1 {
    int perm, pc, pf;
    bool isLarge, curValid;
    bool h2mcMove;  // h2mc integrator: this small step is an H2MC move, made by the second-stage kernels
    bool lensMove;  // h2mc integrator: this small step is a lens mutation (lens_move)
    int cam, lgt;  // large step: the sampled strategy
};
// copyState: copy the current state into the proposal slot for a small step
// (the batched form does it while binning, and begins again without it).
DEV Stage1 stage1_begin(const Chains &ch, const DrParams &P, int chain, Philox &g, bool copyState = true) {
    Stage1 a;
    const int C = ch.C;
    a.perm = ch.perm[chain];
    a.pc = chain + C * perm_cur(a.perm);
    a.pf = chain + C * perm_first(a.perm);
    a.isLarge = flip_coin(P.largeStepProb, g.uni());
    // a current state outside the fixed vertex slots is treated as invalid
    const int cc = ch.c[a.pc], cl = ch.l[a.pc];
    const bool wellFormed = cc >= 1 && cl >= 0 && cc <= kVMax + 1 && cl <= kVMax + 1 && cc + cl - 1 >= 2 && 2 * (cc + cl - 1) <= kDMax;
    a.curValid = ch.valid[a.pc] != 0 && wellFormed;
    a.cam = a.lgt = 0;
    a.h2mcMove = false;
    a.lensMove = false;
    if (!a.curValid || a.isLarge) {
        const int len = sample_discrete(P.lenCdf, P.nLen, g.uni());
        a.lgt = min(max(int(g.uni() * float(len + 1)), 0), len);
        a.cam = len - a.lgt + 1;
    } else {
        // h2mc (H2MCSmallStep::Mutate, h2mc.cpp:448-466): isotropic with
        // probability uniformMixingProbability; else, when the current state
        // has a lens score (known: lens_score_ensure ran), a lens mutation with
        // probability lensPerturbProb, else an H2MC move
        if (P.integrator == 2) {
            a.h2mcMove = !(g.uni() < P.uniformMixing);
            if (a.h2mcMove && P.lensProb > 0.0f && ch.lscore[a.pc] > 0.0f && g.uni() < P.lensProb) {
                a.lensMove = true;
                a.h2mcMove = false;
            }
        }
        if (copyState && !a.h2mcMove && !a.lensMove) copy_state(slot_ref(ch, a.pc), slot_ref(ch, a.pf));
    }
    return a;
}
// Length (edges) of the path stage 1 evaluates.
DEV int stage1_length(const Chains &ch, const Stage1 &a) {
    return (!a.curValid || a.isLarge) ? a.cam + a.lgt - 1 : ch.c[a.pc] + ch.l[a.pc] - 1;
}
// The walk of stage 1 on slot pf (the small step's state was copied there).
DEV void stage1_walk(const Chains &ch, const DrParams &P, const Stage1 &a, Philox &g, Walk &w) {
    w.p = a.pf;
    w.g = &g;
    if (!a.curValid || a.isLarge) {
        w.mode = WK_GEN;
        w.c = a.cam;
        w.l = a.lgt;
    } else {
        w.mode = WK_OFF_NORMAL;
        w.sigma = P.perturbStdDev;
        w.c = ch.c[a.pc];
        w.l = ch.l[a.pc];
    }
}
// Returns the chain's flags; finishes the iteration (splats, swap) when no
// second stage follows.
DEV int stage1_end(const DevScene &Sc, const Chains &ch, const DrParams &P, int chain, const Stage1 &a, const POut<float> &out, Philox &g,
                   float *film, unsigned long long *stats) {
    if (a.h2mcMove) {
        // no first-stage proposal: the second-stage kernels make the H2MC
        // move with a1 = 0 and accept it with min(1, ratio)
        ch.a1[chain] = 0.0f;
        ch.flags[chain] = 2 | 4;
        return 2 | 4;
    }
    if (a.lensMove) {
        // made by lens_move (flag bit 7)
        ch.a1[chain] = 0.0f;
        ch.flags[chain] = 128;
        return 128;
    }
    const int perm = a.perm, pc = a.pc, pf = a.pf;
    const bool isLarge = a.isLarge, curValid = a.curValid;
    float a1 = 0.0f;
    bool acc1 = false;
    ch.lflags[pf] = 0;  // a new state: its lens score and lens Gaussian are not known
    if (!curValid || isLarge) {
        count_state(Sc, a.cam, a.lgt);
        ch.gInit[pf] = 0;
        if (out.valid) {
            (void)g.uni();  // contribution pick (one candidate when multiplexed)
            store_contrib(ch, pf, out);
            ch.hasSplat[pf] = 1;
            if (curValid) {
                const int lc = ch.c[pc] + ch.l[pc] - 1, lp = a.cam + a.lgt - 1;
                const float pmfP = P.lenFunc[lp] / (P.lenFuncInt * float(P.nLen));
                const float pmfC = P.lenFunc[lc] / (P.lenFuncInt * float(P.nLen));
                a1 = clamp_ref((float(lp + 1) * out.lsScore / pmfP) / (float(lc + 1) * ch.ls[pc] / pmfC));
            }
        }
    } else {
        count_state(Sc, ch.c[pc], ch.l[pc]);
        ch.gInit[pf] = 0;
        if (out.valid) {
            store_contrib(ch, pf, out);
            ch.hasSplat[pf] = 1;
            a1 = clamp_ref(out.ssScore / ch.ss[pc]);
        }
    }
    if (is_invalid(a1)) { a1 = 0.0f; acc1 = false; }
    else acc1 = flip_coin(a1, g.uni());
    const bool large = !curValid || isLarge;
    count_if(stats + ST_LARGE, large);
    count_if(stats + ST_LARGE_ACC, large && acc1);
    count_if(stats + ST_S1, !large);
    count_if(stats + ST_S1_ACC, !large && acc1);
    const bool doSecond = P.integrator == 0 && !acc1 && !isLarge && curValid;
    bool isAniso = false;
    if (doSecond) isAniso = flip_coin(P.anisoPerturbProb, g.uni());
    ch.a1[chain] = a1;
    const int flags = (acc1 ? 1 : 0) | (doSecond ? 2 : 0) | (isAniso ? 4 : 0) | (isLarge ? 8 : 0) | (out.valid ? 32 : 0);
    ch.flags[chain] = flags;
    if (!doSecond) {
        const int W = Sc.cam.W, H = Sc.cam.H;
        const float w1 = a1, wx = 1.0f - w1;
        if (curValid && wx > 0.0f) splat(ch, P, W, H, film, pc, wx);
        if (w1 > 0.0f) splat(ch, P, W, H, film, pf, w1);
        if (acc1) {
            ch.perm[chain] = make_perm(perm_first(perm), perm_cur(perm), perm_second(perm));
            ch.gInit[pf] = 0;
            if (!isLarge && P.acceptanceMap) splat_const(ch, W, H, film, pc, 1.0f, 0.0f, 0.0f);
            ch.valid[pf] = 1;
        }
    }
    return flags;
}
// Plain form of stage 1.
DEV int dr_stage1(const DevScene &Sc, const Chains &ch, const DrParams &P, int chain, uint32_t iter, float *film,
                  unsigned long long *stats) {
    Philox g(P.seed, iter, chain, TAG_S1);
    if (P.lensProb > 0.0f) lens_score_ensure(Sc, ch, P, chain);
    const Stage1 a = stage1_begin(ch, P, chain, g);
    if (a.lensMove) {
        lens_move(Sc, ch, P, chain, iter, film, stats);
        return ch.flags[chain];
    }
    const SRef rf = slot_ref(ch, a.pf);
    POut<float> out;
    out.valid = false;
    if (a.h2mcMove) {
        // made by the second stage
    } else if (!a.curValid || a.isLarge) {
        generate_subpath(Sc, P.o, a.cam, a.lgt, rf, g, out);
    } else {
        PerturbU<OffNormal> up{rf.u, rf.s, 0, OffNormal{&g, P.perturbStdDev}};
        perturb_eval<float, false>(Sc, P.o, rf, up, DiscNormal{&g, P.discreteStdDev}, out);
    }
    return stage1_end(Sc, ch, P, chain, a, out, g, film, stats);
}

// Stage-2 proposal (delayed.cpp:448-463): H2MC (drsmallstep.cpp:113-212) or
// isotropic with the second-stage sigma (mutation.cpp:167-208) into the
// second-proposal slot. For H2MC the offset must already be drawn
// (h2mc_offset); its ratio needs the proposal's Gaussian and is finished by
// h2mc_ratio.
struct Stage2 {
    int pc, ps, flags;
    bool isAniso;
};
DEV Stage2 stage2_begin(const Chains &ch, int chain, Philox &g) {
    Stage2 a;
    const int C = ch.C;
    const int perm = ch.perm[chain];
    a.pc = chain + C * perm_cur(perm);
    a.ps = chain + C * perm_second(perm);
    a.flags = ch.flags[chain];
    a.isAniso = (a.flags & 4) != 0;
    copy_state(slot_ref(ch, a.pc), slot_ref(ch, a.ps));
    // the H2MC offset was drawn by h2mc_offset from this stream's first D normals
    if (a.isAniso) g.skip_normals(dim_of(ch.c[a.pc], ch.l[a.pc]));
    return a;
}
DEV void stage2_walk(const Chains &ch, const DrParams &P, int chain, const Stage2 &a, Philox &g, Walk &w) {
    w.p = a.ps;
    w.g = &g;
    w.c = ch.c[a.pc];
    w.l = ch.l[a.pc];
    if (a.isAniso) {
        w.mode = WK_OFF_ARRAY;
        w.offChain = chain;
    } else {
        w.mode = WK_OFF_NORMAL;
        w.sigma = P.anisoStdDev;
    }
}
DEV void stage2_end(const DevScene &Sc, const Chains &ch, int chain, const Stage2 &a, const POut<float> &out) {
    const int pc = a.pc, ps = a.ps;
    if (!a.isAniso) ch.gInit[ps] = 0;
    ch.lflags[ps] = 0;  // h2mc.cpp:583: the proposal's lens Gaussian is not initialised
    count_state(Sc, ch.c[pc], ch.l[pc]);
    double a2 = 0.0;
    if (out.valid) {
        store_contrib(ch, ps, out);
        if (a.isAniso) {
            ch.hasSplat[ps] = (out.lsScore > 0.0f && isfinite(out.lsScore)) ? 1 : 0;
        } else {
            ch.hasSplat[ps] = 1;
            a2 = double(clamp_ref(out.ssScore / ch.ss[pc]));
        }
    }
    ch.a2d[chain] = a2;
    ch.flags[chain] = a.flags | (out.valid ? 16 : 0);
}
// Plain form of the stage-2 proposal.
DEV void dr_stage2_propose(const DevScene &Sc, const Chains &ch, const DrParams &P, int chain, uint32_t iter) {
    Philox g(P.seed, iter, chain, TAG_S2);
    const Stage2 a = stage2_begin(ch, chain, g);
    const SRef rs = slot_ref(ch, a.ps);
    POut<float> out;
    if (a.isAniso) {
        PerturbU<OffArray> up{rs.u, rs.s, 0, OffArray{chain_off(ch, chain), 1}};
        perturb_eval<float, false>(Sc, P.o, rs, up, DiscNormal{&g, P.discreteStdDev}, out);
    } else {
        PerturbU<OffNormal> up{rs.u, rs.s, 0, OffNormal{&g, P.anisoStdDev}};
        perturb_eval<float, false>(Sc, P.o, rs, up, DiscNormal{&g, P.discreteStdDev}, out);
    }
    stage2_end(Sc, ch, chain, a, out);
}

// Second-stage acceptance (delayed.cpp:466-541), splats (548-561) and state
// swap (565-606). For Green & Mira the y* = z + (x - y) state
// (mutation.cpp:210-237) is evaluated in the spare fourth slot.
struct Stage3 {
    int perm, pc, pf, ps, pstar, flags;
    bool isAniso, traceStar;
    float a1;
    double a2;
};
DEV Stage3 stage3_begin(const Chains &ch, const DrParams &P, int chain) {
    Stage3 a;
    const int C = ch.C;
    a.perm = ch.perm[chain];
    a.pc = chain + C * perm_cur(a.perm);
    a.pf = chain + C * perm_first(a.perm);
    a.ps = chain + C * perm_second(a.perm);
    a.pstar = chain + 3 * C;
    a.flags = ch.flags[chain];
    a.isAniso = (a.flags & 4) != 0;
    const bool zValid = (a.flags & 16) != 0;
    a.a1 = ch.a1[chain];
    a.a2 = zValid ? ch.a2d[chain] : 0.0;
    a.traceStar = !is_invalid(a.a2) && P.type == 0 && P.integrator == 0 && a.a2 != 0.0 && ch.valid[a.ps];
    if (a.traceStar) copy_state(slot_ref(ch, a.ps), slot_ref(ch, a.pstar));
    return a;
}
DEV void stage3_walk(const Chains &ch, const DrParams &P, const Stage3 &a, Philox &g, Walk &w) {
    w.p = a.pstar;
    w.g = &g;
    w.mode = WK_OFF_DIFF;
    w.px = a.pc;
    w.py = a.pf;
    w.c = ch.c[a.ps];
    w.l = ch.l[a.ps];
}
// Second-stage acceptance probability from a1, the second stage's own ratio
// a2 (before the delayed-rejection correction), aStar and the Tierney-Mira
// kernel ratio tr: Green & Mira (delayed.cpp:466-500; aStar is y*'s
// acceptance from z, 0 when y* was not traced), Tierney & Mira
// (delayed.cpp:502-541; aStar = min(1, f_y / f_z), tr = Q1(z, y) / Q1(x, y)),
// or the h2mc integrator's min(1, a2) (h2mc.cpp:584-587). draw tells whether
// the reference draws the acceptance coin.
DEV double dr_alpha2(int integrator, int type, float a1, double a2, float aStar, double tr, bool &draw) {
    draw = false;
    if (is_invalid(a2)) return 0.0;
    if (integrator == 2) {
        draw = true;
        return fmin(1.0, a2);
    }
    if (type == 0) {
        if (is_invalid(aStar)) aStar = 0.0f;
        if (aStar == 1.0f) return 0.0;
        draw = true;
        return fmin(1.0, a2 * double(1.0f - aStar) / double(1.0f - a1));
    }
    if (is_invalid(aStar)) aStar = 1.0f;
    if (aStar == 1.0f) return 0.0;
    if (is_invalid(tr)) return 0.0;
    draw = true;
    return fmin(1.0, a2 * tr * double(1.0f - aStar) / double(1.0f - a1));
}
// Q1(z, y) / Q1(x, y) of the isotropic first-stage kernel (delayed.cpp:514-527).
DEV double mira_kernel_ratio(const Chains &ch, const DrParams &P, int pc, int pf, int ps) {
    const int D = dim_of(ch.c[pc], ch.l[pc]);
    const float *uc = slot_u(ch, pc), *uf = slot_u(ch, pf), *us = slot_u(ch, ps);
    double qzy = 0, qxy = 0;
    for (int i = 0; i < D; i++) {
        const double y = uf[i];
        const double dz = double(us[i]) - y, dx = double(uc[i]) - y;
        qzy += dz * dz;
        qxy += dx * dx;
    }
    const double s2 = double(P.perturbStdDev) * double(P.perturbStdDev);
    return exp(-0.5 * qzy / s2 + 0.5 * qxy / s2);
}

DEV void stage3_end(const DevScene &Sc, const Chains &ch, const DrParams &P, int chain, const Stage3 &a, const POut<float> &starOut,
                    Philox &g, float *film, unsigned long long *stats) {
    const int perm = a.perm, pc = a.pc, pf = a.pf, ps = a.ps;
    const bool isAniso = a.isAniso;
    const float a1 = a.a1;
    float aStar = 0.0f;
    double tr = 1.0;
    if (!is_invalid(a.a2) && P.integrator == 0) {
        if (P.type == 0) {
            if (a.traceStar) {
                count_state(Sc, ch.c[ps], ch.l[ps]);
                if (starOut.valid) aStar = clamp_ref(starOut.ssScore / ch.ss[ps]);
            }
        } else {
            aStar = fminf(1.0f, ch.ss[pf] / ch.ss[ps]);
            if (!is_invalid(aStar) && aStar != 1.0f) tr = mira_kernel_ratio(ch, P, pc, pf, ps);
        }
    }
    bool draw;
    const double a2 = dr_alpha2(P.integrator, P.type, a1, a.a2, aStar, tr, draw);
    const bool acc2 = draw && flip_coin(a2, g.uni());
    ch.a2d[chain] = a2;  // the iteration's alpha2, y*'s validity and score (read back by drmlt_read_step)
    if (a.traceStar && starOut.valid) {
        ch.flags[chain] = a.flags | 64;
        ch.ss[a.pstar] = starOut.ssScore;
    }
    count_if(stats + ST_ISO, !isAniso);
    count_if(stats + ST_ISO_ACC, !isAniso && acc2);
    count_if(stats + ST_H2MC, isAniso);
    count_if(stats + ST_H2MC_ACC, isAniso && acc2);
    const int W = Sc.cam.W, H = Sc.cam.H;
    const float w1 = a1, w2 = (1.0f - w1) * float(a2), wx = 1.0f - w1 - w2;
    if (ch.valid[pc] && wx > 0.0f) splat(ch, P, W, H, film, pc, wx);
    if (w1 > 0.0f) splat(ch, P, W, H, film, pf, w1);
    if (w2 > 0.0f) splat(ch, P, W, H, film, ps, w2);
    if (acc2) {
        ch.perm[chain] = make_perm(perm_second(perm), perm_first(perm), perm_cur(perm));
        if (!isAniso) ch.gInit[ps] = 0;
        if (P.acceptanceMap) splat_const(ch, W, H, film, pc, 0.0f, 1.0f, 0.0f);
        ch.valid[ps] = 1;
    }
}
// Plain form of the second-stage acceptance.
DEV void dr_stage2_accept(const DevScene &Sc, const Chains &ch, const DrParams &P, int chain, uint32_t iter, float *film,
                          unsigned long long *stats) {
    Philox g(P.seed, iter, chain, TAG_S3);
    const Stage3 a = stage3_begin(ch, P, chain);
    POut<float> out;
    out.valid = false;
    if (a.traceStar) {
        const SRef rt = slot_ref(ch, a.pstar);
        PerturbU<OffDiff> up{rt.u, rt.s, 0, OffDiff{slot_u(ch, a.pc), slot_u(ch, a.pf), 1, 1}};
        perturb_eval<float, false>(Sc, P.o, rt, up, DiscNormal{&g, P.discreteStdDev}, out);
    }
    stage3_end(Sc, ch, P, chain, a, out, g, film, stats);
}

constexpr int kPreMaxL = 16;  // light-subpath vertices of a pre-pass sample

// records: the drmlt_path_rec layout (134 x 4-byte words)
DEV SRef rec_ref(char *base) {
    SRef r;
    int *ip = (int *)base;
    float *fp = (float *)base;
    r.c = ip; r.l = ip + 1; r.light = ip + 2; r.lprim = ip + 3;
    r.screen = fp + 4;
    r.u = fp + 6;
    r.cprim = ip + 38; r.lgprim = ip + 54;
    r.cdisc = fp + 70; r.ldisc = fp + 86; r.crr = fp + 102; r.lrr = fp + 118;
    r.s = 1;
    r.vs = 1;
    return r;
}
constexpr int kRecBytes = 536;

DEV void contrib8(float *c8, const POut<float> &o) {
    c8[0] = o.valid ? 1.0f : 0.0f;
    if (!o.valid) { for (int k = 1; k < 8; k++) c8[k] = 0.0f; return; }
    c8[1] = o.contrib.x; c8[2] = o.contrib.y; c8[3] = o.contrib.z;
    c8[4] = o.lsScore; c8[5] = o.ssScore; c8[6] = o.screen[0]; c8[7] = o.screen[1];
}

}  // namespace drmlt

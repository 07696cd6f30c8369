// Kernel-side data shared by drmlt_kernels.hip and the C ABI: chain state in
// HBM and the per-step work lists of the batched DR pipeline.
#pragma once
#include "dev_path.h"
#include "dev_gauss.h"

namespace drmlt {

constexpr int kJetK = 2;     // Hessian columns per derivative pass
constexpr int kMaxBins = 16; // path-length bins (path length 0..15)
// Gaussians of dimension up to kGSmallD are computed by groups of
// DRMLT_GAUSS_GROUP lanes, several per wave; larger ones by a whole wave.
constexpr int kGSmallD = 8;
#ifndef DRMLT_GAUSS_GROUP
#define DRMLT_GAUSS_GROUP 8  // lanes per small Gaussian (measured: 8 beats 16 and 32)
#endif

// Chain state: 4 physical slots per chain (current, first- and second-stage
// proposal, Green's y*), all fields structure-of-arrays with stride NS = 4C.
// perm[chain] packs which of the first three physical slots plays current /
// first / second (2 bits each), so accepting a proposal is an index swap.
struct Chains {
    int C, NS, GS;  // chains, state slots (4C), Gaussian slots (3C)
    int *c, *l, *light, *lprim;
    float *screen, *u;
    int *cprim, *lgprim;
    float *cdisc, *ldisc, *crr, *lrr;
    int *valid, *hasSplat, *gInit;
    float *rgb, *ls, *ss;  // rgb [3][NS]
    int *perm;
    double *gauss;         // [3C][kGSlot] Gaussian slots (dev_gauss.h), one per state slot 0..3C-1
    // per chain scratch of the DR step
    float *off;            // [32][C] H2MC offset
    float *grad, *hess;    // [C][24], [C][24*24] derivative output (Hessian: upper triangle)
    float *a1, *a2;        // stage acceptances
    double *a2d;
    int *flags;            // bit0 acc1, bit1 doSecond, bit2 isAniso, bit3 large, bit4 z valid, bit5 y valid, bit6 y* valid
};

DEV SRef slot_ref(const Chains &ch, int p) {
    SRef r;
    r.c = ch.c + p; r.l = ch.l + p; r.light = ch.light + p; r.lprim = ch.lprim + p;
    r.screen = ch.screen + p; r.u = ch.u + p;
    r.cprim = ch.cprim + p; r.lgprim = ch.lgprim + p;
    r.cdisc = ch.cdisc + p; r.ldisc = ch.ldisc + p; r.crr = ch.crr + p; r.lrr = ch.lrr + p;
    r.s = ch.NS;
    return r;
}
DEV GSlot g_ref(const Chains &ch, int p) { return GSlot{ch.gauss + size_t(p) * kGSlot}; }

// DR options the kernels read (src/dptoptions.h:8-37)
struct DrParams {
    float largeStepProb, perturbStdDev, anisoPerturbProb, anisoStdDev, discreteStdDev;
    int type;  // 0 green, 1 mira
    int maxDervDepth;
    int acceptanceMap;
    float b;   // normalisation (average brightness from the pre-pass)
    int nLen;  // length distribution size
    const float *lenFunc, *lenCdf;
    float lenFuncInt;
    uint64_t seed;
    H2Param h2;
    Opts o;
    Opts oDeriv;  // options for the derivative evaluation (chad MIS quirk)
    int binStages;  // path-stage kernels take their chains from the length bins (1) or in chain order (0)
    int integrator;   // 0 drmlt (delayed rejection), 1 mcmc (the first stage alone), 2 h2mc (no lens moves)
    float uniformMixing;  // h2mc: probability of an isotropic small step
    int stageLanes;   // chains per 64-lane workgroup of the path-stage kernels (64, 32 or 16)
    int chainBase, chainCount;  // the chains [chainBase, chainBase + chainCount) the batched stage kernels take
};

// Work lists of the batched pipeline, each binned by path length:
// kind 0: H2MC chains (the current state's Gaussian, then the offset);
// kind 1: chains with a second stage (binned by the current path length);
// kind 2: H2MC chains whose proposal is valid (its Gaussian and the ratio);
// kind 3: every chain, binned by the length of the path its first stage
//         evaluates (the current length for a small step, the sampled one
//         for a large step);
// kind 4: the H2MC chains of kind 0 whose current state needs derivatives
//         (no cached Gaussian, a path within the derivative depth, a
//         positive score).
constexpr int kListKinds = 5;
// most chain ranges of the graph-replayed step on concurrent streams
constexpr int kMaxParts = 4;
struct Lists {
    int *count;  // [kMaxBins] per list kind
    int *items;  // [kMaxBins][C]
};

}  // namespace drmlt

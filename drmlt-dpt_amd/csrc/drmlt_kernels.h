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

// One path-state slot is a record of kSlotWords 32-bit words (nine 64-byte
// lines): a header line (c, l, light, light triangle, valid, hasSplat,
// gInit, lens score, contribution rgb, scores, screen position, lens flags), two lines of primary
// samples u[32], then the camera and the light vertices, each vertex's
// (triangle, discrete sample, Russian-roulette factor) side by side. The
// stage kernels take chains in length-bin order, so neighbouring lanes hold
// unrelated slots; with whole records a lane's reads and writes of one state
// stay within a few lines instead of one line per field.
constexpr int kSlotWords = 144;
enum : int {
    SO_C = 0, SO_L = 1, SO_LIGHT = 2, SO_LPRIM = 3, SO_VALID = 4, SO_HASSPLAT = 5, SO_GINIT = 6,
    SO_LENS = 7,    // lens score (h2mc with lens mutations; valid when SO_LFLAGS bit 0 is set)
    SO_RGB = 8, SO_LS = 11, SO_SS = 12, SO_SCREEN = 13,
    SO_LFLAGS = 15, // bit 0: lens score and lensVertexPos known; bit 1: lens Gaussian initialised
    SO_U = 16,    // u[32]
    SO_CV = 48,   // camera vertex k: words SO_CV + 3k .. + 2
    SO_LV = 96,   // light vertex k: words SO_LV + 3k .. + 2
};
static_assert(SO_LV + 3 * kVMax == kSlotWords && SO_U + kDMax == SO_CV, "slot record layout");
// A per-slot field: the field of slot p sits kSlotWords words after the
// one of slot p - 1.
template <class T> struct SlotField {
    T *base;
    DEV T &operator[](size_t p) const { return base[p * kSlotWords]; }
};

// Chain state: 4 physical slots per chain (current, first- and second-stage
// proposal, Green's y*); slot p = chain + C * k is the record at
// slots + p * kSlotWords. perm[chain] packs which of the first three
// physical slots plays current / first / second (2 bits each), so accepting
// a proposal is an index swap.
struct Chains {
    int C, NS, GS;  // chains, state slots (4C), Gaussian slots (3C)
    uint32_t *slots;       // [NS][kSlotWords] state records
    SlotField<int> c, l, light, lprim, valid, hasSplat, gInit, lflags;
    SlotField<float> rgb, ls, ss, screen, lscore;  // rgb and screen: the first of 3 and 2 words
    int *perm;
    double *gauss;         // [3C][kGSlot] Gaussian slots (dev_gauss.h), one per state slot 0..3C-1
    double *lensG;         // [3C][kLSlot] lens Gaussian and lensVertexPos per state slot (h2mc with lens mutations; else null)
    // per chain scratch of the DR step
    float *off;            // [C][32] H2MC offset
    float *grad, *hess;    // [C][24], [C][24*24] derivative output (Hessian: upper triangle)
    float *a1, *a2;        // stage acceptances
    double *a2d;
    int *flags;            // bit0 acc1, bit1 doSecond, bit2 isAniso, bit3 large, bit4 z valid, bit5 y valid, bit6 y* valid
};
// Points the per-slot fields of `ch` into ch.slots (host side, after allocation).
inline void bind_slot_fields(Chains &ch) {
    int *i = reinterpret_cast<int *>(ch.slots);
    float *f = reinterpret_cast<float *>(ch.slots);
    ch.c.base = i + SO_C; ch.l.base = i + SO_L; ch.light.base = i + SO_LIGHT; ch.lprim.base = i + SO_LPRIM;
    ch.valid.base = i + SO_VALID; ch.hasSplat.base = i + SO_HASSPLAT; ch.gInit.base = i + SO_GINIT;
    ch.lflags.base = i + SO_LFLAGS;
    ch.rgb.base = f + SO_RGB; ch.ls.base = f + SO_LS; ch.ss.base = f + SO_SS; ch.screen.base = f + SO_SCREEN;
    ch.lscore.base = f + SO_LENS;
}
DEV float *slot_u(const Chains &ch, int p) { return reinterpret_cast<float *>(ch.slots + size_t(p) * kSlotWords) + SO_U; }
DEV float *chain_off(const Chains &ch, int chain) { return ch.off + size_t(chain) * 32; }

DEV SRef slot_ref(const Chains &ch, int p) {
    uint32_t *b = ch.slots + size_t(p) * kSlotWords;
    int *ib = reinterpret_cast<int *>(b);
    float *fb = reinterpret_cast<float *>(b);
    SRef r;
    r.c = ib + SO_C; r.l = ib + SO_L; r.light = ib + SO_LIGHT; r.lprim = ib + SO_LPRIM;
    r.screen = fb + SO_SCREEN; r.u = fb + SO_U;
    r.cprim = ib + SO_CV; r.cdisc = fb + SO_CV + 1; r.crr = fb + SO_CV + 2;
    r.lgprim = ib + SO_LV; r.ldisc = fb + SO_LV + 1; r.lrr = fb + SO_LV + 2;
    r.s = 1;
    r.vs = 3;
    return r;
}
DEV GSlot g_ref(const Chains &ch, int p) { return GSlot{ch.gauss + size_t(p) * kGSlot}; }

// Lens slot of a state: its 2-D lens Gaussian (screen coordinates) and
// lensVertexPos, kLSlot doubles (a 128-byte record).
constexpr int kLSlot = 16;
using LSlot = GSlotT<2>;  // V 4, p 2, mean 2, logDet 1 at doubles 0..8
DEV LSlot lens_ref(const Chains &ch, int p) { return LSlot{ch.lensG + size_t(p) * kLSlot}; }
DEV V3<float> lens_pos(const Chains &ch, int p) {
    const double *b = ch.lensG + size_t(p) * kLSlot + 9;
    return V3<float>(float(b[0]), float(b[1]), float(b[2]));
}
DEV void set_lens_pos(const Chains &ch, int p, const V3<float> &v) {
    double *b = ch.lensG + size_t(p) * kLSlot + 9;
    b[0] = v.x; b[1] = v.y; b[2] = v.z;
}

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
    int integrator;   // 0 drmlt (delayed rejection), 1 mcmc (the first stage alone), 2 h2mc
    float uniformMixing;  // h2mc: probability of an isotropic small step
    float lensProb;   // h2mc: lensPerturbProb (0: no lens mutations, no lens bookkeeping)
    H2Param h2lens;   // h2mc: the lens Gaussian's parameters (sigma = lensPerturbStdDev)
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
//         positive score);
// kind 5: h2mc chains making a lens mutation.
constexpr int kListKinds = 6;
// loop counters (the eight of drmlt_stats' DR part, then lens mutations and their acceptances)
constexpr int kStatCounters = 10;
// most chain ranges of the graph-replayed step on concurrent streams
constexpr int kMaxParts = 4;
struct Lists {
    int *count;  // [kMaxBins] per list kind
    int *items;  // [kMaxBins][C]
};

}  // namespace drmlt

// Kernels of the batched DR step (one per stage, persistent over work lists
// binned by path length). Included twice: by k_loop.hip as the product
// kernels and by k_count.hip, with DRMLT_COUNT_WORK defined, as the counting
// kernels (suffix _counted) behind drmlt_count_work.
#pragma once
#include "dr_stages.h"

namespace drmlt {

// The path-stage kernels are compiled twice, bounded to two and to three
// resident waves per SIMD (256 and 168 VGPRs): more waves hide more
// traversal latency, at the price of spills; which pays depends on how many
// waves of chains there are (the schedule's stage_waves picks).
// The same for the Gaussian and derivative kernels.
#ifndef DRMLT_GAUSS_WAVES
#define DRMLT_GAUSS_WAVES 1
#endif
#ifndef DRMLT_DERIV_WAVES
#define DRMLT_DERIV_WAVES 1
#endif

#ifdef DRMLT_COUNT_WORK
#define DRMLT_STAGE(name) name##_counted
#else
#define DRMLT_STAGE(name) name
#endif

// ------------------------------------------------------------ batched form
// The step's work lists (drmlt_kernels.h: Lists) are filled by the kernels as
// they go, binned by path length; each kernel takes its items from the bins
// concatenated in bin order, so the lanes of a wave mostly evaluate paths of
// one length and take the same number of walk rounds.
// Wave-aggregated list_push for every active lane with `pred`: one atomic per
// bin present in the wave, and the wave's items of one bin are stored next to
// each other in lane order (neighbouring chains stay neighbours in the list).
// Called by all active lanes of a wave together. With a key, the wave's
// items of one bin are also grouped by key (the derivative lists group by
// camera-subpath length, so a wave of derivative passes mostly evaluates one
// strategy).
DEV void list_push_wave(const Lists &L, int kind, int C, int bin, int item, bool pred, int key = 0) {
    const int lane = int(threadIdx.x & 63);
    unsigned long long todo = __ballot(pred);
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const int b = __shfl(bin, leader), k = __shfl(key, leader);
        const bool mine = pred && bin == b && key == k;
        const unsigned long long m = __ballot(mine);
        int base = 0;
        if (lane == leader) base = atomicAdd(&L.count[kind * kMaxBins + b], __popcll(m));
        base = __shfl(base, leader);
        if (mine) L.items[(size_t(kind) * kMaxBins + b) * C + base + __popcll(m & ((1ull << lane) - 1))] = item;
        todo &= ~m;
    }
}
DEV int length_bin(int len) { return min(max(len, 0), kMaxBins - 1); }

// The bins of a list concatenated in bin order: its length, and item w of it.
// Consecutive items share a bin (a path length) except at bin boundaries.
DEV int list_length(const Lists &L, int kind) {
    int n = 0;
    for (int b = 0; b < kMaxBins; b++) n += L.count[kind * kMaxBins + b];
    return n;
}
DEV int list_item(const Lists &L, int kind, int C, int w) {
    for (int b = 0; b < kMaxBins; b++) {
        const int n = L.count[kind * kMaxBins + b];
        if (w < n) return L.items[(size_t(kind) * kMaxBins + b) * C + w];
        w -= n;
    }
    return -1;
}

// Stage-1 moves binned by the length of the path each evaluates (list 3); a
// small step's state is copied into its proposal slot here.
__global__ void __launch_bounds__(64) DRMLT_STAGE(k_b_stage1_bin)(Chains ch, DrParams P, const uint32_t *iterp, Lists L) {
    const uint32_t iter = *iterp;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P.chainCount; i += gridDim.x * blockDim.x) {
        const int chain = P.chainBase + i;
        Philox g(P.seed, iter, chain, TAG_S1);
        const Stage1 a = stage1_begin(ch, P, chain, g);
        const int cam = (!a.curValid || a.isLarge) ? a.cam : ch.c[a.pc];  // camera vertices of the path stage 1 evaluates
        list_push_wave(L, 3, ch.C, length_bin(stage1_length(ch, a)), chain, true, cam);
    }
}

// First chain (or list item) of this lane in the path-stage kernels, which
// take P.stageLanes chains per 64-lane workgroup (lanes at and above it take
// none): with fewer chains than the chip has SIMDs x 64, half-filled waves
// give each SIMD two waves whose memory waits overlap, and each walk round
// waits for the slowest of fewer lanes.
DEV int stage_first(const DrParams &P, int n) {
    return int(threadIdx.x) < P.stageLanes ? int(blockIdx.x) * P.stageLanes + int(threadIdx.x) : n;
}

// Stage 1 over every chain: from list 3 (binned) or in chain order.
DEV void stage1_kernel(const DevScene &Sc, const Chains &ch, const DrParams &P, const uint32_t *iterp, float *film, unsigned long long *stats,
                       const Lists &L) {
    const uint32_t iter = *iterp;
    const int n = P.binStages == 1 ? list_length(L, 3) : P.chainCount;
    for (int it = stage_first(P, n); it < n; it += gridDim.x * P.stageLanes) {
        const int chain = P.binStages == 1 ? list_item(L, 3, ch.C, it) : P.chainBase + it;
        Philox g(P.seed, iter, chain, TAG_S1);
        const Stage1 a = stage1_begin(ch, P, chain, g, P.binStages != 1);
        Walk w;
        stage1_walk(ch, P, a, g, w);
        POut<float> out;
        out.valid = false;
        walk_run(Sc, P.o, ch, P.discreteStdDev, w, !a.h2mcMove && !a.lensMove, out);
        const int flags = stage1_end(Sc, ch, P, chain, a, out, g, film, stats);
        // second stage: list 1 (every chain with one), list 0 (H2MC), list 4
        // (H2MC needing the current state's derivatives); list 5: lens mutations
        const bool second = (flags & 2) != 0, h2mc = second && (flags & 4);
        const int pc = a.pc, bin = length_bin(ch.c[pc] + ch.l[pc] - 1);
        list_push_wave(L, 1, ch.C, bin, chain, second, ch.c[pc]);
        list_push_wave(L, 0, ch.C, bin, chain, h2mc);
        list_push_wave(L, 4, ch.C, bin, chain,
                       h2mc && !ch.gInit[pc] && has_derivative(ch.c[pc], ch.l[pc], P.maxDervDepth) && double(ch.ss[pc]) > 1e-15, ch.c[pc]);
        list_push_wave(L, 5, ch.C, bin, chain, (flags & 128) != 0);
    }
}

#ifndef DRMLT_COUNT_WORK
// h2mc with lens mutations: the lens score (and lensVertexPos) of every
// current state that does not carry one yet, before stage 1 draws the lens
// coin.
__global__ void __launch_bounds__(64) DRMLT_STAGE(k_b_lens_score)(DevScene Sc, Chains ch, DrParams P) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P.chainCount; i += gridDim.x * blockDim.x)
        lens_score_ensure(Sc, ch, P, P.chainBase + i);
}
// The lens mutations of list 5, one per lane (each with its own ray queries).
__global__ void __launch_bounds__(64) DRMLT_STAGE(k_b_lens)(DevScene Sc, Chains ch, DrParams P, const uint32_t *iterp, float *film,
                                                             unsigned long long *stats, Lists L) {
    const uint32_t iter = *iterp;
    const int n = list_length(L, 5);
    for (int it = blockIdx.x * blockDim.x + threadIdx.x; it < n; it += gridDim.x * blockDim.x)
        lens_move(Sc, ch, P, list_item(L, 5, ch.C, it), iter, film, stats);
}
#endif

// Gaussian work of one list item (a chain). kind 0 (which 0): the current
// state's Gaussian unless it is cached, then the proposal offset. kind 2
// (which 1): the proposal's Gaussian, then (finishRatio) the H2MC ratio.
template <class X, class W>
DEV void gauss_item(const X &x, const DevScene &Sc, const Chains &ch, const DrParams &P, int chain, int which, int finishRatio, uint32_t iter,
                    W &w) {
    const int perm = ch.perm[chain];
    const int p = chain + ch.C * (which == 0 ? perm_cur(perm) : perm_second(perm));
    const int D = dim_of(ch.c[p], ch.l[p]);
    if (D > W::kMaxD) return;  // not reached: items are routed by their path-length bin
    const bool cached = which == 0 && ch.gInit[p] != 0;
    x.sync();
    if (!cached) {
        gaussian_from_derivs(x, ch, P, chain, p, w);
        // gradient and upper Hessian read (float); V, p, mean, logDet written (double)
        if (x.lane == 0) count_bytes(Sc, (unsigned long long)(4 * (D + D * (D + 1) / 2) + 8 * (D * D + 2 * D + 1)));
    }
    if (which == 0) h2mc_offset(x, ch, P, chain, iter, w);
    if (finishRatio) h2mc_ratio(x, ch, chain, w);
}

// Bins 0..kGSmallBins-1 of a list hold paths of length <= 4 (D = 2 max(b, 2)
// <= kGSmallD).
constexpr int kGSmallBins = kGSmallD / 2 + 1;
DEV int list_prefix(const Lists &L, int kind, int nbins) {
    int n = 0;
    for (int b = 0; b < nbins; b++) n += L.count[kind * kMaxBins + b];
    return n;
}

// Items [begin, end) of list `kind`, 64 / G per wave, a group of G lanes and
// the work area w[group] each.
template <int G, class W>
DEV void gauss_groups(const DevScene &Sc, const Chains &ch, const DrParams &P, const Lists &L, int kind, int which, int finishRatio,
                      uint32_t iter, int begin, int end, W *w) {
    constexpr int NG = 64 / G;
    const GroupExec<G> x{int(threadIdx.x) % G};
    const int grp = threadIdx.x / G;
    for (int it = begin + blockIdx.x * NG + grp; it < end; it += gridDim.x * NG)
        gauss_item(x, Sc, ch, P, list_item(L, kind, ch.C, it), which, finishRatio, iter, w[grp]);
}

// One-wave workgroups over a Gaussian list: first the small Gaussians
// (D <= 8), 64 / DRMLT_GAUSS_GROUP per wave, a group of lanes and an LDS work
// area each (a whole wave would leave most lanes idle on such matrices);
// then the larger ones, one wave each. Every executor gives the same results.
__global__ void __launch_bounds__(64, DRMLT_GAUSS_WAVES) DRMLT_STAGE(k_b_gauss)(DevScene Sc, Chains ch, DrParams P, Lists L, int kind, int which,
                                                             int finishRatio, const uint32_t *iterp) {
    constexpr int G = DRMLT_GAUSS_GROUP, NG = 64 / G;
    union Work {
        GWork big;
        GWorkT<kGSmallD> small[NG];
    };
    __shared__ Work w;
    const uint32_t iter = *iterp;
    const int n = list_length(L, kind), nSmall = list_prefix(L, kind, kGSmallBins);
    gauss_groups<G>(Sc, ch, P, L, kind, which, finishRatio, iter, 0, nSmall, w.small);
    __syncthreads();
    const WaveExec x{int(threadIdx.x)};
    for (int it = nSmall + blockIdx.x; it < n; it += gridDim.x)
        gauss_item(x, Sc, ch, P, list_item(L, kind, ch.C, it), which, finishRatio, iter, w.big);
}
// Second-stage proposals: list 1 (binned), or every chain in chain order with
// the lanes of chains without a second stage idle in the walk rounds.
DEV void stage2_propose_kernel(const DevScene &Sc, const Chains &ch, const DrParams &P, const uint32_t *iterp, const Lists &L) {
    const uint32_t iter = *iterp;
    const int n = P.binStages ? list_length(L, 1) : P.chainCount;
    for (int it = stage_first(P, n); it < n; it += gridDim.x * P.stageLanes) {
        const int chain = P.binStages ? list_item(L, 1, ch.C, it) : P.chainBase + it;
        const bool act = P.binStages || (ch.flags[chain] & 2) != 0;
        Philox g(P.seed, iter, chain, TAG_S2);
        Stage2 a;
        Walk w;
        if (act) {
            a = stage2_begin(ch, chain, g);
            stage2_walk(ch, P, chain, a, g, w);
        }
        POut<float> out;
        out.valid = false;
        walk_run(Sc, P.o, ch, P.discreteStdDev, w, act, out);
        int bin = 0, flags = 0, cam = 0;
        if (act) {
            stage2_end(Sc, ch, chain, a, out);
            flags = ch.flags[chain];
            bin = length_bin(ch.c[a.pc] + ch.l[a.pc] - 1);
            cam = ch.c[a.pc];
        }
        list_push_wave(L, 2, ch.C, bin, chain, (flags & 4) && (flags & 16), cam);
    }
}

// Second-stage acceptance over the same chains; Green's y* walks where they
// are needed.
DEV void stage2_accept_kernel(const DevScene &Sc, const Chains &ch, const DrParams &P, const uint32_t *iterp, float *film,
                              unsigned long long *stats, const Lists &L) {
    const uint32_t iter = *iterp;
    const int n = P.binStages ? list_length(L, 1) : P.chainCount;
    for (int it = stage_first(P, n); it < n; it += gridDim.x * P.stageLanes) {
        const int chain = P.binStages ? list_item(L, 1, ch.C, it) : P.chainBase + it;
        const bool act = P.binStages || (ch.flags[chain] & 2) != 0;
        Philox g(P.seed, iter, chain, TAG_S3);
        Stage3 a;
        Walk w;
        a.traceStar = false;
        if (act) {
            a = stage3_begin(ch, P, chain);
            if (a.traceStar) stage3_walk(ch, P, a, g, w);
        }
        POut<float> out;
        out.valid = false;
        walk_run(Sc, P.o, ch, P.discreteStdDev, w, a.traceStar, out);
        if (act) stage3_end(Sc, ch, P, chain, a, out, g, film, stats);
    }
}

// The path-stage kernels at 2 (the names without suffix) and 3 resident
// waves per SIMD (suffix _w3); the counting build has the first only.
#define DRMLT_PATH_STAGE_KERNELS(WAVES, SFX)                                                                                          \
    __global__ void __launch_bounds__(64, WAVES) DRMLT_STAGE(k_b_stage1##SFX)(DevScene Sc, Chains ch, DrParams P, const uint32_t *iterp, \
                                                                              float *film, unsigned long long *stats, Lists L) {      \
        stage1_kernel(Sc, ch, P, iterp, film, stats, L);                                                                             \
    }                                                                                                                                 \
    __global__ void __launch_bounds__(64, WAVES) DRMLT_STAGE(k_b_stage2_propose##SFX)(DevScene Sc, Chains ch, DrParams P,               \
                                                                                      const uint32_t *iterp, Lists L) {              \
        stage2_propose_kernel(Sc, ch, P, iterp, L);                                                                                  \
    }                                                                                                                                 \
    __global__ void __launch_bounds__(64, WAVES) DRMLT_STAGE(k_b_stage2_accept##SFX)(DevScene Sc, Chains ch, DrParams P, uint32_t *iterp, \
                                                                                     float *film, unsigned long long *stats, Lists L) { \
        stage2_accept_kernel(Sc, ch, P, iterp, film, stats, L);                                                                      \
    }
DRMLT_PATH_STAGE_KERNELS(2, )
#ifndef DRMLT_COUNT_WORK
DRMLT_PATH_STAGE_KERNELS(3, _w3)
#endif
#undef DRMLT_PATH_STAGE_KERNELS

// Derivative passes over list `kind` (4 with which 0: the current states that
// need them; 2 with which 1: the proposals), persistent over (chain, pass)
// work items bin by bin.
__global__ void __launch_bounds__(64, DRMLT_DERIV_WAVES) DRMLT_STAGE(k_b_deriv)(DevScene Sc, Chains ch, DrParams P, Lists L, int kind, int which) {
    // bin b holds paths of length b (D = 2 max(b, 2)); its work items are
    // (chain, pass) pairs, all bins concatenated
    long total = 0;
    for (int b = 0; b < kMaxBins; b++) {
        const int D = 2 * (b > 2 ? b : 2);
        if (D <= kGMax) total += long(L.count[kind * kMaxBins + b]) * jet_passes(D);
    }
    const long gsz = long(gridDim.x) * blockDim.x;
    for (long w = long(blockIdx.x) * blockDim.x + threadIdx.x; w < total; w += gsz) {
        long r = w;
        int bin = 0, D = 0, np = 0;
        for (; bin < kMaxBins; bin++) {
            D = 2 * (bin > 2 ? bin : 2);
            np = D <= kGMax ? jet_passes(D) : 0;
            const long nb = long(L.count[kind * kMaxBins + bin]) * np;
            if (r < nb) break;
            r -= nb;
        }
        const int chain = L.items[(size_t(kind) * kMaxBins + bin) * ch.C + r / np];
        const int q = int(r % np);
        const int perm = ch.perm[chain];
        const int p = chain + ch.C * (which == 0 ? perm_cur(perm) : perm_second(perm));
        if (!has_derivative(ch.c[p], ch.l[p], P.maxDervDepth) || !(double(ch.ss[p]) > 1e-15)) continue;
        if (which == 0 && ch.gInit[p]) continue;  // cached Gaussian
        int row, col0;
        jet_pass_index(D, q, row, col0);
        float *gr = ch.grad + size_t(chain) * kGMax, *hs = ch.hess + size_t(chain) * kGMax * kGMax;
        const bool ok = jet_pass<float>(Sc, P.oDeriv, slot_ref(ch, p), D, row, col0,
                                        [&](int cc, float gv) { if (row == 0) gr[cc] = gv; },
                                        [&](int cc, float hv) { if (cc >= row) hs[row * kGMax + cc] = hv; });
        if (!ok) gr[0] = __builtin_nanf("");
        // state read by the pass, gradient (row 0) and Hessian values written
        count_bytes(Sc, (unsigned long long)(state_bytes(ch.c[p], ch.l[p]) + 4 * kJetK * (row == 0 ? 2 : 1)));
    }
}

#undef DRMLT_STAGE

}  // namespace drmlt

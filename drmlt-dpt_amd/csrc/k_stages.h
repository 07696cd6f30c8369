// Kernels of the batched DR step (one per stage, persistent over work lists
// binned by path length). Included twice: by k_loop.hip as the product
// kernels and by k_count.hip, with DRMLT_COUNT_WORK defined, as the counting
// kernels (suffix _counted) behind drmlt_count_work.
#pragma once
#include "dr_stages.h"

namespace drmlt {

// Minimum waves per SIMD the path-stage kernels are compiled for (bounds
// their registers); more resident waves hide traversal latency.
#ifndef DRMLT_STAGE_WAVES
#define DRMLT_STAGE_WAVES 2
#endif

#ifdef DRMLT_COUNT_WORK
#define DRMLT_STAGE(name) name##_counted
#else
#define DRMLT_STAGE(name) name
#endif

// ------------------------------------------------------------ batched form
// Work lists, binned by path length: kind 0 = H2MC chains (current Gaussian
// and offset), kind 2 = H2MC chains with a valid proposal (its Gaussian and
// the ratio). The stage kernels proper run over all chains in index order.
DEV void list_push(const Lists &L, int kind, int C, int bin, int item) {
    const int slot = atomicAdd(&L.count[kind * kMaxBins + bin], 1);
    L.items[(size_t(kind) * kMaxBins + bin) * C + slot] = item;
}
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

__global__ void __launch_bounds__(64, DRMLT_STAGE_WAVES) DRMLT_STAGE(k_b_stage1)(DevScene Sc, Chains ch, DrParams P, const uint32_t *iterp, float *film,
                                                 unsigned long long *stats, Lists L) {
    const uint32_t iter = *iterp;
    for (int chain = blockIdx.x * blockDim.x + threadIdx.x; chain < ch.C; chain += gridDim.x * blockDim.x) {
        Philox g(P.seed, iter, chain, TAG_S1);
        const Stage1 a = stage1_begin(ch, P, chain, g);
        Walk w;
        stage1_walk(ch, P, a, g, w);
        POut<float> out;
        walk_run(Sc, P.o, w, true, out);
        const int flags = stage1_end(Sc, ch, P, chain, a, out, g, film, stats);
        if (flags & 2) {
            const int perm = ch.perm[chain];
            const int pc = chain + ch.C * perm_cur(perm);
            const int bin = min(ch.c[pc] + ch.l[pc] - 1, kMaxBins - 1);
            if (flags & 4) list_push(L, 0, ch.C, bin, chain);
        }
    }
}

// One wave per Gaussian (one-wave workgroups, the work matrices in LDS).
// kind 0 (which 0): every H2MC chain; the current state's Gaussian unless it
// is cached, then the proposal offset. kind 2 (which 1): the proposal's
// Gaussian, then (finishRatio) the H2MC ratio.
__global__ void __launch_bounds__(64) DRMLT_STAGE(k_b_gauss)(DevScene Sc, Chains ch, DrParams P, Lists L, int kind, int which,
                                                             int finishRatio, const uint32_t *iterp) {
    __shared__ GWork w;
    const WaveExec x{int(threadIdx.x)};
    const uint32_t iter = *iterp;
    const int n = list_length(L, kind);
    {
        for (int it = blockIdx.x; it < n; it += gridDim.x) {
            const int chain = list_item(L, kind, ch.C, it);
            const int perm = ch.perm[chain];
            const int p = chain + ch.C * (which == 0 ? perm_cur(perm) : perm_second(perm));
            const bool cached = which == 0 && ch.gInit[p] != 0;
            x.sync();
            if (!cached) {
                gaussian_from_derivs(x, ch, P, chain, p, w);
                if (x.lane == 0) {
                    // gradient and upper Hessian read (float); V, p, mean, logDet written (double)
                    const int D = dim_of(ch.c[p], ch.l[p]);
                    count_bytes(Sc, (unsigned long long)(4 * (D + D * (D + 1) / 2) + 8 * (D * D + 2 * D + 1)));
                }
            }
            if (which == 0) h2mc_offset(x, ch, P, chain, iter, w);
            if (finishRatio) h2mc_ratio(x, ch, chain, w);
        }
    }
}
__global__ void __launch_bounds__(64, DRMLT_STAGE_WAVES) DRMLT_STAGE(k_b_stage2_propose)(DevScene Sc, Chains ch, DrParams P, const uint32_t *iterp, Lists L) {
    const uint32_t iter = *iterp;
    const int gsz = gridDim.x * blockDim.x;
    // chains in index order (coalesced state access); those without a second stage idle
    for (int chain = blockIdx.x * blockDim.x + threadIdx.x; chain < ch.C; chain += gsz) {
        // every lane of the wave takes part in the walk rounds; those without
        // a second stage have no walk
        const bool act = (ch.flags[chain] & 2) != 0;
        Philox g(P.seed, iter, chain, TAG_S2);
        Stage2 a;
        Walk w;
        if (act) {
            a = stage2_begin(ch, chain, g);
            stage2_walk(ch, P, chain, a, g, w);
        }
        POut<float> out;
        out.valid = false;
        walk_run(Sc, P.o, w, act, out);
        if (!act) continue;
        stage2_end(Sc, ch, chain, a, out);
        const int flags = ch.flags[chain];
        if ((flags & 4) && (flags & 16)) list_push(L, 2, ch.C, min(ch.c[a.pc] + ch.l[a.pc] - 1, kMaxBins - 1), chain);
    }
}

__global__ void __launch_bounds__(64, DRMLT_STAGE_WAVES) DRMLT_STAGE(k_b_stage2_accept)(DevScene Sc, Chains ch, DrParams P, uint32_t *iterp, float *film,
                                                        unsigned long long *stats, Lists L) {
    const uint32_t iter = *iterp;
    const int gsz = gridDim.x * blockDim.x;
    for (int chain = blockIdx.x * blockDim.x + threadIdx.x; chain < ch.C; chain += gsz) {
        const bool act = (ch.flags[chain] & 2) != 0;
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
        walk_run(Sc, P.o, w, act && a.traceStar, out);
        if (act) stage3_end(Sc, ch, P, chain, a, out, g, film, stats);
    }
}

// Derivative passes over list `kind` (slot: 0 current, 2 second), persistent
// over (chain, pass) work items bin by bin.
__global__ void __launch_bounds__(64) DRMLT_STAGE(k_b_deriv)(DevScene Sc, Chains ch, DrParams P, Lists L, int kind, int which) {
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

// Kernels of the batched DR step (one per stage, persistent over work lists
// binned by path length). Included twice: by k_loop.hip as the product
// kernels and by k_count.hip, with DRMLT_COUNT_WORK defined, as the counting
// kernels (suffix _counted) behind drmlt_count_work.
#pragma once
#include "dr_stages.h"

namespace drmlt {

#ifdef DRMLT_COUNT_WORK
#define DRMLT_STAGE(name) name##_counted
#else
#define DRMLT_STAGE(name) name
#endif

// ------------------------------------------------------------ batched form
// List kinds: 0 = needs current Gaussian, 1 = second stage (aniso or iso),
// 2 = needs proposal Gaussian. Each is binned by path length.
DEV void list_push(const Lists &L, int kind, int C, int bin, int item) {
    const int slot = atomicAdd(&L.count[kind * kMaxBins + bin], 1);
    L.items[(size_t(kind) * kMaxBins + bin) * C + slot] = item;
}

__global__ void __launch_bounds__(64) DRMLT_STAGE(k_b_stage1)(DevScene Sc, Chains ch, DrParams P, const uint32_t *iterp, float *film,
                                                 unsigned long long *stats, Lists L) {
    const uint32_t iter = *iterp;
    for (int chain = blockIdx.x * blockDim.x + threadIdx.x; chain < ch.C; chain += gridDim.x * blockDim.x) {
        const int flags = dr_stage1(Sc, ch, P, chain, iter, film, stats);
        if (flags & 2) {
            const int perm = ch.perm[chain];
            const int pc = chain + ch.C * perm_cur(perm);
            const int bin = min(ch.c[pc] + ch.l[pc] - 1, kMaxBins - 1);
            list_push(L, 1, ch.C, bin, chain);
            if ((flags & 4) && !ch.gInit[pc]) list_push(L, 0, ch.C, bin, chain);
        }
    }
}

__global__ void __launch_bounds__(64) DRMLT_STAGE(k_b_gauss)(DevScene Sc, Chains ch, DrParams P, Lists L, int kind, int which, int finishRatio) {
    const int gsz = gridDim.x * blockDim.x;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    for (int bin = 0; bin < kMaxBins; bin++) {
        const int n = L.count[kind * kMaxBins + bin];
        const int *items = L.items + (size_t(kind) * kMaxBins + bin) * ch.C;
        for (int w = tid; w < n; w += gsz) {
            const int chain = items[w];
            const int perm = ch.perm[chain];
            const int p = chain + ch.C * (which == 0 ? perm_cur(perm) : perm_second(perm));
            const int c = ch.c[p], l = ch.l[p], D = dim_of(c, l);
            const GRef G = g_ref(ch, p);
            const int C = ch.C;
            if (!has_derivative(c, l, P.maxDervDepth) || D > kGMax) {
                iso_gaussian(D, P.h2.sigma, G);
            } else {
                const double ssv = ch.ss[p];
                const float *gr = ch.grad + chain, *hs = ch.hess + chain;
                bool fin = ssv > 1e-15;
                if (fin) {
                    for (int i = 0; i < D && fin; i++) fin = isfinite(gr[i * C]);
                    for (int i = 0; i < D && fin; i++)
                        for (int j = i; j < D && fin; j++) fin = isfinite(hs[(i * kGMax + j) * C]);
                }
                if (fin) {
                    compute_gaussian(P.h2, ssv, D, [&](int i) { return double(gr[i * C]); },
                                     [&](int i, int j) { return double(hs[(i * kGMax + j) * C]); }, G);
                } else {
                    compute_gaussian(P.h2, ssv, D, [&](int) { return 0.0; }, [&](int, int) { return 0.0; }, G);
                }
            }
            ch.gInit[p] = 1;
            // gradient and upper Hessian read (float), mean, covL, invCov, logDet written (double)
            count_bytes(Sc, (unsigned long long)(4 * (D + D * (D + 1) / 2) + 8 * (D + 2 * D * D + 1)));
            if (finishRatio) dr_stage2_ratio(ch, chain);
        }
    }
}

__global__ void __launch_bounds__(64) DRMLT_STAGE(k_b_stage2_propose)(DevScene Sc, Chains ch, DrParams P, const uint32_t *iterp, Lists L) {
    const uint32_t iter = *iterp;
    const int gsz = gridDim.x * blockDim.x;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    for (int bin = 0; bin < kMaxBins; bin++) {
        const int n = L.count[1 * kMaxBins + bin];
        const int *items = L.items + (size_t(1) * kMaxBins + bin) * ch.C;
        for (int w = tid; w < n; w += gsz) {
            const int chain = items[w];
            dr_stage2_propose(Sc, ch, P, chain, iter);
            const int flags = ch.flags[chain];
            if ((flags & 4) && (flags & 16)) list_push(L, 2, ch.C, bin, chain);
        }
    }
}

__global__ void __launch_bounds__(64) DRMLT_STAGE(k_b_stage2_accept)(DevScene Sc, Chains ch, DrParams P, uint32_t *iterp, float *film,
                                                        unsigned long long *stats, Lists L) {
    const uint32_t iter = *iterp;
    const int gsz = gridDim.x * blockDim.x;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    for (int bin = 0; bin < kMaxBins; bin++) {
        const int n = L.count[1 * kMaxBins + bin];
        const int *items = L.items + (size_t(1) * kMaxBins + bin) * ch.C;
        for (int w = tid; w < n; w += gsz) dr_stage2_accept(Sc, ch, P, items[w], iter, film, stats);
    }
}

// Derivative passes over list `kind` (slot: 0 current, 2 second), persistent
// over (chain, pass) work items bin by bin.
__global__ void __launch_bounds__(64) DRMLT_STAGE(k_b_deriv)(DevScene Sc, Chains ch, DrParams P, Lists L, int kind, int which) {
    const int gsz = gridDim.x * blockDim.x;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    for (int bin = 0; bin < kMaxBins; bin++) {
        const int n = L.count[kind * kMaxBins + bin];
        if (n == 0) continue;
        const int *items = L.items + (size_t(kind) * kMaxBins + bin) * ch.C;
        const int chain0 = items[0];
        const int perm0 = ch.perm[chain0];
        const int p0 = chain0 + ch.C * (which == 0 ? perm_cur(perm0) : perm_second(perm0));
        const int c0 = ch.c[p0], l0 = ch.l[p0];
        (void)l0;
        // every chain in a bin has the same path length, hence the same D
        const int D = dim_of(c0, ch.l[p0]);
        if (D > kGMax) continue;
        const int np = jet_passes(D);
        const long total = long(n) * np;
        for (long w = tid; w < total; w += gsz) {
            const int chain = items[w / np];
            const int q = int(w % np);
            const int perm = ch.perm[chain];
            const int p = chain + ch.C * (which == 0 ? perm_cur(perm) : perm_second(perm));
            if (!has_derivative(ch.c[p], ch.l[p], P.maxDervDepth) || !(double(ch.ss[p]) > 1e-15)) continue;
            int row, col0;
            jet_pass_index(D, q, row, col0);
            float *gr = ch.grad + chain, *hs = ch.hess + chain;
            const int C = ch.C;
            const bool ok = jet_pass<float>(Sc, P.oDeriv, slot_ref(ch, p), D, row, col0,
                                            [&](int cc, float gv) { if (row == 0) gr[cc * C] = gv; },
                                            [&](int cc, float hv) { if (cc >= row) hs[(row * kGMax + cc) * C] = hv; });
            if (!ok) gr[0] = __builtin_nanf("");
            // state read by the pass, gradient (row 0) and Hessian values written
            count_bytes(Sc, (unsigned long long)(state_bytes(ch.c[p], ch.l[p]) + 4 * kJetK * (row == 0 ? 2 : 1)));
        }
    }
}

#undef DRMLT_STAGE

}  // namespace drmlt

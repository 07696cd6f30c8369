// Derivative kernels (FP32 jets): the batched pipeline's per-(chain, pass) work items and the parity surface.
#include "dr_stages.h"

namespace drmlt {

// Derivative passes over list `kind` (slot: 0 current, 2 second), persistent
// over (chain, pass) work items bin by bin.
__global__ void __launch_bounds__(64) k_b_deriv(DevScene Sc, Chains ch, DrParams P, Lists L, int kind, int which) {
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
        }
    }
}

// Gradient / Hessian of log f at each record, one thread per (record, pass).
template <class S>
DEV void gradhess_rec(int w, const DevScene &Sc, const Opts &o, int n, char *recs, const int *passOff, int totalPasses, S *grad, S *hess, int *okOut) {
    if (w >= totalPasses) return;
    int lo = 0, hi = n;  // record with passOff[r] <= w < passOff[r+1]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (passOff[mid] <= w) lo = mid; else hi = mid;
    }
    const int r = lo, q = w - passOff[r];
    const SRef ref = rec_ref(recs + size_t(r) * kRecBytes);
    const int D = dim_of(*ref.c, *ref.l);
    int row, col0;
    jet_pass_index(D, q, row, col0);
    const bool ok = jet_pass<S>(Sc, o, ref, D, row, col0,
                                [&](int cc, S gv) { if (row == 0) grad[size_t(r) * 32 + cc] = gv; },
                                [&](int cc, S hv) { if (cc >= row) hess[(size_t(r) * 32 + row) * 32 + cc] = hv; });
    if (!ok) okOut[r] = 0;
}

__global__ void __launch_bounds__(64) k_gradhess_rec_f32(DevScene Sc, Opts o, int n, char *recs, const int *passOff, int totalPasses, float *grad, float *hess,
                                   int *okOut) {
    gradhess_rec<float>(blockIdx.x * blockDim.x + threadIdx.x, Sc, o, n, recs, passOff, totalPasses, grad, hess, okOut);
}

}  // namespace drmlt

// Batched DR pipeline (persistent per-stage kernels over path-length bins), pre-pass, direct pass and the primal parity surface.
#include "dr_stages.h"
#include "kernels_decl.h"

namespace drmlt {

// ------------------------------------------------------------ batched form
// List kinds: 0 = needs current Gaussian, 1 = second stage (aniso or iso),
// 2 = needs proposal Gaussian. Each is binned by path length.
DEV void list_push(const Lists &L, int kind, int C, int bin, int item) {
    const int slot = atomicAdd(&L.count[kind * kMaxBins + bin], 1);
    L.items[(size_t(kind) * kMaxBins + bin) * C + slot] = item;
}

__global__ void __launch_bounds__(64) k_b_stage1(DevScene Sc, Chains ch, DrParams P, const uint32_t *iterp, float *film,
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

__global__ void __launch_bounds__(64) k_b_gauss(DevScene Sc, Chains ch, DrParams P, Lists L, int kind, int which, int finishRatio) {
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
            if (finishRatio) dr_stage2_ratio(ch, chain);
        }
    }
}

__global__ void __launch_bounds__(64) k_b_stage2_propose(DevScene Sc, Chains ch, DrParams P, const uint32_t *iterp, Lists L) {
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

__global__ void __launch_bounds__(64) k_b_stage2_accept(DevScene Sc, Chains ch, DrParams P, uint32_t *iterp, float *film,
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

// Clears the list counters for the next step and advances the iteration
// counter (a fill kernel, not a memset, so the step can be graph-captured).
__global__ void __launch_bounds__(64) k_b_reset(Lists L, uint32_t *iterp, int advance) {
    const int t = threadIdx.x;
    if (t < 3 * kMaxBins) L.count[t] = 0;
    if (t == 0 && advance) *iterp += 1;
}

// ------------------------------------------------------------ pre-pass
// GeneratePathBidir per sample (delayed.cpp:65-118): up to kPreMaxC
// contributions per sample; lengthContrib and the score total accumulate in
// double.
struct PreSink {
    int n;
    int *cl;
    float *ls;
    double *lenSum;
    double *total;
    DEV void operator()(const POut<float> &o, int c, int l) {
        if (n < kPreMaxC) {
            cl[n] = c | (l << 8);
            ls[n] = o.lsScore;
        }
        n++;
        const int len = c + l - 1;
        if (len < kMaxBins) atomicAdd(lenSum + len, double(o.lsScore));
        atomicAdd(total, double(o.lsScore));
    }
};
__global__ void __launch_bounds__(64) k_prepass_gen(DevScene Sc, Opts o, int numInit, int maxDepth, uint64_t seed, int *cnt, int *cl, float *ls, double *lenSum,
                              double *total) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= numInit) return;
    Philox g(seed, 0, i, TAG_PRE);
    PreSink sink{0, cl + size_t(i) * kPreMaxC, ls + size_t(i) * kPreMaxC, lenSum, total};
    generate_path_bidir<kPreMaxL>(Sc, o, 3, maxDepth, g, sink, 0, 0, (const SRef *)nullptr);
    cnt[i] = min(sink.n, kPreMaxC);
}
struct NullSink {
    DEV void operator()(const POut<float> &, int, int) {}
};
struct KeepSink {
    int c, l;
    POut<float> kept;
    DEV void operator()(const POut<float> &o, int cc, int ll) { if (cc == c && ll == l) kept = o; }
};
// Seed replay (delayed.cpp:125-167): chain i regenerates sample sel[i] and
// keeps the strategy (c, l) as its current state.
__global__ void __launch_bounds__(64) k_prepass_seed(DevScene Sc, Opts o, Chains ch, int maxDepth, uint64_t seed, const int *selSample, const int *selCl) {
    const int chain = blockIdx.x * blockDim.x + threadIdx.x;
    if (chain >= ch.C) return;
    Philox g(seed, 0, selSample[chain], TAG_PRE);
    KeepSink sink;
    sink.c = selCl[chain] & 255;
    sink.l = selCl[chain] >> 8;
    sink.kept.valid = false;
    const int p = chain;  // physical slot 0 is current
    const SRef ref = slot_ref(ch, p);
    generate_path_bidir<kPreMaxL>(Sc, o, 3, maxDepth, g, sink, sink.c, sink.l, &ref);
    const bool found = ch.c[p] != 0 && sink.kept.valid;
    if (!found) {
        // replay did not reproduce the strategy: leave a well-formed, invalid state
        ch.c[p] = 2; ch.l[p] = 2; ch.light[p] = 0; ch.lprim[p] = Sc.lightTris[0];
        for (int k = 0; k < kVMax; k++) { ch.cprim[k * ch.NS + p] = 0; ch.lgprim[k * ch.NS + p] = 0; }
        sink.kept.valid = false;
        sink.kept.contrib = V3<float>(0, 0, 0);
        sink.kept.lsScore = sink.kept.ssScore = 0.0f;
        sink.kept.screen[0] = sink.kept.screen[1] = 0.5f;
    }
    store_contrib(ch, p, sink.kept);
    ch.valid[p] = found ? 1 : 0;
    ch.hasSplat[p] = 0;  // seeds carry no splat list
    ch.gInit[p] = 0;
    for (int k = 1; k < 4; k++) {
        ch.valid[p + k * ch.C] = 0;
        ch.hasSplat[p + k * ch.C] = 0;
        ch.gInit[p + k * ch.C] = 0;
        ch.c[p + k * ch.C] = ch.c[p];
        ch.l[p + k * ch.C] = ch.l[p];
    }
    ch.perm[chain] = make_perm(0, 1, 2);
}

// ------------------------------------------------------------ direct pass
__global__ void __launch_bounds__(128) k_direct(DevScene Sc, Opts o, int spp, uint64_t seed, float *film) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    const long npix = long(Sc.cam.W) * Sc.cam.H;
    if (i >= npix * spp) return;
    const int pix = int(i % npix), s = int(i / npix);
    Philox g(seed, uint32_t(s), uint32_t(pix), TAG_DIRECT);
    generate_path_uni_direct(Sc, o, pix % Sc.cam.W, pix / Sc.cam.W, g, film);
}

// ------------------------------------------------------------ parity surface
__global__ void __launch_bounds__(128) k_trace(DevScene Sc, int n, const float *org, const float *dir, const float *tmin, const float *tmax, int any, int *prim,
                        float *t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float o[3] = {org[3 * i], org[3 * i + 1], org[3 * i + 2]}, d[3] = {dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]};
    if (any) {
        prim[i] = trace_any(Sc, o, d, tmin[i], tmax[i]) ? 1 : 0;
    } else {
        float tb;
        prim[i] = trace_closest(Sc, o, d, tmin[i], tmax[i], tb);
        t[i] = tb;
    }
}

__global__ void __launch_bounds__(64) k_perturb_rec(DevScene Sc, Opts o, int n, char *recs, const float *off, const float *disc, float *c8) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const SRef r = rec_ref(recs + size_t(i) * kRecBytes);
    PerturbU<OffArray> up{r.u, 1, 0, OffArray{off + 32 * i, 1}};
    POut<float> out;
    perturb_eval<float, false>(Sc, o, r, up, DiscArray{disc + 32 * i, 1}, out);
    if (out.valid) { r.screen[0] = out.screen[0]; r.screen[1] = out.screen[1]; }
    contrib8(c8 + 8 * i, out);
}

struct ListSink {
    float *cs;
    int maxc, n;
    DEV void operator()(const POut<float> &o, int c, int l) {
        if (n < maxc) {
            float *q = cs + size_t(n) * 10;
            q[0] = float(c); q[1] = float(l); q[2] = o.contrib.x; q[3] = o.contrib.y; q[4] = o.contrib.z;
            q[5] = o.lsScore; q[6] = o.ssScore; q[7] = o.screen[0]; q[8] = o.screen[1]; q[9] = o.mis;
        }
        n++;
    }
};
__global__ void __launch_bounds__(64) k_genbidir_rec(DevScene Sc, Opts o, int n, int minDepth, int maxDepth, const float *uni, int nuni, int maxc, float *cs, int *nc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    GenArray g{uni + size_t(i) * nuni, nuni, 0};
    ListSink sink{cs + size_t(i) * maxc * 10, maxc, 0};
    generate_path_bidir<kPreMaxL>(Sc, o, minDepth, maxDepth, g, sink, 0, 0, (const SRef *)nullptr);
    nc[i] = sink.n;
}

__global__ void __launch_bounds__(64) k_gauss_rec(int n, H2Param h2, const int *Ds, const double *sc, const double *grad, const double *hess, double *mean,
                            double *covL, double *invCov, double *logDet) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int D = Ds[i];
    GRef G;
    G.mean = mean; G.covL = covL; G.invCov = invCov; G.logDet = logDet; G.gs = n; G.k = i;
    compute_gaussian(h2, sc[i], D, [&](int k) { return grad[size_t(i) * 32 + k]; },
                     [&](int a, int b) { return hess[(size_t(i) * 32 + a) * 32 + b]; }, G);
}

__global__ void __launch_bounds__(64) k_logpdf_rec(int n, const int *Ds, const double *x, double sign, const double *mean, const double *covL, const double *invCov,
                             const double *logDet, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    GRef G;
    G.mean = const_cast<double *>(mean); G.covL = const_cast<double *>(covL); G.invCov = const_cast<double *>(invCov);
    G.logDet = const_cast<double *>(logDet); G.gs = n; G.k = i;
    out[i] = gaussian_logpdf(Ds[i], [&](int k) { return x[size_t(i) * 32 + k]; }, sign, G);
}

// Second-stage acceptance arithmetic alone (delayed.cpp:466-541), batched:
// in[i] = (type, a1, a2, aStar, transRatio, secondValid); out = alpha2.
__global__ void __launch_bounds__(64) k_alpha2(int n, const double *in, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *q = in + 6 * size_t(i);
    const int type = int(q[0]);
    const float a1 = float(q[1]);
    double a2 = q[2];
    if (is_invalid(a2)) { out[i] = 0.0; return; }
    if (type == 0) {
        float aStar = q[5] != 0.0 ? float(q[3]) : 0.0f;
        if (is_invalid(aStar)) aStar = 0.0f;
        out[i] = aStar == 1.0f ? 0.0 : fmin(1.0, a2 * double(1.0f - aStar) / double(1.0f - a1));
    } else {
        float aStar = float(q[3]);
        if (is_invalid(aStar)) aStar = 1.0f;
        if (aStar == 1.0f) { out[i] = 0.0; return; }
        const double tr = q[4];
        out[i] = is_invalid(tr) ? 0.0 : fmin(1.0, a2 * tr * double(1.0f - aStar) / double(1.0f - a1));
    }
}

}  // namespace drmlt

// Batched DR pipeline (persistent per-stage kernels over path-length bins), pre-pass, direct pass and the primal parity surface.
#include "k_stages.h"
#include "kernels_decl.h"

namespace drmlt {

// Clears the list counters for the next step and advances the iteration
// counter (a fill kernel, not a memset, so the step can be graph-captured).
__global__ void __launch_bounds__(64) k_b_reset(Lists L, uint32_t *iterp, int advance) {
    const int t = threadIdx.x;
    for (int k = t; k < kListKinds * kMaxBins; k += blockDim.x) L.count[k] = 0;
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
        for (int k = 0; k < kVMax; k++) { ref.cprim[k * ref.vs] = 0; ref.lgprim[k * ref.vs] = 0; }
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
__global__ void __launch_bounds__(64) k_direct(DevScene Sc, Opts o, int spp, uint64_t seed, float *film) {
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    const long npix = long(Sc.cam.W) * Sc.cam.H;
    if (i >= npix * spp) return;
    const int pix = int(i % npix), s = int(i / npix);
    Philox g(seed, uint32_t(s), uint32_t(pix), TAG_DIRECT);
    generate_path_uni_direct(Sc, o, pix % Sc.cam.W, pix / Sc.cam.W, g, film);
}

// ------------------------------------------------------------ parity surface
__global__ void __launch_bounds__(64) k_trace(DevScene Sc, int n, const float *org, const float *dir, const float *tmin, const float *tmax, int any, int *prim,
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

// ComputeGaussian of one record per one-wave workgroup, the same code the
// batched step runs; slots [n][kGSlot].
__global__ void __launch_bounds__(64) k_gauss_rec(int n, H2Param h2, const int *Ds, const double *sc, const double *grad, const double *hess,
                                                  double *slots) {
    __shared__ GWork w;
    const WaveExec x{int(threadIdx.x)};
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        x.sync();
        gauss_compute(x, h2, sc[i], Ds[i], [&](int k) { return grad[size_t(i) * 32 + k]; },
                      [&](int a, int b) { return hess[(size_t(i) * 32 + a) * 32 + b]; }, w, GSlot{slots + size_t(i) * kGSlot});
    }
}
// ComputeGaussian, GaussianLogPdf of sign * x and GenerateSample from the
// standard normals nrm with the loop's own device functions (gauss_compute,
// gauss_logpdf, gauss_sample: what k_b_gauss, h2mc_offset and h2mc_ratio run)
// for the records idx[0..nidx): small = 1 takes them in groups of
// DRMLT_GAUSS_GROUP lanes with a work area each, as k_b_gauss takes the
// Gaussians of dimension <= kGSmallD; small = 0 one wave each.
template <class X, class W>
DEV void gauss_eval_item(const X &x, const H2Param &h2, int i, const int *Ds, const double *sc, const double *grad, const double *hess,
                         const double *xv, double sign, const double *nrm, double *slots, double *logpdf, double *sample, W &w) {
    const int D = Ds[i];
    const GSlot G{slots + size_t(i) * kGSlot};
    x.sync();
    gauss_compute(x, h2, sc[i], D, [&](int k) { return grad[size_t(i) * 32 + k]; },
                  [&](int a, int b) { return hess[(size_t(i) * 32 + a) * 32 + b]; }, w, G);
    const double lp = gauss_logpdf(x, D, [&](int k) { return xv[size_t(i) * 32 + k]; }, sign, G, w);
    if (x.lane == 0) logpdf[i] = lp;
    gauss_sample(x, D, G, [&](int j) { return nrm[size_t(i) * 32 + j]; }, w, [&](int k, double v) { sample[size_t(i) * 32 + k] = v; });
}
__global__ void __launch_bounds__(64) k_gauss_eval_rec(int nidx, const int *idx, int small, H2Param h2, const int *Ds, const double *sc,
                                                       const double *grad, const double *hess, const double *xv, double sign,
                                                       const double *nrm, double *slots, double *logpdf, double *sample) {
    constexpr int G = DRMLT_GAUSS_GROUP, NG = 64 / G;
    union Work {
        GWork big;
        GWorkT<kGSmallD> small[NG];
    };
    __shared__ Work w;
    if (small) {
        const GroupExec<G> x{int(threadIdx.x) % G};
        const int grp = threadIdx.x / G;
        for (int it = blockIdx.x * NG + grp; it < nidx; it += gridDim.x * NG)
            gauss_eval_item(x, h2, idx[it], Ds, sc, grad, hess, xv, sign, nrm, slots, logpdf, sample, w.small[grp]);
    } else {
        const WaveExec x{int(threadIdx.x)};
        for (int it = blockIdx.x; it < nidx; it += gridDim.x)
            gauss_eval_item(x, h2, idx[it], Ds, sc, grad, hess, xv, sign, nrm, slots, logpdf, sample, w.big);
    }
}

// The reference's representation of a factored Gaussian: mean (n x 32),
// covL = V diag(p^-1/2) and invCov = V diag(p) V^T (n x 24 x 24), logDet.
__global__ void __launch_bounds__(64) k_gauss_expand(int n, const int *Ds, const double *slots, double *mean, double *covL, double *invCov,
                                                     double *logDet) {
    const long t = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= long(n) * kGMax * kGMax) return;
    const int r = int(t / (kGMax * kGMax)), e = int(t % (kGMax * kGMax)), i = e / kGMax, j = e % kGMax;
    const int D = Ds[r];
    const GSlot g{const_cast<double *>(slots) + size_t(r) * kGSlot};
    double L = 0, Q = 0;
    if (i < D && j < D) {
        L = g.V(i, j) / sqrt(g.p(j));
        for (int k = 0; k < D; k++) Q += g.V(i, k) * g.p(k) * g.V(j, k);
    }
    covL[t] = L;
    invCov[t] = Q;
    if (j == 0) mean[size_t(r) * 32 + i] = i < D ? g.m(i) : 0.0;
    if (e == 0) {
        logDet[r] = g.ld();
        for (int k = kGMax; k < 32; k++) mean[size_t(r) * 32 + k] = 0.0;
    }
}

// GaussianLogPdf (gaussian.cpp:18-26) over the reference's representation:
// mean (n x 32), invCov (n x 24 x 24), logDet, of sign * x.
__global__ void __launch_bounds__(64) k_logpdf_rec(int n, const int *Ds, const double *x, double sign, const double *mean,
                                                   const double *invCov, const double *logDet, double *out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int D = Ds[r];
    const double *m = mean + size_t(r) * 32, *xv = x + size_t(r) * 32, *Q = invCov + size_t(r) * kGMax * kGMax;
    double q = 0;
    for (int i = 0; i < D; i++) {
        double si = 0;
        const double di = sign * xv[i] - m[i];
        for (int j = 0; j < D; j++) si += Q[i * kGMax + j] * (sign * xv[j] - m[j]);
        q += di * si;
    }
    out[r] = double(D) * (-0.5 * log(2.0 * 3.14159265358979323846)) + 0.5 * logDet[r] - 0.5 * q;
}

// The loop's second-stage acceptance (dr_alpha2, the function stage3_end
// calls), batched: in[i] = (type, a1, a2, aStar, transRatio, secondValid);
// out = alpha2. For Green & Mira secondValid = 0 means y* was not traced
// (aStar = 0), as in stage3_end.
__global__ void __launch_bounds__(64) k_alpha2(int n, const double *in, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *q = in + 6 * size_t(i);
    const int type = int(q[0]);
    const float aStar = (type == 0 && q[5] == 0.0) ? 0.0f : float(q[3]);
    bool draw;
    out[i] = dr_alpha2(0, type, float(q[1]), q[2], aStar, q[4], draw);
}

}  // namespace drmlt

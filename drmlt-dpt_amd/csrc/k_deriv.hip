// FP32 derivative kernel of the parity surface (one thread per record and pass).
#include "dr_stages.h"

namespace drmlt {

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

// log f at each record with its triangles fixed (the function the derivative
// passes differentiate), in T; ok[i] = 0 where it has no contribution.
template <class T>
DEV void logf_rec(int i, const DevScene &Sc, const Opts &o, int n, const char *recs, double *logf, int *okOut) {
    if (i >= n) return;
    const SRef ref = rec_ref(const_cast<char *>(recs) + size_t(i) * kRecBytes);
    PlainU<T> up{ref.u, ref.s, 0};
    POut<T> out;
    perturb_eval<T, true>(Sc, o, ref, up, DiscNone{}, out);
    const bool ok = out.valid && out.ssScore > T(0);
    logf[i] = ok ? double(log(out.ssScore)) : 0.0;
    okOut[i] = ok ? 1 : 0;
}
__global__ void __launch_bounds__(64) k_logf_rec_f32(DevScene Sc, Opts o, int n, const char *recs, double *logf, int *okOut) {
    logf_rec<float>(blockIdx.x * blockDim.x + threadIdx.x, Sc, o, n, recs, logf, okOut);
}
__global__ void __launch_bounds__(64) k_logf_rec_f64(DevScene Sc, Opts o, int n, const char *recs, double *logf, int *okOut) {
    logf_rec<double>(blockIdx.x * blockDim.x + threadIdx.x, Sc, o, n, recs, logf, okOut);
}

// The lens bookkeeping of each record -- lens score and lensVertexPos of its
// fixed re-evaluation (lens_state_eval) -- and the gradient and Hessian of
// the lens function at screen position scr[2i], scr[2i + 1] (lens_jets):
// out9[9i..] = lensScore, lensPos (3), d/ds0, d/ds1, H00, H01, H11;
// ok[i] = 0 where the lens function has no contribution there.
__global__ void __launch_bounds__(64) k_lens_rec(DevScene Sc, Opts o, int n, char *recs, const float *scr, const float *lpIn, float *out9,
                                                 int *okOut) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const SRef ref = rec_ref(recs + size_t(i) * kRecBytes);
    V3<float> lp;
    const float ls = lens_state_eval(Sc, o, ref, lp);
    float g[2] = {0.0f, 0.0f}, h[3] = {0.0f, 0.0f, 0.0f};
    const float s2[2] = {scr[2 * i], scr[2 * i + 1]};
    const V3<float> lpJ = lpIn ? V3<float>(lpIn[3 * i], lpIn[3 * i + 1], lpIn[3 * i + 2]) : lp;
    const bool ok = lens_jets(Sc, o, ref, s2, lpJ, g, h);
    float *q = out9 + 9 * size_t(i);
    q[0] = ls; q[1] = lp.x; q[2] = lp.y; q[3] = lp.z;
    q[4] = g[0]; q[5] = g[1]; q[6] = h[0]; q[7] = h[1]; q[8] = h[2];
    okOut[i] = ok ? 1 : 0;
}

__global__ void __launch_bounds__(64) k_gradhess_rec_f32(DevScene Sc, Opts o, int n, char *recs, const int *passOff, int totalPasses, float *grad, float *hess,
                                   int *okOut) {
    gradhess_rec<float>(blockIdx.x * blockDim.x + threadIdx.x, Sc, o, n, recs, passOff, totalPasses, grad, hess, okOut);
}

}  // namespace drmlt

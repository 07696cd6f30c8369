// FP64 H2MC Gaussian on the device: ComputeGaussian (drsmallstep.cpp:18-93,
// 219-266), IsotropicGaussian / GaussianLogPdf / GenerateSample
// (gaussian.cpp:4-47).
//
// A Gaussian is stored factored, one contiguous slot of kGSlot doubles per
// state: the eigenvectors V of the Hessian (column k = k-th eigenvector), the
// precisions p along them, the mean and logDet. The reference's matrices are
// products of these: invCov = V diag(p) V^T, covL = V diag(p^-1/2)
// (drsmallstep.cpp:70-92).
//
// Every routine is written once over an executor X that names the lanes
// working on one Gaussian: SeqExec (one thread does everything, the plain
// form), WaveExec (the 64 lanes of a one-wave workgroup share the work, the
// matrices in LDS) or GroupExec (a group of lanes of one wave, several
// Gaussians per wave). Work is split by index, every element is computed by the
// same expression in the same order in both, and reductions run in index
// order on one lane, so both executors produce bit-identical results.
// Floating-point contraction is off in these routines for the same reason.
//
// The eigen-decomposition is a parallel-order cyclic Jacobi iteration
// (Golub & Van Loan, Matrix Computations, 4th ed., 8.5.4-8.5.5: round-robin
// ordering; each round applies D/2 disjoint rotations, rows then columns).
#pragma once
#include "dev_math.h"
#include <cmath>

namespace drmlt {

constexpr int kGMax = 24;          // largest D with derivatives (path length 12)
constexpr int kGSlot = 640;        // doubles per Gaussian slot (V, p, mean, logDet, padding to 64 B)

// A Gaussian of dimension up to DM stored at b: V (DM x DM), p, mean, logDet.
template <int DM> struct GSlotT {
    double *b;
    DEV double &V(int i, int k) const { return b[i * DM + k]; }
    DEV double &p(int k) const { return b[DM * DM + k]; }
    DEV double &m(int i) const { return b[DM * DM + DM + i]; }
    DEV double &ld() const { return b[DM * DM + 2 * DM]; }
};
using GSlot = GSlotT<kGMax>;

// drsmallstep.h:26-42; posOffset is 0.5 * (e^L + e^-L - 1) as computed there
struct H2Param {
    double sigma, L, posScale, posOffset, negScale, negOffset;
};
inline H2Param make_h2param(double sigma) {
    H2Param p;
    p.sigma = sigma;
    p.L = M_PI / 2.0;
    p.posScale = 0.5 * (std::exp(p.L) - std::exp(-p.L)) * 0.5 * (std::exp(p.L) - std::exp(-p.L));
    p.posOffset = 0.5 * (std::exp(p.L) + std::exp(-p.L) - 1.0);
    p.negScale = std::sin(p.L) * std::sin(p.L);
    p.negOffset = -(std::cos(p.L) - 1.0);
    return p;
}

struct SeqExec {
    int lane = 0;
    static constexpr int n = 1;
    DEV void sync() const {}
};
struct WaveExec {
    int lane;
    static constexpr int n = 64;
    DEV void sync() const { __syncthreads(); }
};
// G consecutive lanes of a one-wave workgroup; the groups of a wave work on
// different Gaussians and may diverge, so sync() is no barrier: the fences
// complete this lane's memory operations and keep the compiler from moving
// memory accesses across it, which is all lanes of one wave need to see
// each other's writes.
template <int G> struct GroupExec {
    int lane;  // lane within the group
    static constexpr int n = G;
    DEV void sync() const {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
};

// Scratch of one Gaussian computation of dimension up to DM (private for
// SeqExec, LDS otherwise). Work matrices have row stride kS = DM + 1 (odd for
// even DM, so rows fall on different LDS banks); the stride changes no
// arithmetic, so every size gives the same results.
template <int DM> struct GWorkT {
    static constexpr int kS = DM + 1;
    static constexpr int kMaxD = DM;
    double A[DM * kS], V[DM * kS];
    double vec[DM], vec2[DM], part[DM];
    double rc[DM / 2], rs[DM / 2];
    int rp[DM / 2], rq[DM / 2];
    int flag;
};
using GWork = GWorkT<kGMax>;

// Pair k of round r of the round-robin schedule over D (even) indices:
// position 0 stays, positions 1..D-1 rotate.
DEV void rr_pair(int D, int r, int k, int &p, int &q) {
    auto pos = [&](int i) { return i == 0 ? 0 : 1 + (i - 1 + r) % (D - 1); };
    const int a = pos(k), b = pos(D - 1 - k);
    p = a < b ? a : b;
    q = a < b ? b : a;
}

// gaussian.cpp:4-16 in factored form: V = I, p = 1/sigma^2, mean 0.
template <class X, class GS> DEV void gauss_iso(const X &x, int D, double sigma, const GS &g) {
#pragma clang fp contract(off)
    for (int t = x.lane; t < D * D; t += x.n) g.V(t / D, t % D) = (t / D == t % D) ? 1.0 : 0.0;
    for (int i = x.lane; i < D; i += x.n) {
        g.p(i) = 1.0 / (sigma * sigma);
        g.m(i) = 0.0;
    }
    if (x.lane == 0) {
        double ld = 0;
        for (int i = 0; i < D; i++) ld += log(1.0 / (sigma * sigma));
        g.ld() = ld;
    }
}

// Eigen-decomposition of w.A (D x D symmetric, row stride W::kS) into w.V; the
// eigenvalues end on the diagonal of w.A. Converged when the off-diagonal
// mass is below 1e-30 of the diagonal mass.
template <class X, class W> DEV void gauss_eigen(const X &x, int D, W &w) {
#pragma clang fp contract(off)
    constexpr int kGS = W::kS;
    for (int t = x.lane; t < D * D; t += x.n) w.V[(t / D) * kGS + t % D] = (t / D == t % D) ? 1.0 : 0.0;
    const int half = D / 2;
    for (int sweep = 0; sweep < 60; sweep++) {
        x.sync();
        for (int i = x.lane; i < D; i += x.n) {
            double o = 0;
            for (int j = 0; j < D; j++)
                if (j != i) o += w.A[i * kGS + j] * w.A[i * kGS + j];
            w.part[i] = o;
            w.vec[i] = w.A[i * kGS + i] * w.A[i * kGS + i];
        }
        x.sync();
        if (x.lane == 0) {
            double off = 0, diag = 0;
            for (int i = 0; i < D; i++) { off += w.part[i]; diag += w.vec[i]; }
            w.flag = (off <= 1e-30 * diag || off == 0) ? 1 : 0;
        }
        x.sync();
        if (w.flag) break;
        for (int r = 0; r < D - 1; r++) {
            for (int k = x.lane; k < half; k += x.n) {
                int p, q;
                rr_pair(D, r, k, p, q);
                const double apq = w.A[p * kGS + q];
                double c = 1.0, s = 0.0;
                if (apq != 0) {
                    const double app = w.A[p * kGS + p], aqq = w.A[q * kGS + q];
                    const double theta = (aqq - app) / (2 * apq);
                    const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
                    c = 1 / sqrt(t * t + 1);
                    s = t * c;
                }
                w.rp[k] = p; w.rq[k] = q; w.rc[k] = c; w.rs[k] = s;
            }
            x.sync();
            for (int t = x.lane; t < half * D; t += x.n) {
                const int k = t / D, j = t % D, p = w.rp[k], q = w.rq[k];
                const double c = w.rc[k], s = w.rs[k];
                const double a = w.A[p * kGS + j], b = w.A[q * kGS + j];
                w.A[p * kGS + j] = c * a - s * b;
                w.A[q * kGS + j] = s * a + c * b;
            }
            x.sync();
            for (int t = x.lane; t < half * D; t += x.n) {
                const int k = t / D, i = t % D, p = w.rp[k], q = w.rq[k];
                const double c = w.rc[k], s = w.rs[k];
                const double a = w.A[i * kGS + p], b = w.A[i * kGS + q];
                w.A[i * kGS + p] = c * a - s * b;
                w.A[i * kGS + q] = s * a + c * b;
                const double va = w.V[i * kGS + p], vb = w.V[i * kGS + q];
                w.V[i * kGS + p] = c * va - s * vb;
                w.V[i * kGS + q] = s * va + c * vb;
            }
            x.sync();
        }
    }
}

// drsmallstep.cpp:219-266 + 18-93. grad(i) and hess(i, j) for i <= j (the
// upper triangle of the row-major Hessian; Eigen reads the lower triangle of
// the column-major view of the same buffer). Writes the Gaussian to g.
template <class X, class GP, class HP, class W, class GS>
DEV void gauss_compute(const X &x, const H2Param &pr, double sc, int D, GP grad, HP hess, W &w, const GS &g) {
#pragma clang fp contract(off)
    constexpr int kGS = W::kS;
    for (int t = x.lane; t < D * D; t += x.n) {
        const int i = t / D, j = t % D;
        w.A[i * kGS + j] = i <= j ? hess(i, j) : hess(j, i);
    }
    for (int i = x.lane; i < D; i += x.n) {
        double fr = 0;
        for (int j = i; j < D; j++) { const double h = hess(i, j); fr += (i == j ? 1.0 : 2.0) * h * h; }
        w.part[i] = fr;
    }
    x.sync();
    if (x.lane == 0) {
        double fro = 0;
        for (int i = 0; i < D; i++) fro += w.part[i];
        fro = sqrt(fro);
        w.flag = (sc <= 1e-15 || fro < 0.5 / (pr.sigma * pr.sigma)) ? 1 : 0;
    }
    x.sync();
    if (w.flag) {
        gauss_iso(x, D, pr.sigma, g);
        x.sync();
        return;
    }
    gauss_eigen(x, D, w);
    const double invS2 = 1.0 / (pr.sigma * pr.sigma);
    for (int i = x.lane; i < D; i += x.n) {
        const double lam = w.A[i * kGS + i];
        double e = fabs(lam) > 1e-10 ? 1.0 / fabs(lam) : 0.0;
        double vg = 0;
        for (int k = 0; k < D; k++) vg += w.V[k * kGS + i] * grad(k);
        const double off = e * vg;
        double s2, o;
        if (fabs(lam) > 1e-10) {
            o = off;
            if (lam > 0) { s2 = pr.posScale; o *= pr.posOffset; }
            else { s2 = pr.negScale; o *= pr.negOffset; }
        } else {
            s2 = pr.L * pr.L;
            o = 0.5 * off * pr.L * pr.L;
        }
        e *= s2;
        e = e > 1e-10 ? 1.0 / e : 0.0;
        const double p = e + invS2;
        w.vec[i] = p;
        w.vec2[i] = (e / p) * o;
        g.p(i) = p;
    }
    x.sync();
    for (int i = x.lane; i < D; i += x.n) {
        double mi = 0;
        for (int k = 0; k < D; k++) mi += w.V[i * kGS + k] * w.vec2[k];
        g.m(i) = mi;
    }
    for (int t = x.lane; t < D * D; t += x.n) g.V(t / D, t % D) = w.V[(t / D) * kGS + t % D];
    if (x.lane == 0) {
        double ld = 0;
        for (int i = 0; i < D; i++) ld += log(w.vec[i]);
        g.ld() = ld;
    }
    x.sync();
}

// GenerateSample (gaussian.cpp:36-47): x = mean + covL n with
// covL = V diag(p^-1/2); nrm(j) supplies the j-th standard normal.
template <class X, class GS, class NP, class W, class OW>
DEV void gauss_sample(const X &x, int D, const GS &g, NP nrm, W &w, OW out) {
#pragma clang fp contract(off)
    for (int j = x.lane; j < D; j += x.n) w.vec[j] = nrm(j) / sqrt(g.p(j));
    x.sync();
    for (int i = x.lane; i < D; i += x.n) {
        double xi = g.m(i);
        for (int j = 0; j < D; j++) xi += g.V(i, j) * w.vec[j];
        out(i, xi);
    }
    x.sync();
}

// GaussianLogPdf (gaussian.cpp:18-26) of sign * xv(i):
// -D/2 log 2pi + logDet/2 - 1/2 sum_k p_k (V^T (x - mean))_k^2.
template <class X, class XP, class W, class GS>
DEV double gauss_logpdf(const X &x, int D, XP xv, double sign, const GS &g, W &w) {
#pragma clang fp contract(off)
    for (int i = x.lane; i < D; i += x.n) w.vec[i] = sign * xv(i) - g.m(i);
    x.sync();
    for (int k = x.lane; k < D; k += x.n) {
        double s = 0;
        for (int i = 0; i < D; i++) s += g.V(i, k) * w.vec[i];
        w.vec2[k] = g.p(k) * s * s;
    }
    x.sync();
    double q = 0;
    for (int k = 0; k < D; k++) q += w.vec2[k];
    const double r = double(D) * (-0.5 * log(2.0 * 3.14159265358979323846)) + 0.5 * g.ld() - 0.5 * q;
    x.sync();
    return r;
}

}  // namespace drmlt

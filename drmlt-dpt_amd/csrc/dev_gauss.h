// FP64 H2MC Gaussian on the device: ComputeGaussian (drsmallstep.cpp:18-93,
// 219-266), IsotropicGaussian / GaussianLogPdf / GenerateSample
// (gaussian.cpp:4-47). The symmetric eigen-decomposition is a cyclic Jacobi
// iteration (Golub & Van Loan 8.5) run by one thread per matrix in double.
// Gaussian fields live in HBM as structure-of-arrays: element e of slot k at
// [e * gs + k] (gs = number of Gaussian slots), so a wave's stores coalesce.
#pragma once
#include "dev_math.h"
#include <cmath>

namespace drmlt {

constexpr int kGMax = 24;  // largest D with derivatives (path length 12)

struct GRef {
    double *mean, *covL, *invCov, *logDet;  // [e * gs + k]
    long gs;
    long k;
    DEV double &m(int i) const { return mean[i * gs + k]; }
    DEV double &L(int i, int j) const { return covL[(i * kGMax + j) * gs + k]; }
    DEV double &Q(int i, int j) const { return invCov[(i * kGMax + j) * gs + k]; }
    DEV double &ld() const { return logDet[k]; }
};

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

// gaussian.cpp:4-16
DEV void iso_gaussian(int D, double sigma, const GRef &g) {
    double ld = 0;
    for (int i = 0; i < D; i++) {
        g.m(i) = 0;
        for (int j = 0; j < D; j++) {
            g.L(i, j) = i == j ? sigma : 0.0;
            g.Q(i, j) = i == j ? 1.0 / (sigma * sigma) : 0.0;
        }
        ld += log(1.0 / (sigma * sigma));
    }
    g.ld() = ld;
}

// Cyclic Jacobi on A (D x D, row-major, private), eigenvectors in V columns.
DEV void jacobi(int D, double *A, double *V) {
    for (int i = 0; i < D * D; i++) V[i] = 0;
    for (int i = 0; i < D; i++) V[i * D + i] = 1;
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0, diag = 0;
        for (int i = 0; i < D; i++)
            for (int j = 0; j < D; j++) {
                const double a = A[i * D + j];
                if (i == j) diag += a * a; else off += a * a;
            }
        if (off <= 1e-30 * diag || off == 0) break;
        for (int p = 0; p < D - 1; p++)
            for (int q = p + 1; q < D; q++) {
                const double apq = A[p * D + q];
                if (apq == 0) continue;
                const double app = A[p * D + p], aqq = A[q * D + q];
                const double theta = (aqq - app) / (2 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
                const double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < D; k++) {
                    const double akp = A[k * D + p], akq = A[k * D + q];
                    A[k * D + p] = c * akp - s * akq;
                    A[k * D + q] = s * akp + c * akq;
                }
                for (int k = 0; k < D; k++) {
                    const double apk = A[p * D + k], aqk = A[q * D + k];
                    A[p * D + k] = c * apk - s * aqk;
                    A[q * D + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < D; k++) {
                    const double vkp = V[k * D + p], vkq = V[k * D + q];
                    V[k * D + p] = c * vkp - s * vkq;
                    V[k * D + q] = s * vkp + c * vkq;
                }
            }
    }
}

// drsmallstep.cpp:219-266 + 18-93. grad[i * gst], hess[(i * hld + j) * hst]
// (row-major); the eigen-solver reads the upper triangle, as Eigen reads the
// lower triangle of the column-major view of the same buffer.
template <class GP, class HP>
DEV void compute_gaussian(const H2Param &pr, double sc, int D, GP grad, HP hess, const GRef &g) {
    // Frobenius norm of the symmetric Hessian from its upper triangle
    // (hess(i, j) is only read for i <= j)
    double fro = 0;
    for (int i = 0; i < D; i++)
        for (int j = i; j < D; j++) { const double h = hess(i, j); fro += (i == j ? 1.0 : 2.0) * h * h; }
    fro = sqrt(fro);
    if (sc <= 1e-15 || fro < 0.5 / (pr.sigma * pr.sigma)) {
        iso_gaussian(D, pr.sigma, g);
        return;
    }
    double A[kGMax * kGMax], V[kGMax * kGMax];
    for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++) A[i * D + j] = i <= j ? hess(i, j) : hess(j, i);
    jacobi(D, A, V);
    double e[kGMax], off[kGMax], p[kGMax], lam[kGMax];
    const double invS2 = 1.0 / (pr.sigma * pr.sigma);
    for (int i = 0; i < D; i++) {
        lam[i] = A[i * D + i];
        e[i] = fabs(lam[i]) > 1e-10 ? 1.0 / fabs(lam[i]) : 0.0;
        double vg = 0;
        for (int k = 0; k < D; k++) vg += V[k * D + i] * grad(k);
        off[i] = e[i] * vg;
    }
    double ld = 0;
    for (int i = 0; i < D; i++) {
        double s2, o;
        if (fabs(lam[i]) > 1e-10) {
            o = off[i];
            if (lam[i] > 0) { s2 = pr.posScale; o *= pr.posOffset; }
            else { s2 = pr.negScale; o *= pr.negOffset; }
        } else {
            s2 = pr.L * pr.L;
            o = 0.5 * off[i] * pr.L * pr.L;
        }
        e[i] *= s2;
        e[i] = e[i] > 1e-10 ? 1.0 / e[i] : 0.0;
        off[i] = o;
        p[i] = e[i] + invS2;
        ld += log(p[i]);
    }
    for (int i = 0; i < D; i++) {
        double mi = 0;
        for (int k = 0; k < D; k++) mi += V[i * D + k] * (e[k] / p[k]) * off[k];
        g.m(i) = mi;
        for (int j = 0; j < D; j++) {
            double q = 0;
            for (int k = 0; k < D; k++) q += V[i * D + k] * p[k] * V[j * D + k];
            g.Q(i, j) = q;
            g.L(i, j) = V[i * D + j] / sqrt(p[j]);
        }
    }
    g.ld() = ld;
}

// gaussian.cpp:18-26 (x given by functor, sign applied to x)
template <class XP> DEV double gaussian_logpdf(int D, XP x, double sign, const GRef &g) {
    double q = 0;
    for (int i = 0; i < D; i++) {
        double si = 0;
        const double di = sign * x(i) - g.m(i);
        for (int j = 0; j < D; j++) si += g.Q(i, j) * (sign * x(j) - g.m(j));
        q += di * si;
    }
    return double(D) * (-0.5 * log(2.0 * 3.14159265358979323846)) + 0.5 * g.ld() - 0.5 * q;
}

}  // namespace drmlt

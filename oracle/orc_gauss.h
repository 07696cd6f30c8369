// ORACLE (test infrastructure only): H2MC Gaussian construction
// (drsmallstep.cpp:18-93, 219-266) and Gaussian helpers (gaussian.cpp:4-47),
// in double. The symmetric eigen-decomposition is a cyclic Jacobi solver
// (Golub & Van Loan, Matrix Computations, section 8.5) standing in for
// Eigen's SelfAdjointEigenSolver; only similarity-invariant outputs (mean,
// inverse covariance, covariance, log-determinant) are compared, since
// eigenvector sign and order are not unique.
#pragma once
#include <cmath>
#include <vector>
#include <algorithm>

namespace orc {

// A is n x n row-major symmetric; on return A's diagonal holds eigenvalues and
// V (row-major, V[i*n+k] = component i of eigenvector k) the eigenvectors.
inline void JacobiEigen(int n, double *A, double *V, int maxSweeps = 100) {
    for (int i = 0; i < n * n; i++) V[i] = 0;
    for (int i = 0; i < n; i++) V[i * n + i] = 1;
    for (int sweep = 0; sweep < maxSweeps; sweep++) {
        double off = 0, diag = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) (i == j ? diag : off) += A[i * n + j] * A[i * n + j];
        if (off <= 1e-30 * diag || off == 0) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double apq = A[p * n + q];
                if (apq == 0) continue;
                double app = A[p * n + p], aqq = A[q * n + q];
                double theta = (aqq - app) / (2 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
                double c = 1 / std::sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; k++) {
                    double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
}

struct Gauss {
    int D = 0;
    std::vector<double> mean, covL, invCov;  // covL, invCov row-major D x D
    double logDet = 0;
};

// drsmallstep.h:26-42 (Param). posOffsetFactor is 0.5*(e^L + e^-L - 1) as the
// code computes it (its comment says cosh(L) - 1).
struct H2Param {
    double sigma, L, posScale, posOffset, negScale, negOffset;
    explicit H2Param(double s = 0.01, double L_ = M_PI / 2) : sigma(s), L(L_) {
        posScale = 0.5 * (std::exp(L) - std::exp(-L)) * 0.5 * (std::exp(L) - std::exp(-L));
        posOffset = 0.5 * (std::exp(L) + std::exp(-L) - 1.0);
        negScale = std::sin(L) * std::sin(L);
        negOffset = -(std::cos(L) - 1.0);
    }
};

// gaussian.cpp:4-16
inline void IsotropicGaussian(int D, double sigma, Gauss &g) {
    g.D = D;
    g.mean.assign(D, 0.0);
    g.covL.assign(D * D, 0.0);
    g.invCov.assign(D * D, 0.0);
    g.logDet = 0;
    for (int i = 0; i < D; i++) {
        g.covL[i * D + i] = sigma;
        g.invCov[i * D + i] = 1.0 / (sigma * sigma);
        g.logDet += std::log(1.0 / (sigma * sigma));
    }
}

// drsmallstep.cpp:219-266 + 18-93. hess is the row-major D x D output of the
// derivative function; the eigen-solver reads the upper triangle (Eigen maps
// the row-major buffer column-major and reads its lower triangle).
inline void ComputeGaussian(const H2Param &pr, double sc, int D, const double *grad, const double *hess, Gauss &g) {
    double fro = 0;
    for (int i = 0; i < D * D; i++) fro += hess[i] * hess[i];
    fro = std::sqrt(fro);
    if (sc <= 1e-15 || fro < 0.5 / (pr.sigma * pr.sigma)) {
        IsotropicGaussian(D, pr.sigma, g);
        return;
    }
    std::vector<double> A(D * D), V(D * D);
    for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++) A[i * D + j] = hess[std::min(i, j) * D + std::max(i, j)];
    JacobiEigen(D, A.data(), V.data());
    std::vector<double> lam(D), e(D), off(D), p(D);
    for (int i = 0; i < D; i++) lam[i] = A[i * D + i];
    const double invS2 = 1.0 / (pr.sigma * pr.sigma);
    for (int i = 0; i < D; i++) e[i] = std::fabs(lam[i]) > 1e-10 ? 1.0 / std::fabs(lam[i]) : 0.0;
    for (int i = 0; i < D; i++) {
        double vg = 0;
        for (int k = 0; k < D; k++) vg += V[k * D + i] * grad[k];
        off[i] = e[i] * vg;
    }
    for (int i = 0; i < D; i++) {
        double s2 = 1, o = 0;
        if (std::fabs(lam[i]) > 1e-10) {
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
    }
    g.D = D;
    g.mean.assign(D, 0.0);
    g.covL.assign(D * D, 0.0);
    g.invCov.assign(D * D, 0.0);
    for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++) {
            double s = 0;
            for (int k = 0; k < D; k++) s += V[i * D + k] * p[k] * V[j * D + k];
            g.invCov[i * D + j] = s;
            g.covL[i * D + j] = V[i * D + j] / std::sqrt(p[j]);
        }
    for (int i = 0; i < D; i++) {
        double s = 0;
        for (int k = 0; k < D; k++) s += V[i * D + k] * (e[k] / p[k]) * off[k];
        g.mean[i] = s;
    }
    g.logDet = 0;
    for (int i = 0; i < D; i++) g.logDet += std::log(p[i]);
}

// gaussian.cpp:18-26
inline double GaussianLogPdf(const double *x, const Gauss &g) {
    const int D = g.D;
    double q = 0;
    for (int i = 0; i < D; i++) {
        double s = 0;
        for (int j = 0; j < D; j++) s += g.invCov[i * D + j] * (x[j] - g.mean[j]);
        q += (x[i] - g.mean[i]) * s;
    }
    return double(D) * (-0.5 * std::log(2.0 * M_PI)) + 0.5 * g.logDet - 0.5 * q;
}
// gaussian.cpp:36-47 with the normals supplied by the caller
inline void GenerateSample(const Gauss &g, const double *n, double *x) {
    for (int i = 0; i < g.D; i++) {
        double s = g.mean[i];
        for (int j = 0; j < g.D; j++) s += g.covL[i * g.D + j] * n[j];
        x[i] = s;
    }
}

}  // namespace orc

"""Pins the oracle's H2MC Gaussian (drsmallstep.cpp:18-93, 219-266) and
Gaussian helpers (gaussian.cpp:4-47) against numpy / scipy: an independent
symmetric eigen-solver, the multivariate normal log-density, and the
invariants the construction must satisfy."""
import numpy as np
import pytest
from scipy.stats import multivariate_normal

import orc

SIG = 0.01
L = np.pi / 2
POS_SCALE = np.sinh(L) ** 2
POS_OFF = 0.5 * (np.exp(L) + np.exp(-L) - 1.0)  # drsmallstep.h:31 as computed


def _sym(rng, D, scale, neg_frac=0.3):
    Q, _ = np.linalg.qr(rng.normal(size=(D, D)))
    lam = rng.uniform(0.2, 2.0, D) * scale
    lam[rng.uniform(size=D) < neg_frac] *= -1
    return (Q * lam) @ Q.T


def _pad(g, H):
    D = len(g)
    G = np.zeros(32); G[:D] = g
    HH = np.zeros((32, 32)); HH[:D, :D] = H
    return G, HH


@pytest.mark.parametrize("D", [4, 6, 10, 18, 24])
def test_jacobi_matches_numpy_eigh(D):
    rng = np.random.default_rng(D)
    A = _sym(rng, D, 50.0)
    lam, V = orc.jacobi_eigen(A)
    np.testing.assert_allclose(np.sort(lam), np.linalg.eigvalsh(A), rtol=1e-10, atol=1e-9)
    np.testing.assert_allclose(V @ np.diag(lam) @ V.T, A, atol=1e-9)
    np.testing.assert_allclose(V.T @ V, np.eye(D), atol=1e-12)


@pytest.mark.parametrize("D", [4, 8, 14, 20, 24])
def test_gaussian_against_eigh_construction(D):
    """mean / invCov / logDet from numpy's eigh with the H2MC bounds."""
    rng = np.random.default_rng(100 + D)
    H = _sym(rng, D, 3e4)  # well above the isotropic threshold 0.5 / sigma^2 in Frobenius norm
    g = rng.normal(size=D) * 50
    G, HH = _pad(g, H)
    mean, covL, inv, ld = orc.compute_gaussian([D], [1.0], SIG, G[None], HH[None])
    lam, V = np.linalg.eigh(H)
    e = np.where(np.abs(lam) > 1e-10, 1 / np.abs(lam), 0.0)
    gp = e * (V.T @ g)
    s2 = np.where(lam > 0, POS_SCALE, 1.0)
    o = gp * np.where(lam > 0, POS_OFF, 1.0 - np.cos(L))
    e = 1.0 / (e * s2)
    p = e + 1 / SIG ** 2
    np.testing.assert_allclose(inv[0, :D, :D], (V * p) @ V.T, rtol=1e-9, atol=1e-6)
    np.testing.assert_allclose(mean[0, :D], V @ (e / p * o), rtol=1e-8, atol=1e-12)
    np.testing.assert_allclose(ld[0], np.sum(np.log(p)), rtol=1e-12)
    # invariants: covL covL^T is the covariance, logDet = log det(invCov)
    Lm = covL[0, :D, :D]
    np.testing.assert_allclose(Lm @ Lm.T @ inv[0, :D, :D], np.eye(D), atol=1e-8)
    np.testing.assert_allclose(ld[0], np.linalg.slogdet(inv[0, :D, :D])[1], rtol=1e-10)


def test_isotropic_branches():
    D = 8
    rng = np.random.default_rng(3)
    small = _sym(rng, D, 1.0)  # ||H||_F < 0.5 / sigma^2 = 5000: isotropic (drsmallstep.cpp:244-252)
    big = _sym(rng, D, 1e4)
    g = rng.normal(size=D)
    for H, sc, iso in ((small, 1.0, True), (big, 1e-16, True), (big, 1.0, False)):
        G, HH = _pad(g, H)
        mean, covL, inv, ld = orc.compute_gaussian([D], [sc], SIG, G[None], HH[None])
        if iso:
            np.testing.assert_allclose(mean[0, :D], 0)
            np.testing.assert_allclose(inv[0, :D, :D], np.eye(D) / SIG ** 2)
            np.testing.assert_allclose(covL[0, :D, :D], np.eye(D) * SIG)
            np.testing.assert_allclose(ld[0], D * np.log(1 / SIG ** 2))
        else:
            assert not np.allclose(inv[0, :D, :D], np.eye(D) / SIG ** 2)


def test_zero_eigenvalue_branch():
    """|lambda| <= 1e-10 uses s2 = L^2 and o = g' L^2 / 2 with e = 0 -> p = 1/sigma^2."""
    D = 6
    rng = np.random.default_rng(4)
    Q, _ = np.linalg.qr(rng.normal(size=(D, D)))
    lam = np.array([4e4, -3e4, 2e4, 0.0, 5e4, -6e4])
    H = (Q * lam) @ Q.T
    H = 0.5 * (H + H.T)
    G, HH = _pad(rng.normal(size=D), H)
    mean, covL, inv, ld = orc.compute_gaussian([D], [1.0], SIG, G[None], HH[None])
    # along the null eigenvector the precision is the prior's
    v0 = Q[:, 3]
    assert v0 @ inv[0, :D, :D] @ v0 == pytest.approx(1 / SIG ** 2, rel=1e-6)


@pytest.mark.parametrize("D", [3, 10, 24])
def test_logpdf_matches_scipy(D):
    rng = np.random.default_rng(200 + D)
    A = rng.normal(size=(D, D))
    cov = A @ A.T + D * np.eye(D)
    inv = np.linalg.inv(cov)
    mean = rng.normal(size=D)
    inv32 = np.zeros((32, 32)); inv32[:D, :D] = inv
    for _ in range(5):
        x = rng.normal(size=D) * 2
        got = orc.gaussian_logpdf(D, x, mean, inv32, -np.linalg.slogdet(cov)[1])
        assert got == pytest.approx(multivariate_normal(mean, cov).logpdf(x), rel=1e-12, abs=1e-12)

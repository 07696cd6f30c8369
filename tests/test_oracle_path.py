"""Pins the oracle's path space (src/path.cpp): re-evaluating a generated
subpath at zero offset reproduces its contribution (PerturbPathBidir is the
fixed-topology restatement of GenerateSubpath), the light-coordinate
Jacobian quirk of l = 0 paths, the primary-sample layout, finite-difference
derivatives, and an unbiasedness check of the MIS-weighted strategies."""
import numpy as np
import pytest

import orc

CLS = [(c, l) for P in (3, 4, 5) for l in range(0, P + 1) for c in [P - l + 1] if not (c == 1 and l < 2)]


@pytest.fixture(scope="module")
def gen(cfg1):
    s = orc.Scene(cfg1)
    rng = np.random.default_rng(11)
    N = 3000
    cl = np.array([CLS[i % len(CLS)] for i in range(N)], np.int32)
    uni = rng.uniform(0, 1, (N, 96)).astype(np.float32)
    recs, c8 = s.generate_subpath(cl, uni, dbl=True)
    return s, cl, recs, c8


def test_dimension_layout(gen):
    """D = 2 * max(c + l - 1, 2) (path.cpp:2286-2289 minus the time dimension)
    and every generated primary sample lies in [0, 1)."""
    s, cl, recs, c8 = gen
    v = c8[:, 0] > 0
    assert v.sum() > 500
    for r in recs[v]:
        D = 2 * max(r["c"] + r["l"] - 1, 2)
        assert np.all((r["u"][:D] >= 0) & (r["u"][:D] < 1))
        assert np.all(r["u"][D:] == 0)


def test_zero_offset_reproduces_generation(gen):
    s, cl, recs, c8 = gen
    v = c8[:, 0] > 0
    rv = recs[v]
    z = np.zeros((len(rv), 32), np.float32)
    out, p8 = s.perturb(rv, z, z, dbl=True)
    assert np.all(p8[:, 0] > 0)
    np.testing.assert_array_equal(out["cprim"], rv["cprim"])
    np.testing.assert_array_equal(out["lgprim"], rv["lgprim"])
    l0 = rv["l"] == 0
    # lsScore and contribution are reproduced (u is stored in float)
    np.testing.assert_allclose(p8[:, 4], c8[v, 4], rtol=5e-5)
    np.testing.assert_allclose(p8[~l0, 5], c8[v][~l0, 5], rtol=5e-5)
    # l = 0 paths generated with light-coordinate sampling carry the light's
    # sample pdf (1 / area) in ssJacobian (path.cpp:1197-1211 vs 1430-1478)
    area = 0.5 * 0.5
    np.testing.assert_allclose(p8[l0, 5] / c8[v][l0, 5], area, rtol=5e-5)


def test_small_perturbation_is_continuous(gen):
    """f(u + eps) -> f(u) as eps -> 0 for states away from occlusion edges."""
    s, cl, recs, c8 = gen
    v = np.where(c8[:, 0] > 0)[0][:300]
    rv = recs[v]
    rng = np.random.default_rng(3)
    z = np.zeros((len(rv), 32), np.float32)
    _, base = s.perturb(rv, z, z, dbl=True)
    off = rng.normal(0, 1e-6, (len(rv), 32)).astype(np.float32)
    _, p8 = s.perturb(rv, off, z, dbl=True)
    ok = p8[:, 0] > 0
    assert ok.mean() > 0.95
    rel = np.abs(p8[ok, 5] - base[ok, 5]) / base[ok, 5]
    assert np.median(rel) < 1e-3


def test_fd_derivatives_consistent(gen):
    """FD gradient agrees with a directional difference of log f; FD Hessian is symmetric."""
    s, cl, recs, c8 = gen
    v = np.where(c8[:, 0] > 0)[0][:60]
    rv = recs[v]
    logf, g, H = s.grad_hess_fd(rv, 1e-4)
    rng = np.random.default_rng(9)
    D = 2 * np.maximum(rv["c"] + rv["l"] - 1, 2)
    errs = []
    for i in range(len(rv)):
        d = D[i]
        assert np.all(H[i, :d, :d] == H[i, :d, :d].T)
        dirv = rng.normal(size=d); dirv /= np.linalg.norm(dirv)
        h = 1e-5
        off_p = np.zeros((1, 32)); off_p[0, :d] = h * dirv
        fp = s.logf_fixed(rv[i:i + 1], off_p)[0]
        fm = s.logf_fixed(rv[i:i + 1], -off_p)[0]
        if np.isfinite(fp) and np.isfinite(fm):
            errs.append(abs((fp - fm) / (2 * h) - g[i, :d] @ dirv) / max(1.0, abs(g[i, :d] @ dirv)))
    assert np.median(errs) < 1e-4


def test_bdpt_strategies_unbiased(cfg1):
    """Per path length, GenerateSubpath over every strategy (c, l) (each MIS
    weighted) and GeneratePathBidir (all strategies of one sample) estimate the
    same integral: the MIS weights of a length sum to one in expectation."""
    s = orc.Scene(cfg1)
    rng = np.random.default_rng(21)
    N = 60000
    uni = rng.uniform(0, 1, (N, 160)).astype(np.float32)
    cs, nc = s.generate_bidir(uni, 3, 5, dbl=False)
    tot_bidir = {}
    for P in (3, 4):
        sel = (cs[:, :, 0] + cs[:, :, 1] - 1 == P) & (np.arange(cs.shape[1])[None] < nc[:, None])
        tot_bidir[P] = (cs[:, :, 5] * sel).sum() / N
    for P in (3, 4):
        cls = [(c, l) for l in range(0, P + 1) for c in [P - l + 1] if not (c == 1 and l < 2)]
        acc = 0.0
        for (c, l) in cls:
            cl = np.tile(np.array([[c, l]], np.int32), (N, 1))
            u2 = rng.uniform(0, 1, (N, 96)).astype(np.float32)
            _, c8 = s.generate_subpath(cl, u2, dbl=False)
            acc += c8[:, 4].sum() / N
        assert acc == pytest.approx(tot_bidir[P], rel=0.06), (P, acc, tot_bidir[P])

"""Pins the oracle's DR chain loop (delayed.cpp:396-606) end to end: the
image a DRMLT run converges to has the same block-wise brightness as an
independent bidirectional Monte-Carlo estimate of the same (path length >= 3)
integral, and per-stage statistics have the structure the loop implies."""
import numpy as np
import pytest

import orc
from drmlt import scenes


def _small_cfg1():
    sc = scenes.config1()
    sc["cam"] = sc["cam"].copy()
    sc["cam"][10] = sc["cam"][11] = 64
    return sc


def _blocks(img, k=8):
    H, W = img.shape[:2]
    lum = img @ np.array([0.212671, 0.715160, 0.072169])
    return lum.reshape(k, H // k, k, W // k).sum(axis=(1, 3))


@pytest.fixture(scope="module")
def small():
    sc = _small_cfg1()
    return sc, orc.Scene(sc)


def test_drmlt_converges_to_bdpt(small):
    sc, s = small
    W = H = 64
    # independent estimate: all strategies of N bidirectional samples
    rng = np.random.default_rng(0)
    N = 300000
    uni = rng.uniform(0, 1, (N, 160)).astype(np.float32)
    cs, nc = s.generate_bidir(uni, 3, 5, dbl=False)
    sel = np.arange(cs.shape[1])[None] < nc[:, None]
    c = cs[sel]
    ix = np.clip((c[:, 7] * W).astype(int), 0, W - 1)
    iy = np.clip((c[:, 8] * H).astype(int), 0, H - 1)
    ref = np.zeros((H, W, 3))
    np.add.at(ref, (iy, ix), c[:, 2:5])
    ref *= (W * H) / N
    # DRMLT
    b, lf, seeds, c8, ncand = s.prepass(50000, 5, 256, seed=3)
    iters = 4000
    film, stats, sec = s.render(seeds, c8, lf, b, iters, threads=8, seed=5)
    spp = iters * 256 / (W * H)
    img = film / spp
    assert b == pytest.approx(ref.mean(axis=(0, 1)) @ [0.212671, 0.715160, 0.072169], rel=0.05)
    rb, ib = _blocks(ref), _blocks(img)
    mask = rb > 0.05 * rb.max()
    rel = np.abs(ib[mask] - rb[mask]) / rb[mask]
    assert np.median(rel) < 0.08, np.median(rel)
    assert ib.sum() == pytest.approx(rb.sum(), rel=0.05)


def test_mcmc_converges_to_bdpt(small):
    """The oracle's mcmc integrator (the first stage alone, mcmc.cpp:125-160)
    converges to the same independent bidirectional estimate, and makes no
    second-stage moves."""
    sc, s = small
    W = H = 64
    rng = np.random.default_rng(1)
    N = 300000
    uni = rng.uniform(0, 1, (N, 160)).astype(np.float32)
    cs, nc = s.generate_bidir(uni, 3, 5, dbl=False)
    c = cs[np.arange(cs.shape[1])[None] < nc[:, None]]
    ix = np.clip((c[:, 7] * W).astype(int), 0, W - 1)
    iy = np.clip((c[:, 8] * H).astype(int), 0, H - 1)
    ref = np.zeros((H, W, 3))
    np.add.at(ref, (iy, ix), c[:, 2:5])
    ref *= (W * H) / N
    b, lf, seeds, c8, _ = s.prepass(50000, 5, 256, seed=3)
    iters = 4000
    film, stats, _ = s.render(seeds, c8, lf, b, iters, threads=8, seed=6, opts={"secondstage": 0})
    assert stats[4:].sum() == 0 and stats[0] + stats[2] == iters * 256
    img = film / (iters * 256 / (W * H))
    rb, ib = _blocks(ref), _blocks(img)
    mask = rb > 0.05 * rb.max()
    rel = np.abs(ib[mask] - rb[mask]) / rb[mask]
    assert np.median(rel) < 0.08, np.median(rel)
    assert ib.sum() == pytest.approx(rb.sum(), rel=0.05)


def test_h2mc_converges_to_bdpt(small):
    """The oracle's h2mc integrator without lens moves (h2mc.cpp:196-600:
    isotropic small steps with probability 0.2, else single-stage H2MC
    moves accepted with min(1, ratio)) converges to the same independent
    bidirectional estimate."""
    sc, s = small
    W = H = 64
    rng = np.random.default_rng(2)
    N = 300000
    uni = rng.uniform(0, 1, (N, 160)).astype(np.float32)
    cs, nc = s.generate_bidir(uni, 3, 5, dbl=False)
    c = cs[np.arange(cs.shape[1])[None] < nc[:, None]]
    ix = np.clip((c[:, 7] * W).astype(int), 0, W - 1)
    iy = np.clip((c[:, 8] * H).astype(int), 0, H - 1)
    ref = np.zeros((H, W, 3))
    np.add.at(ref, (iy, ix), c[:, 2:5])
    ref *= (W * H) / N
    b, lf, seeds, c8, _ = s.prepass(50000, 5, 256, seed=3)
    iters = 2000
    film, stats, _ = s.render(seeds, c8, lf, b, iters, threads=8, seed=7, opts={"h2mc": 1})
    assert stats[4] == stats[5] == 0 and stats[0] + stats[2] + stats[6] == iters * 256
    small_moves = stats[2] + stats[6]
    assert abs(stats[2] / small_moves - 0.2) < 0.01  # the uniform-mixing share
    img = film / (iters * 256 / (W * H))
    rb, ib = _blocks(ref), _blocks(img)
    mask = rb > 0.05 * rb.max()
    rel = np.abs(ib[mask] - rb[mask]) / rb[mask]
    assert np.median(rel) < 0.08, np.median(rel)
    assert ib.sum() == pytest.approx(rb.sum(), rel=0.05)


def test_stats_structure(small):
    sc, s = small
    b, lf, seeds, c8, _ = s.prepass(20000, 5, 64, seed=4)
    iters = 2000
    film, st, sec = s.render(seeds, c8, lf, b, iters, threads=4, seed=9)
    large, large_acc, s1, s1_acc, iso, iso_acc, h2, h2_acc = (int(x) for x in st[:8])
    total = 64 * iters
    assert large + s1 == total
    assert large / total == pytest.approx(0.3, abs=0.02)  # largestepprob
    # a second stage follows every rejected small step, split by anisoperturbprob
    assert iso + h2 == s1 - s1_acc
    assert h2 / (iso + h2) == pytest.approx(0.5, abs=0.03)
    assert 0 < s1_acc < s1 and 0 < h2_acc < h2 and 0 < iso_acc < iso


def test_h2mc_derivative_step_is_converged(cfg2_small):
    """The oracle's H2MC Gaussians take their derivatives from central
    differences of its FP64 fixed-topology log f (the reference differentiates
    exactly, with chad). On glossy and refractive vertices a step of 1e-4 is
    too coarse: the H2MC acceptance rate moves measurably. At the oracle's
    step (2e-6) halving it leaves the rate unchanged within Monte Carlo
    error, while 1e-4 lies clearly away from it."""
    O = orc.Scene(cfg2_small)
    maxd = int(cfg2_small["options"]["maxdepth"])
    b, lf, seeds, c8, _ = O.prepass(8000, maxd, 1024, seed=13, threads=8)
    rate = {}
    for h in (1e-4, 2e-6, 1e-6):
        acc = tot = 0
        for seed in (19, 23):
            _, st, _ = O.render(seeds, c8, lf, b, 16, threads=8, seed=seed, opts={"fdstep": h})
            acc += int(st[7]); tot += int(st[6])
        rate[h] = (acc / tot, tot)
    sd = lambda r, n: np.sqrt(r * (1 - r) / n)
    (r2, n2), (r1, n1), (rc, nc) = rate[2e-6], rate[1e-6], rate[1e-4]
    tol = 4 * np.hypot(sd(r2, n2), sd(r1, n1))
    assert abs(r2 - r1) < tol, rate
    assert r2 - rc > tol, rate

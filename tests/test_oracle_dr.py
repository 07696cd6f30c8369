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


def test_stats_structure(small):
    sc, s = small
    b, lf, seeds, c8, _ = s.prepass(20000, 5, 64, seed=4)
    iters = 2000
    film, st, sec = s.render(seeds, c8, lf, b, iters, threads=4, seed=9)
    large, large_acc, s1, s1_acc, iso, iso_acc, h2, h2_acc = (int(x) for x in st)
    total = 64 * iters
    assert large + s1 == total
    assert large / total == pytest.approx(0.3, abs=0.02)  # largestepprob
    # a second stage follows every rejected small step, split by anisoperturbprob
    assert iso + h2 == s1 - s1_acc
    assert h2 / (iso + h2) == pytest.approx(0.5, abs=0.03)
    assert 0 < s1_acc < s1 and 0 < h2_acc < h2 and 0 < iso_acc < iso

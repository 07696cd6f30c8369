"""Edge cases of the C ABI on the device: empty batches, argument errors on a
live context, and the longest paths the derivative path supports (path
length 12, D = 24)."""
import numpy as np
import pytest

import orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def deep(cfg1):
    """The config-1 room at maximum depth 12 (the largest path length with
    derivatives)."""
    from drmlt import capi

    sc = dict(cfg1)
    sc["options"] = dict(cfg1["options"], maxdepth=12)
    opt = capi.options_from_dpt(sc["options"], num_init_samples=8192, num_chains=1024, seed=4, max_derivatives_depth=12)
    ctx = capi.Context(sc, opt)
    yield sc, ctx, orc.Scene(sc)
    ctx.close()


def test_empty_batches(deep):
    """Zero-length batches and zero iterations are no-ops."""
    from drmlt import capi

    sc, ctx, O = deep
    z3 = np.zeros((0, 3), np.float32)
    z1 = np.zeros(0, np.float32)
    say("empty trace / eval_logf / grad_hess")
    prim, t = ctx.trace(z3, z3, z1, z1)
    assert prim.shape == (0,) and t.shape == (0,)
    recs = np.zeros(0, capi.PATH_REC)
    lf, ok = ctx.eval_logf(recs)
    assert lf.shape == (0,)
    g, h, ok = ctx.grad_hess(recs)
    assert g.shape == (0, 32)
    ctx.prepass()
    before = ctx.stats()
    ctx.run(0, capi.BATCHED_GRAPH)
    ctx.run(0, capi.PLAIN)
    assert ctx.stats() == before


def test_argument_errors(deep):
    """Out-of-range inputs are rejected with DRMLT_ERR_ARG before any device
    work, and the context stays usable."""
    from drmlt import capi

    sc, ctx, O = deep
    rec = np.zeros(1, capi.PATH_REC)
    rec["c"], rec["l"] = 14, 0  # path length 13: D = 26, beyond the derivative bound
    rec["cprim"][0, :13] = 0
    with pytest.raises(capi.DrmltError):
        ctx.grad_hess(rec)
    bad = np.zeros(1, capi.PATH_REC)
    bad["c"], bad["l"], bad["light"], bad["lprim"] = 2, 2, 99, 0  # no light 99
    bad["cprim"][0, 0] = 0
    bad["lgprim"][0, 0] = 0
    with pytest.raises(capi.DrmltError):
        ctx.eval_logf(bad)
    with pytest.raises(capi.DrmltError):
        capi.check(capi.lib().drmlt_run(ctx.h, capi.C.c_int64(1), 7))  # unknown mode
    with pytest.raises(capi.DrmltError):
        ctx.set_schedule(stage_blocks=-1)
    with pytest.raises(capi.DrmltError):
        ctx.set_schedule(stage_lanes=65)
    with pytest.raises(capi.DrmltError):
        ctx.set_schedule(stage_split=5)
    with pytest.raises(capi.DrmltError):
        ctx.set_schedule(bin_stages=3)
    with pytest.raises(capi.DrmltError):
        ctx.set_schedule(stage_waves=4)
    with pytest.raises(capi.DrmltError):
        ctx.read_chains(10 ** 7)
    with pytest.raises(capi.DrmltError):
        out = np.zeros((8, 8, 3), np.float32)
        capi.check(capi.lib().drmlt_read_film(ctx.h, capi.P(out), 8, 8, 0))
    long = np.zeros(1, capi.PATH_REC)
    long["c"], long["l"] = 14, 0  # longer than max_depth 12
    long["cprim"][0, :13] = 0
    with pytest.raises(capi.DrmltError):
        ctx.seed_chains(long, np.ones((1, 8), np.float32), 1.0, np.ones(16, np.float32))
    with pytest.raises(capi.DrmltError):
        ctx.autotune(steps=0)
    # still usable
    ctx.prepass()
    ctx.run(1, capi.BATCHED)
    assert ctx.stats()["stage1"] + ctx.stats()["large_step"] >= 1024


def test_longest_paths_with_derivatives(deep):
    """Path length 12 (D = 24): log f against the oracle's FP64 within 1e-4
    relative in f; the FP64 jet gradient against central differences of the
    oracle's log f; the loop runs at maximum depth 12."""
    from drmlt import capi

    sc, ctx, O = deep
    rng = np.random.default_rng(12)
    n = 20000
    cls = [(12 - l + 1, l) for l in range(0, 13) if 12 - l + 1 >= 2]
    cl = np.array([cls[i % len(cls)] for i in range(n)], np.int32)
    uni = rng.uniform(0, 1, (n, 160)).astype(np.float32)
    st, c8 = O.generate_subpath(cl, uni, dbl=True)
    recs = st[c8[:, 0] > 0][:200]
    assert len(recs) >= 50, len(recs)
    assert (2 * (recs["c"] + recs["l"] - 1) == 24).all()
    say(f"deep paths n={len(recs)}")
    l64, ok64 = ctx.eval_logf(recs, fp64=True)
    l32, ok32 = ctx.eval_logf(recs, fp64=False)
    lo = O.logf_fixed(recs, np.zeros((len(recs), 32)), chad_quirk=True)
    both = (ok64 == 1) & (ok32 == 1) & np.isfinite(lo)
    assert both.mean() > 0.9
    d32 = np.abs(l32[both] - lo[both])
    say("fp32 log f vs oracle: median", np.median(d32), "p95", np.percentile(d32, 95))
    assert np.median(d32) < 1e-5 and np.percentile(d32, 95) < 1e-4
    g, h, okd = ctx.grad_hess(recs[:40], fp64=True)
    assert okd.mean() > 0.9
    _, gfd, _ = O.grad_hess_fd(recs[:40], 1e-5, chad_quirk=True)
    rel = [np.linalg.norm(g[i, :24] - gfd[i, :24]) / max(np.linalg.norm(gfd[i, :24]), 1e-9) for i in range(40) if okd[i]]
    say("grad vs FD: median", np.median(rel))
    assert np.median(rel) < 1e-4
    H = h[okd == 1, :24, :24]
    assert np.allclose(H, np.transpose(H, (0, 2, 1)))
    ctx.prepass()
    say("loop at maxdepth 12: 4 graph steps")
    ctx.run(4, capi.BATCHED_GRAPH)
    film = ctx.film()
    assert np.isfinite(film).all() and film.mean() > 0

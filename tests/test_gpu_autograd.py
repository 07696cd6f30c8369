"""Fixed-topology log f (drmlt_eval_logf), its torch.autograd binding
(drmlt.torch_ad) and the chain read-back, through the C ABI.

The binding's gradient and Hessian (backward and double backward) are the
jet kernels' values; they are checked against central finite differences
with h = 1e-4 of the device's own FP64 log f, and log f itself against the
oracle's FP64 restatement. The reference's chad-generated derivatives cannot
be produced here (the reference does not build in this environment), so
agreement with them is not tested."""
import numpy as np
import pytest

import orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def c2(cfg2_small):
    from drmlt import capi

    opt = capi.options_from_dpt(cfg2_small["options"], num_init_samples=4096, num_chains=4096, seed=3)
    ctx = capi.Context(cfg2_small, opt)
    yield cfg2_small, ctx, orc.Scene(cfg2_small)
    ctx.close()


def _valid_recs(O, n, seed, P_range=(3, 4, 5)):
    """Valid path states from the oracle's GenerateSubpath (double)."""
    rng = np.random.default_rng(seed)
    cls = [(P - l + 1, l) for P in P_range for l in range(0, P + 1) if not (P - l + 1 == 1 and l < 2)]
    cl = np.array([cls[i % len(cls)] for i in range(n)], np.int32)
    uni = rng.uniform(0, 1, (n, 96)).astype(np.float32)
    st, c8 = O.generate_subpath(cl, uni, dbl=True)
    return st[c8[:, 0] > 0]


def test_eval_logf_vs_oracle(c2):
    """drmlt_eval_logf against the oracle's FP64 fixed-topology log f (the AD
    twin's function, chad MIS quirk on), to the bar of 1e-4 relative in f
    (|d log f| < 1e-4) for FP32 and for the double evaluation (scene
    constants such as light pick probabilities and emission enter both in
    float, as in the primal)."""
    sc, ctx, O = c2
    recs = _valid_recs(O, 40000, 11)[:3000]
    say(f"eval_logf n={len(recs)}")
    l64, ok64 = ctx.eval_logf(recs, fp64=True)
    l32, ok32 = ctx.eval_logf(recs, fp64=False)
    lo = O.logf_fixed(recs, np.zeros((len(recs), 32), np.float32), chad_quirk=True)
    both = (ok64 == 1) & np.isfinite(lo)
    assert both.sum() > len(recs) // 2
    assert (ok64 != np.isfinite(lo)).mean() < 2e-3
    d64 = np.abs(l64[both] - lo[both])
    say("fp64 vs oracle: median", np.median(d64), "p95", np.percentile(d64, 95), "p99", np.percentile(d64, 99))
    assert np.median(d64) < 1e-5 and np.percentile(d64, 95) < 1e-4
    b32 = both & (ok32 == 1)
    # log f differences are relative differences of f
    d32 = np.abs(l32[b32] - l64[b32])
    say("fp32 vs fp64: median", np.median(d32), "p95", np.percentile(d32, 95), "p99", np.percentile(d32, 99))
    assert np.median(d32) < 1e-5 and np.percentile(d32, 95) < 1e-4
    d32o = np.abs(l32[b32] - lo[b32])
    say("fp32 vs oracle: median", np.median(d32o), "p95", np.percentile(d32o, 95))
    assert np.median(d32o) < 1e-5 and np.percentile(d32o, 95) < 1e-4


def _one_dimension(recs, D, n):
    dims = 2 * np.maximum(recs["c"] + recs["l"] - 1, 2)
    sel = recs[dims == D]
    assert len(sel) >= n // 2, (D, len(sel))
    return sel[:n]


@pytest.mark.parametrize("D", [6, 8, 10])
def test_autograd_gradient_and_hessian(c2, D):
    """log_f through autograd: backward gives the jet gradient, double
    backward the jet Hessian (exactly), and both agree with central
    differences (h = 1e-4) of the device's FP64 log f."""
    import torch
    from drmlt import torch_ad

    sc, ctx, O = c2
    recs = _one_dimension(_valid_recs(O, 40000, 20 + D), D, 48)
    n = len(recs)
    u = torch.tensor(recs["u"][:, :D].astype(np.float64), requires_grad=True)
    say(f"autograd D={D} n={n}")
    lf = torch_ad.log_f(u, ctx, recs, fp64=True)
    keep = torch.isfinite(lf)
    assert keep.sum() > n // 2
    (g,) = torch.autograd.grad(torch.where(keep, lf, torch.zeros_like(lf)).sum(), u, create_graph=True)
    gj, hj, okj = ctx.grad_hess(recs, fp64=True)
    k = keep.numpy()
    assert np.array_equal(g.detach().numpy()[k], gj[k, :D])
    H = np.stack([torch.autograd.grad(g[:, j].sum(), u, retain_graph=True)[0].numpy() for j in range(D)], axis=1)
    assert np.array_equal(H[k], hj[k, :D, :D])
    # central differences of the device's own FP64 log f; u lives in float32,
    # so the step actually taken is read back from the perturbed records
    h = 1e-4
    fd_g = np.full((n, D), np.nan)
    fd_h = np.full((n, D, D), np.nan)
    for j in range(D):
        rp, rm = recs.copy(), recs.copy()
        rp["u"][:, j] = (recs["u"][:, j].astype(np.float64) + h).astype(np.float32)
        rm["u"][:, j] = (recs["u"][:, j].astype(np.float64) - h).astype(np.float32)
        step = rp["u"][:, j].astype(np.float64) - rm["u"][:, j].astype(np.float64)
        lp, okp = ctx.eval_logf(rp, fp64=True)
        lm, okm = ctx.eval_logf(rm, fp64=True)
        gp, _, okgp = ctx.grad_hess(rp, fp64=True)
        gm, _, okgm = ctx.grad_hess(rm, fp64=True)
        good = (okp == 1) & (okm == 1) & (okgp == 1) & (okgm == 1)
        fd_g[good, j] = (lp[good] - lm[good]) / step[good]
        fd_h[good, j, :] = (gp[good, :D] - gm[good, :D]) / step[good, None]
    rel = lambda a, b: np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-9)
    eg = [rel(gj[i, :D], fd_g[i]) for i in range(n) if k[i] and np.isfinite(fd_g[i]).all()]
    eh = [rel(hj[i, :D, :D], fd_h[i]) for i in range(n) if k[i] and np.isfinite(fd_h[i]).all()]
    say(f"D={D}: {len(eg)} points, median rel grad/FD {np.median(eg):.2e}, hess/FD {np.median(eh):.2e}")
    assert len(eg) > n // 3
    # the bar is 1e-4 relative; the difference quotients of log f carry the
    # float rounding of the scene constants (about 1e-7 in log f, divided by
    # the 2e-4 step), those of the gradient do not
    assert np.median(eg) < 1e-4 and np.percentile(eg, 90) < 1e-2
    assert np.median(eh) < 1e-4 and np.percentile(eh, 90) < 1e-2
    # FP32 jets against FP64 (the FP32 path's derivative accuracy)
    g32, h32, ok32 = ctx.grad_hess(recs, fp64=False)
    e32 = [rel(h32[i, :D, :D], hj[i, :D, :D]) for i in range(n) if k[i] and ok32[i]]
    say(f"D={D}: median rel hess32/hess64 {np.median(e32):.2e}")
    assert np.median(e32) < 1e-4


def test_read_chains_matches_loop_scores(cfg1):
    """drmlt_read_chains returns the loop's current states: records the
    function-level entry points accept, whose fixed-topology f reproduces the
    ssScore the loop stored for them (both without the chad MIS quirk)."""
    from drmlt import capi

    opt = capi.options_from_dpt(cfg1["options"], num_init_samples=8192, num_chains=2048, seed=9, quirk_chad_mis=0)
    ctx = capi.Context(cfg1, opt)
    ctx.prepass()
    say("read_chains cfg1: prepass + 4 iterations")
    ctx.run(4, capi.BATCHED)
    recs, c8 = ctx.read_chains()
    valid = c8[:, 0] > 0
    assert valid.mean() > 0.9
    recs, c8 = recs[valid], c8[valid]
    P = recs["c"] + recs["l"] - 1
    assert P.min() >= 2 and P.max() <= int(cfg1["options"]["maxdepth"])
    lf, ok = ctx.eval_logf(recs, fp64=False)
    m = (ok == 1) & (c8[:, 5] > 0)
    assert m.mean() > 0.9
    rel = np.abs(np.exp(lf[m]) - c8[m, 5]) / c8[m, 5]
    say("exp(logf) vs stored ssScore: median", np.median(rel), "share within 1e-5", np.mean(rel < 1e-5))
    # states a large step generated and no small step has re-evaluated since
    # hold a score with the relative light parameterisation (the light area
    # factor); every other state reproduces exactly up to float rounding
    assert np.median(rel) < 1e-5 and np.mean(rel < 1e-5) > 0.6
    # the read-back seeds another context (invalid states included: they
    # seed empty states that the first iteration replaces by a large step)
    allrecs, allc8 = ctx.read_chains()
    b = 0.5
    ctx2 = capi.Context(cfg1, opt)
    ctx2.seed_chains(allrecs, allc8, b, np.ones(16, np.float32))
    say("seed_chains from read_chains: 2 iterations")
    ctx2.run(2, capi.BATCHED_GRAPH)
    st = ctx2.stats()
    assert st["large_step"] + st["stage1"] == 2 * len(allrecs)
    back, back8 = ctx2.read_chains()
    assert (back8[:, 0] > 0).mean() > 0.9
    ctx2.close()
    ctx.close()

"""Function-level parity of the HIP library against the oracle (CPU
restatement), through the C ABI. Each test prints the case it is about to run
on the GPU first."""
import numpy as np
import pytest

import orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def c1(cfg1):
    from drmlt import capi

    opt = capi.options_from_dpt(cfg1["options"], num_init_samples=4096, seed=3)
    ctx = capi.Context(cfg1, opt)
    yield cfg1, ctx, orc.Scene(cfg1)
    ctx.close()


@pytest.fixture(scope="module")
def c2(cfg2_small):
    from drmlt import capi

    opt = capi.options_from_dpt(cfg2_small["options"], num_init_samples=4096, num_chains=4096, seed=3)
    ctx = capi.Context(cfg2_small, opt)
    yield cfg2_small, ctx, orc.Scene(cfg2_small)
    ctx.close()


@pytest.fixture(scope="module")
def c4(cfg4_small):
    from drmlt import capi

    opt = capi.options_from_dpt(cfg4_small["options"], num_init_samples=4096, num_chains=4096, seed=3)
    ctx = capi.Context(cfg4_small, opt)
    yield cfg4_small, ctx, orc.Scene(cfg4_small)
    ctx.close()


@pytest.fixture(scope="module")
def c5(cfg5_small):
    from drmlt import capi

    opt = capi.options_from_dpt(cfg5_small["options"], num_init_samples=4096, num_chains=4096, seed=3)
    ctx = capi.Context(cfg5_small, opt)
    yield cfg5_small, ctx, orc.Scene(cfg5_small)
    ctx.close()


def _rays(rng, n, center, half):
    org = rng.uniform(-half, half, (n, 3)) + np.asarray(center)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return org.astype(np.float32), d.astype(np.float32)


@pytest.mark.parametrize("which", ["c1", "c2"])
def test_trace_bit_exact(which, request):
    """Closest and any hit: identical triangle index and t bits as the float
    CPU search (ties to the smaller index)."""
    sc, ctx, O = request.getfixturevalue(which)
    rng = np.random.default_rng(0)
    n = 50000
    center, half = ([0, 1, 0], 0.95) if which == "c1" else ([0, 0.5, 0], 1.5)
    org, d = _rays(rng, n, center, half)
    tmin = np.full(n, 1e-3, np.float32)
    tmax = np.full(n, np.inf, np.float32)
    say(f"trace closest {which} n={n}")
    gp, gt = ctx.trace(org, d, tmin, tmax)
    op, ot = O.trace(org, d, tmin, tmax)
    assert (op >= 0).sum() > n // 4
    assert np.array_equal(gp, op)
    hit = op >= 0
    assert np.array_equal(gt[hit].view(np.int32), ot[hit].view(np.int32))
    tmax2 = rng.uniform(0.05, 3, n).astype(np.float32)
    say(f"trace any {which} n={n}")
    ga, _ = ctx.trace(org, d, tmin, tmax2, any_hit=True)
    oa = O.occluded(org, d, tmin, tmax2)
    assert 0 < oa.sum() < n
    assert np.array_equal(ga, oa)


def _valid_recs(sc, O, n, seed, P_range=(3, 4, 5)):
    """Valid path states from the oracle's GenerateSubpath (double)."""
    rng = np.random.default_rng(seed)
    cls = [(P - l + 1, l) for P in P_range for l in range(0, P + 1) if not (P - l + 1 == 1 and l < 2)]
    cl = np.array([cls[i % len(cls)] for i in range(n)], np.int32)
    uni = rng.uniform(0, 1, (n, 96)).astype(np.float32)
    st, c8 = O.generate_subpath(cl, uni, dbl=True)
    ok = c8[:, 0] > 0
    return st[ok], c8[ok]


@pytest.mark.parametrize("which", ["c1", "c2"])
def test_perturb_parity(which, request):
    """PerturbPathBidir: same validity, triangles and scores as the oracle in
    double (float32 path arithmetic: relative error of the scores ~1e-6)."""
    pytest.skip("drmlt_perturb is switched off: its kernel made a GPU memory fault whose cause is not established")
    sc, ctx, O = request.getfixturevalue(which)
    recs, _ = _valid_recs(sc, O, 6000, 1)
    rng = np.random.default_rng(2)
    n = len(recs)
    off = rng.normal(0, 0.01, (n, 32)).astype(np.float32)
    disc = rng.normal(0, 0.01, (n, 32)).astype(np.float32)
    say(f"perturb {which} n={n}")
    gs, g8 = ctx.perturb(recs, off, disc)
    os_, o8 = O.perturb(recs, off, disc, dbl=True)
    vg, vo = g8[:, 0] > 0, o8[:, 0] > 0
    assert vo.sum() > n // 3
    assert (vg != vo).mean() < 2e-3
    both = vg & vo
    same = both & (gs["cprim"] == os_["cprim"]).all(1) & (gs["lgprim"] == os_["lgprim"]).all(1)
    assert same.sum() > 0.99 * both.sum()
    for k in (4, 5):  # lsScore, ssScore
        rel = np.abs(g8[same, k] - o8[same, k]) / np.maximum(np.abs(o8[same, k]), 1e-30)
        assert np.median(rel) < 1e-5 and np.percentile(rel, 99) < 1e-3, (k, np.median(rel), np.percentile(rel, 99))
    assert np.abs(gs["u"][same] - os_["u"][same]).max() < 1e-4


def test_generate_subpath_parity(c1):
    pytest.skip("drmlt_generate_subpath is switched off: its kernel made a GPU memory fault whose cause is not established")


@pytest.mark.parametrize("which", ["c1", "c2", "c4", "c5"])
def test_generate_bidir_parity(which, request):
    """GeneratePathBidir (all strategies, MIS weights): the same contribution
    list as the oracle (FP64) for nearly every sample. On the smooth-lobe
    scenes at depths 10 and 12 (c4, c5) float arithmetic itself errs more:
    there the device's FP32 errors are held to those of the oracle's float
    build (the reference as shipped) on the same samples."""
    sc, ctx, O = request.getfixturevalue(which)
    rng = np.random.default_rng(5)
    n = 8000 if which in ("c4", "c5") else 2000  # small lights: fewer random paths contribute
    maxd = int(sc["options"]["maxdepth"])
    uni = rng.uniform(0, 1, (n, 160)).astype(np.float32)
    say(f"generate_bidir {which} n={n} maxdepth={maxd}")
    gcs, gnc = ctx.generate_bidir(uni, 3, maxd)
    ocs, onc = O.generate_bidir(uni, 3, maxd, dbl=True)
    fcs, fnc = O.generate_bidir(uni, 3, maxd, dbl=False)
    assert onc.sum() > n // 20

    def errors(cs, nc):
        same = nc == onc
        idx = np.nonzero(same)[0]
        a = np.concatenate([cs[i, :nc[i]] for i in idx])
        o = np.concatenate([ocs[i, :onc[i]] for i in idx])
        assert np.array_equal(a[:, :2], o[:, :2])  # strategies (c, l)
        return same.mean(), {k: np.abs(a[:, k] - o[:, k]) / np.maximum(np.abs(o[:, k]), 1e-20) for k in (2, 3, 4, 6, 9)}

    same_g, eg = errors(gcs, gnc)
    same_f, ef = errors(fcs, fnc)
    say(f"generate_bidir {which}: same count gpu {same_g:.4f} oracle-float {same_f:.4f}; ssScore median rel gpu "
        f"{np.median(eg[6]):.3g} oracle-float {np.median(ef[6]):.3g}")
    smooth = which in ("c4", "c5")
    assert same_g > (min(0.99, same_f - 0.005) if smooth else 0.99)
    for k in (2, 3, 4, 6, 9):  # rgb, ssScore, MIS weight
        if smooth:
            assert np.median(eg[k]) <= max(1e-5, 1.5 * np.median(ef[k])), (k, np.median(eg[k]), np.median(ef[k]))
            assert np.mean(eg[k] < 1e-3) >= min(0.99, np.mean(ef[k] < 1e-3) - 0.01), (k, np.mean(eg[k] < 1e-3), np.mean(ef[k] < 1e-3))
        else:
            assert np.median(eg[k]) < 1e-5, (k, np.median(eg[k]))
            assert np.mean(eg[k] < 1e-3) > 0.99, (k, np.mean(eg[k] < 1e-3))


@pytest.mark.parametrize("which", ["c1", "c2"])
def test_grad_hess_vs_finite_differences(which, request):
    """Jet derivatives of log f at fixed triangles (FP64 on the device) against
    central differences of the oracle's fixed-topology log f; FP32 against
    FP64."""
    sc, ctx, O = request.getfixturevalue(which)
    recs, _ = _valid_recs(sc, O, 3000 if which == "c1" else 40000, 7)
    recs = recs[:400]
    say(f"grad_hess fp64 {which} n={len(recs)}")
    g64, h64, ok64 = ctx.grad_hess(recs, fp64=True)
    say(f"grad_hess fp32 {which} n={len(recs)}")
    g32, h32, ok32 = ctx.grad_hess(recs, fp64=False)
    # central differences at two steps; points where the two disagree are
    # where the difference quotient itself is unreliable (high curvature)
    lf, gfd, hfd = O.grad_hess_fd(recs, 2e-5, chad_quirk=True)
    _, gfd2, hfd2 = O.grad_hess_fd(recs, 1e-5, chad_quirk=True)
    D = 2 * np.maximum(recs["c"] + recs["l"] - 1, 2)
    rel = lambda a, b: np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-9)
    eg, eh, e32g, e32h = [], [], [], []
    for i in range(len(recs)):
        d = D[i]
        G1, G2, H1, H2 = gfd[i, :d], gfd2[i, :d], hfd[i, :d, :d], hfd2[i, :d, :d]
        if not (ok64[i] and np.isfinite(lf[i]) and np.isfinite(H1).all() and np.isfinite(H2).all()):
            continue
        if rel(G1, G2) > 1e-5 or rel(H1, H2) > 1e-3:
            continue
        eg.append(rel(g64[i, :d], G2))
        eh.append(rel(h64[i, :d, :d], H2))
        if ok32[i]:
            e32g.append(rel(g32[i, :d], g64[i, :d]))
            e32h.append(rel(h32[i, :d, :d], h64[i, :d, :d]))
    assert len(eg) >= 20 and len(eg) > len(recs) // 3, (len(eg), len(recs))
    say("n", len(eg), "median rel grad64/FD", np.median(eg), "hess64/FD", np.median(eh), "grad32/64", np.median(e32g),
        "hess32/64", np.median(e32h))
    assert np.median(eg) < 1e-5 and np.percentile(eg, 95) < 1e-3
    assert np.median(eh) < 1e-4 and np.percentile(eh, 95) < 1e-2
    assert np.median(e32g) < 1e-5 and np.median(e32h) < 1e-4
    # symmetry of the assembled Hessian
    for i in range(min(50, len(recs))):
        d = D[i]
        H = h64[i, :d, :d]
        assert np.allclose(H, H.T, rtol=1e-9, atol=1e-9 * max(np.abs(H).max(), 1.0))


def _roughness(sc, prim, disc):
    # Roughness per vertex (phong.cpp:148-155, roughdielectric.h:61-63, lambertian.h:39-41)
    m = sc["mats"][sc["tri_mat"][prim]]
    return float(orc.lib().orc_roughness(orc.P(np.ascontiguousarray(m)), orc.C.c_double(float(disc))))


@pytest.mark.parametrize("which", ["cfg4_small", "cfg5_small"])
def test_logf_and_derivatives_on_smooth_lobes(which, request):
    """log f (drmlt_eval_logf, FP64) and its gradient and Hessian
    (drmlt_grad_hess, FP64 and FP32) at path states the loop reached on the
    reduced config-4 and config-5 scenes -- dimensions up to 20 and 24,
    vertices on Phong n in [600, 1000] and dielectric alpha in [0.02, 0.03)
    in the relative parameterisation (path.cpp:616-660) -- against the
    oracle's FP64 log f and its central differences at two steps (where the
    two agree)."""
    from drmlt import capi

    sc = request.getfixturevalue(which)
    O = orc.Scene(sc)
    ctx = capi.Context(sc, capi.options_from_dpt(sc["options"], num_init_samples=20000, num_chains=8192, seed=21))
    say(f"smooth-lobe states {which}: pre-pass + 3 iterations")
    ctx.prepass()
    ctx.run(3, capi.BATCHED_GRAPH)
    recs, c8 = ctx.read_chains()
    P = recs["c"] + recs["l"] - 1
    recs = recs[(c8[:, 0] > 0) & (c8[:, 5] > 1e-15) & (recs["c"] + recs["l"] >= 3) & (P <= 12)]
    D = 2 * np.maximum(recs["c"] + recs["l"] - 1, 2)
    smooth = np.array([any(_roughness(sc, r["cprim"][k], r["cdisc"][k]) < 0.03 for k in range(r["c"] - 1)) or
                       any(_roughness(sc, r["lgprim"][k], r["ldisc"][k]) < 0.03 for k in range(r["l"] - 1)) for r in recs])
    say(f"{which}: {len(recs)} states, D histogram {np.bincount(D).tolist()}, {int(smooth.sum())} with a smooth vertex")
    long_d = 12 if which == "cfg4_small" else 16  # the loop's config-4 paths are shorter
    assert smooth.sum() > len(recs) // 10 and (D >= long_d).sum() > 50
    say(f"eval_logf fp64 {which} n={len(recs)}")
    l64, ok64 = ctx.eval_logf(recs, fp64=True)
    lo = O.logf_fixed(recs, np.zeros((len(recs), 32), np.float32), chad_quirk=True)
    both = (ok64 == 1) & np.isfinite(lo)
    assert both.mean() > 0.99
    d64 = np.abs(l64[both] - lo[both])
    say(f"{which} log f fp64 vs oracle: median {np.median(d64):.3g} p95 {np.percentile(d64, 95):.3g}")
    assert np.median(d64) < 1e-5 and np.percentile(d64, 95) < 1e-4
    # derivatives: states with a smooth vertex first, long ones among them
    order = np.lexsort((-D, ~smooth))
    sub = recs[order[:300]]
    Ds = D[order[:300]]
    say(f"grad_hess fp64/fp32 {which} n={len(sub)}")
    g64, h64, okg = ctx.grad_hess(sub, fp64=True)
    g32, h32, ok32 = ctx.grad_hess(sub, fp64=False)
    lf, gfd, hfd = O.grad_hess_fd(sub, 2e-5, chad_quirk=True)
    _, gfd2, hfd2 = O.grad_hess_fd(sub, 1e-5, chad_quirk=True)
    rel = lambda a, b: np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-9)
    eg, eh, e32, big = [], [], [], 0
    for i in range(len(sub)):
        d = Ds[i]
        if not (okg[i] and np.isfinite(lf[i]) and np.isfinite(hfd[i]).all() and np.isfinite(hfd2[i]).all()):
            continue
        if rel(gfd[i, :d], gfd2[i, :d]) > 1e-5 or rel(hfd[i, :d, :d], hfd2[i, :d, :d]) > 1e-3:
            continue
        eg.append(rel(g64[i, :d], gfd2[i, :d]))
        eh.append(rel(h64[i, :d, :d], hfd2[i, :d, :d]))
        big += d >= long_d
        if ok32[i]:
            e32.append(rel(h32[i, :d, :d], h64[i, :d, :d]))
    say(f"{which}: {len(eg)} derivative states compared ({big} with D >= {long_d}): median rel grad {np.median(eg):.3g}, "
        f"hess {np.median(eh):.3g}, fp32 vs fp64 hess {np.median(e32):.3g}")
    assert len(eg) >= 60 and big >= 10
    assert np.median(eg) < 1e-5 and np.percentile(eg, 90) < 1e-3
    assert np.median(eh) < 1e-4 and np.percentile(eh, 90) < 1e-2
    assert np.median(e32) < 1e-3
    ctx.close()


def test_gaussian_and_logpdf():
    """ComputeGaussian / GaussianLogPdf in FP64 against the oracle on random
    gradients and symmetric Hessians of every dimension 4..24."""
    from drmlt import capi

    rng = np.random.default_rng(11)
    n = 600
    D = rng.integers(2, 13, n).astype(np.int32) * 2
    grad = rng.normal(0, 50, (n, 32))
    A = rng.normal(0, 3e3, (n, 32, 32))
    hess = -(A + A.transpose(0, 2, 1)) / 2
    sc = np.where(rng.uniform(size=n) < 0.05, 1e-16, 1.0)
    for i in range(n):
        grad[i, D[i]:] = 0
        hess[i, D[i]:, :] = 0
        hess[i, :, D[i]:] = 0
    say(f"compute_gaussian n={n}")
    m, L, Q, ld = capi.compute_gaussian(D, sc, 0.01, grad, hess)
    om, oL, oQ, old = orc.compute_gaussian(D, sc, 0.01, grad, hess)
    for i in range(n):
        d = D[i]
        assert np.allclose(m[i, :d], om[i, :d], rtol=1e-9, atol=1e-12 * max(1, np.abs(om[i, :d]).max()))
        assert np.allclose(Q[i, :d, :d], oQ[i, :d, :d], rtol=1e-9, atol=1e-9 * np.abs(oQ[i, :d, :d]).max())
        assert abs(ld[i] - old[i]) <= 1e-9 * max(1.0, abs(old[i]))
    x = rng.normal(0, 0.01, (n, 32))
    say(f"gaussian_logpdf n={n}")
    lp = capi.gaussian_logpdf(D, x, 1.0, m, L, Q, ld)
    for i in range(0, n, 7):
        d = D[i]
        ref = orc.gaussian_logpdf(int(d), x[i, :d], om[i, :d], oQ[i, :32, :32], float(old[i]))
        assert abs(lp[i] - ref) <= 1e-8 * max(1.0, abs(ref))


def test_loop_gaussian_code_vs_oracle():
    """The loop's own Gaussian code (drmlt_gaussian_eval_f64: gauss_compute,
    gauss_logpdf and gauss_sample, the functions k_b_gauss, h2mc_offset and
    h2mc_ratio run, dimensions <= 8 in 8-lane groups) against the oracle's
    ComputeGaussian / GaussianLogPdf / GenerateSample: log-densities within
    1e-9 relative; a sample from zero normals is the mean; any sample's
    Mahalanobis norm under the oracle's inverse covariance is the normals'
    squared norm (the covariance factor itself is fixed only up to
    eigenvector signs, so the sample is compared through that invariant)."""
    from drmlt import capi

    rng = np.random.default_rng(17)
    n = 800
    D = rng.integers(2, 13, n).astype(np.int32) * 2
    grad = rng.normal(0, 50, (n, 32))
    A = rng.normal(0, 3e3, (n, 32, 32))
    hess = -(A + A.transpose(0, 2, 1)) / 2
    sc = np.where(rng.uniform(size=n) < 0.05, 1e-16, 1.0)
    x = rng.normal(0, 0.01, (n, 32))
    nrm = rng.normal(size=(n, 32))
    nrm[::2] = 0.0
    for i in range(n):
        grad[i, D[i]:] = 0
        hess[i, D[i]:, :] = 0
        hess[i, :, D[i]:] = 0
    om, _, oQ, old = orc.compute_gaussian(D, sc, 0.01, grad, hess)
    for sign in (1.0, -1.0):
        say(f"gaussian_eval n={n} sign={sign}")
        lp, smp = capi.gaussian_eval(D, sc, 0.01, grad, hess, x, sign, nrm)
        assert (D <= 8).sum() > 100 and (D > 8).sum() > 100  # both executors
        for i in range(n):
            d = int(D[i])
            ref = orc.gaussian_logpdf(d, sign * x[i, :d], om[i, :d], oQ[i, :32, :32], float(old[i]))
            assert abs(lp[i] - ref) <= 1e-9 * max(1.0, abs(ref)), (i, d, lp[i], ref)
            dl = smp[i, :d] - om[i, :d]
            if i % 2 == 0:
                assert np.allclose(smp[i, :d], om[i, :d], rtol=1e-9, atol=1e-12 * max(1.0, np.abs(om[i, :d]).max()))
            else:
                maha = dl @ oQ[i, :d, :d] @ dl
                ref_n = nrm[i, :d] @ nrm[i, :d]
                assert abs(maha - ref_n) <= 1e-7 * ref_n, (i, d, maha, ref_n)
            assert not smp[i, d:].any()


def test_alpha2_closed_form():
    """Second-stage acceptance (delayed.cpp:466-541) against its closed form."""
    from drmlt import capi

    rng = np.random.default_rng(3)
    n = 4000
    typ = rng.integers(0, 2, n).astype(np.float64)
    a1 = rng.uniform(0, 0.999, n).astype(np.float32).astype(np.float64)
    a2 = rng.uniform(0, 3, n)
    ast = np.where(rng.uniform(size=n) < 0.1, 1.0, rng.uniform(0, 1, n).astype(np.float32).astype(np.float64))
    tr = rng.uniform(0.1, 3, n)
    sv = (rng.uniform(size=n) < 0.9).astype(np.float64)
    rows = np.stack([typ, a1, a2, ast, tr, sv], 1)
    say(f"alpha2 n={n}")
    out = capi.dr_alpha2(rows)
    one_minus = lambda v: np.float64(np.float32(1) - np.float32(v))  # 1 - a in float, as the reference
    for i in range(n):
        if typ[i] == 0:
            # Green-Mira: an invalid y* makes Trace return 0 (delayed.cpp:488-503)
            a_star = np.float32(ast[i]) if sv[i] != 0 else np.float32(0)
            exp = 0.0 if a_star == 1 else min(1.0, a2[i] * one_minus(a_star) / one_minus(a1[i]))
        else:
            # Tierney-Mira with the isotropic transition ratio (delayed.cpp:505-541)
            a_star = np.float32(ast[i])
            exp = 0.0 if a_star == 1 else min(1.0, a2[i] * tr[i] * one_minus(a_star) / one_minus(a1[i]))
        assert abs(out[i] - exp) <= 1e-12 * max(1.0, exp), (i, rows[i], out[i], exp)
    # special cases: a* == 1 rejects, a1 == 0 with a* == 0 gives min(1, a2)
    r = np.array([[0, 0.5, 2.0, 1.0, 1.0, 1.0], [1, 0.5, 2.0, 1.0, 1.0, 1.0], [0, 0.0, 0.25, 0.0, 1.0, 1.0],
                  [0, 0.0, 1.5, 0.0, 1.0, 1.0], [1, 0.0, 0.25, 0.0, 2.0, 1.0]])
    assert np.array_equal(capi.dr_alpha2(r), [0.0, 0.0, 0.25, 1.0, 0.5])

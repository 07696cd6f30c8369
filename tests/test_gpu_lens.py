"""The h2mc integrator's lens mutation on the GPU (h2mc.cpp:463-530,
PerturbLensBidir path.cpp:1914-2067): the lens score, lensVertexPos and lens
derivatives per state against the oracle; the plain, batched, graph and split
graph forms advancing identically; the loop against the oracle's h2mc loop
with lens mutations; the statistics text with the lens line."""
import numpy as np
import pytest

import orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _lens_states(O, n, seed, lengths=(3, 4, 5)):
    rng = np.random.default_rng(seed)
    cls = [(P - l + 1, l) for P in lengths for l in range(0, P + 1) if not (P - l + 1 == 1 and l < 2)]
    cl = np.array([cls[i % len(cls)] for i in range(n)], np.int32)
    uni = rng.uniform(0, 1, (n, 96)).astype(np.float32)
    st, c8 = O.generate_subpath(cl, uni, dbl=True)
    tracked = O.generate_subpath_lens(cl, uni, dbl=True)
    ok = (c8[:, 0] > 0) & (tracked > 0)
    return st[ok], c8[ok], tracked[ok]


@pytest.mark.parametrize("which", ["cfg1", "cfg2_small", "cfg4_small"])
def test_lens_score_and_derivatives_vs_oracle(which, request):
    """Per state: the device's lens score and lensVertexPos (its FP32
    fixed-topology re-evaluation) against the oracle's FP64 one, and the FP32
    jet gradient and Hessian of the lens function against central
    differences of the oracle's FP64 lens function, at the state's own
    screen position."""
    from drmlt import capi

    sc = request.getfixturevalue(which)
    O = orc.Scene(sc)
    # smooth lobes (Phong n >= 600 on cfg4_small) have no lens score: more samples there
    st, c8, tracked = _lens_states(O, {"cfg1": 3000, "cfg2_small": 12000}.get(which, 60000), 3)
    assert len(st) > 60
    ctx = capi.Context(sc, capi.options_from_dpt(dict(sc["options"], integrator="h2mc"), num_chains=1024, seed=3))
    say(f"lens_grad_hess {which} n={len(st)}")
    out, ok = ctx.lens_grad_hess(st, c8[:, 6:8])
    # the same lensVertexPos passed in explicitly gives the same derivatives
    out2, ok2 = ctx.lens_grad_hess(st, c8[:, 6:8], out[:, 1:4])
    ctx.close()
    assert np.array_equal(ok, ok2) and np.array_equal(out, out2)
    ref = O.lens_score(st, dbl=True)
    rel = np.abs(out[:, 0] - ref[:, 0]) / ref[:, 0]
    say("lens score rel median / p99 / max", np.median(rel), np.percentile(rel, 99), rel.max())
    assert np.median(rel) < 1e-5 and np.percentile(rel, 99) < 1e-3
    pos_err = np.abs(out[:, 1:4] - ref[:, 1:4]).max(1)
    c3 = st["c"] >= 3
    assert pos_err[c3].max() < 1e-3, pos_err[c3].max()
    logl, g, H = O.lens_grad_hess_fd(st, c8[:, 6:8], h=2e-6)
    fin = ok.astype(bool) & np.isfinite(logl) & np.isfinite(g).all(1) & np.isfinite(H).reshape(-1, 4).all(1)
    assert fin.mean() > 0.9
    gs = np.abs(g[fin]).max(1) + 1.0
    gerr = np.abs(out[fin, 4:6] - g[fin]).max(1) / gs
    hv = np.stack([H[fin, 0, 0], H[fin, 0, 1], H[fin, 1, 1]], 1)
    hs = np.abs(hv).max(1) + 1.0
    herr = np.abs(out[fin, 6:9] - hv).max(1) / hs
    say("gradient rel median / p95", np.median(gerr), np.percentile(gerr, 95), "Hessian", np.median(herr), np.percentile(herr, 95))
    assert np.median(gerr) < 1e-3 and np.percentile(gerr, 95) < 2e-2
    assert np.median(herr) < 1e-2 and np.percentile(herr, 95) < 1e-1


@pytest.mark.parametrize("which", ["cfg1", "cfg2_small"])
def test_lens_forms_agree(which, request):
    """h2mc at lensperturbprob 0.3: the plain, batched, graph and split graph
    forms make the same moves with the same outcomes."""
    from drmlt import capi

    sc = request.getfixturevalue(which)
    dpt = dict(sc["options"], integrator="h2mc", lensperturbprob=0.3)
    res, films = {}, {}
    for mode in (capi.PLAIN, capi.BATCHED, capi.BATCHED_GRAPH, "split"):
        ctx = capi.Context(sc, capi.options_from_dpt(dpt, num_init_samples=8192, num_chains=4096, seed=5))
        say(f"{which} h2mc+lens form {mode}: prepass + 3 iterations")
        ctx.prepass()
        if mode == "split":
            ctx.set_schedule(stage_split=3)
        ctx.run(3, capi.BATCHED_GRAPH if mode == "split" else mode)
        res[mode] = ctx.stats()
        films[mode] = ctx.film(1)
        ctx.close()
    say("h2mc+lens", res[capi.PLAIN])
    assert res[capi.PLAIN] == res[capi.BATCHED] == res[capi.BATCHED_GRAPH] == res["split"]
    st = res[capi.PLAIN]
    assert st["lens_step"] > 0 and st["lens_step_accepted"] > 0 and st["stage2_h2mc"] > 0
    assert st["large_step"] + st["stage1"] + st["stage2_h2mc"] + st["lens_step"] == 3 * 4096
    for m in (capi.BATCHED, capi.BATCHED_GRAPH, "split"):
        assert np.allclose(films[m], films[capi.PLAIN], rtol=1e-4, atol=1e-6 * films[capi.PLAIN].max())


@pytest.mark.parametrize("which", ["cfg1", "cfg2_small"])
def test_lens_loop_vs_oracle(which, request):
    """From the oracle's seeds, the device's h2mc loop with lens mutations
    (lensperturbprob 0.3) and the oracle's: acceptance rates of large steps,
    isotropic moves, H2MC moves and lens mutations, the share of lens
    mutations and the film mean agree within Monte Carlo error."""
    from drmlt import capi

    sc = request.getfixturevalue(which)
    O = orc.Scene(sc)
    maxd = int(sc["options"]["maxdepth"])
    chains, iters = 2048, 32
    b, lf, seeds, c8, _ = O.prepass(20000, maxd, chains, seed=13)
    dpt = dict(sc["options"], integrator="h2mc", lensperturbprob=0.3)
    ctx_opt = capi.options_from_dpt(dpt, num_chains=chains, seed=17)
    ctx = capi.Context(sc, ctx_opt)
    say(f"h2mc+lens loop {which} {chains} chains x {iters} iterations vs the oracle")
    ctx.seed_chains(seeds, c8, b, lf)
    ctx.run(iters, capi.BATCHED_GRAPH)
    gfilm, gs = ctx.film(1), ctx.stats()
    ctx.close()
    # the scene's anisoperturbstddev also sets the lens sigma (parsescene.cpp:553-554)
    lens_sigma = float(ctx_opt.lens_perturb_std_dev)
    say("lens sigma", lens_sigma)
    ofilm, ost, _ = O.render(seeds, c8, lf, b, iters, seed=19, opts={"h2mc": 1, "lensperturbprob": 0.3, "lensperturbstddev": lens_sigma})
    names = ["large_step", "large_step_accepted", "stage1", "stage1_accepted", "stage2_iso", "stage2_iso_accepted",
             "stage2_h2mc", "stage2_h2mc_accepted", "lens_step", "lens_step_accepted"]
    os_ = dict(zip(names, [int(v) for v in ost]))
    say("gpu", gs, "\noracle", os_)
    for a, t in (("large_step_accepted", "large_step"), ("stage1_accepted", "stage1"), ("stage2_h2mc_accepted", "stage2_h2mc"),
                 ("lens_step_accepted", "lens_step")):
        rg, ro = gs[a] / max(gs[t], 1), os_[a] / max(os_[t], 1)
        tol = 4 * np.sqrt(ro * (1 - ro) / max(os_[t], 1)) + 0.01
        assert abs(rg - ro) < tol, (a, rg, ro, tol)
    glens = gs["lens_step"] / (gs["lens_step"] + gs["stage2_h2mc"])
    olens = os_["lens_step"] / (os_["lens_step"] + os_["stage2_h2mc"])
    assert abs(glens - olens) < 0.03, (glens, olens)
    assert os_["lens_step"] > 1000
    gm, om = gfilm.mean(axis=(0, 1)), ofilm.mean(axis=(0, 1))
    say("film mean gpu", gm, "oracle", om)
    assert np.allclose(gm, om, rtol=0.05)


def test_lens_stats_text(cfg1, tmp_path):
    from drmlt import capi

    opt = capi.options_from_dpt(dict(cfg1["options"], integrator="h2mc", lensperturbprob=0.3), num_init_samples=8192,
                                num_chains=1024, seed=25)
    ctx = capi.Context(cfg1, opt)
    ctx.prepass()
    ctx.render(8, capi.BATCHED_GRAPH, partial_dir=tmp_path / "h2mc_partial", base_name="h2mc")
    st = ctx.stats()
    ctx.close()
    assert st["lens_step"] > 0
    text = (tmp_path / "h2mc_partial" / "h2mc_stats.txt").read_text()
    assert text == capi.stats_text(st, integrator="h2mc")
    assert f"({st['lens_step']})" in text.splitlines()[4]

"""Loop-level tests of the HIP library: pre-pass, direct lighting and the DR
mutation loop in its three execution forms, against the oracle's CPU loop
(statistically: the two draw from different random streams) and against each
other (the forms draw identical streams)."""
import numpy as np
import pytest

import orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _ctx(sc, **kw):
    from drmlt import capi

    opt = capi.options_from_dpt(sc["options"], **kw)
    return capi.Context(sc, opt)


def test_prepass_and_direct_vs_oracle(cfg1):
    """Normalisation b and the path-length distribution (delayed.cpp:14-173)
    within Monte Carlo error of the oracle's; direct-lighting film mean
    likewise (integrator.h:33-74)."""
    ctx = _ctx(cfg1, num_init_samples=100000, num_chains=1024, seed=5)
    O = orc.Scene(cfg1)
    say("prepass cfg1 100000 samples")
    b, lf = ctx.prepass(n_len=16)
    ob, olf, _, _, _ = O.prepass(100000, int(cfg1["options"]["maxdepth"]), 1024, seed=9)
    say("b gpu", b, "oracle", ob)
    assert abs(b - ob) < 0.03 * ob
    n = min(len(olf), len(lf))
    assert np.allclose(lf[:n] / lf[:n].sum(), olf[:n] / olf[:n].sum(), atol=0.02)
    say("direct cfg1")
    ctx.direct_lighting()
    d = ctx.film(2)
    spp = int(ctx.opt.direct_spp)
    od = O.direct_pass(spp, seed=3)
    gm, om = d.mean(axis=(0, 1)), od.mean(axis=(0, 1))
    say("direct mean gpu", gm, "oracle", om)
    assert np.allclose(gm, om, rtol=0.02)
    ctx.close()


def test_forms_agree(cfg1):
    """Plain, batched and graph-replayed forms advance the same chains with
    the same random streams: acceptance counts agree and the films match
    closely."""
    from drmlt import capi

    res = {}
    for mode in (capi.PLAIN, capi.BATCHED, capi.BATCHED_GRAPH):
        ctx = _ctx(cfg1, num_init_samples=8192, num_chains=2048, seed=21)
        say(f"form {mode}: prepass + 8 iterations")
        ctx.prepass()
        ctx.run(8, mode)
        res[mode] = (ctx.film(1), ctx.stats())
        ctx.close()
    f0, s0 = res[capi.PLAIN]
    say("stats", capi.PLAIN, s0)
    assert res[capi.BATCHED][1] == res[capi.BATCHED_GRAPH][1]
    for mode in (capi.BATCHED, capi.BATCHED_GRAPH):
        f, s = res[mode]
        say("stats", mode, s)
        assert s["large_step"] + s["stage1"] == 8 * 2048
        for k in s:
            assert abs(s[k] - s0[k]) <= max(3, 0.01 * s0[k]), (mode, k, s[k], s0[k])
        assert abs(f.sum() - f0.sum()) <= 0.02 * f0.sum()
    fb, _ = res[capi.BATCHED]
    fg, _ = res[capi.BATCHED_GRAPH]
    assert np.allclose(fb.sum(), fg.sum(), rtol=1e-4)


@pytest.mark.parametrize("which", ["cfg1", "cfg2_small"])
def test_batched_paths_equal_plain_paths(which, request):
    """One iteration from identical seeds: the batched stage kernels evaluate
    paths as walks that share one trace site per round, the plain form with
    the sequential path code. Stage-1 moves (large and small steps), the
    isotropic second stage (with Green's y*) and the H2MC second stage (the
    same jet derivatives; Gaussians computed by one thread in the plain form,
    by lane groups or whole waves in the batched one, which give bit-identical
    results) depend on nothing but those evaluations and the random streams,
    so their counts must be identical."""
    from drmlt import capi

    sc = request.getfixturevalue(which)
    res = {}
    for mode in (capi.PLAIN, capi.BATCHED):
        ctx = _ctx(sc, num_init_samples=8192, num_chains=4096, seed=5)
        say(f"{which} form {mode}: prepass + 1 iteration")
        ctx.prepass()
        ctx.run(1, mode)
        res[mode] = ctx.stats()
        ctx.close()
    say("plain", res[capi.PLAIN])
    say("batched", res[capi.BATCHED])
    for k in ("large_step", "large_step_accepted", "stage1", "stage1_accepted", "stage2_iso", "stage2_iso_accepted", "stage2_h2mc",
              "stage2_h2mc_accepted"):
        assert res[capi.PLAIN][k] == res[capi.BATCHED][k], k
    assert res[capi.PLAIN]["stage1_accepted"] > 0 and res[capi.PLAIN]["stage2_iso_accepted"] > 0
    assert res[capi.PLAIN]["stage2_h2mc_accepted"] > 0


def test_schedules_agree_and_autotune(cfg2_small):
    """Binned (all path-stage kernels, or only the second-stage ones) and
    chain-order scheduling of the path-stage kernels, 64, 32 or
    16 chains per workgroup, other workgroup counts, the path-stage kernels
    built for three resident waves per SIMD, and the graph-replayed
    step split into 2-4 ranges of the chains on as many streams (also for a
    chain count that none of them divides) advance the chains identically (iterations from the same
    seeds give the same counts, films within float summation order); the
    autotuner returns a schedule the context then runs under,
    graph-replayed."""
    from drmlt import capi

    for chains in (4096, 4097):
        res, films = [], []
        for sched, mode in (({}, capi.BATCHED), ({"bin_stages": 1}, capi.BATCHED),
                            ({"bin_stages": 1, "stage_blocks": 16, "deriv_blocks": 64, "gauss_blocks": 32}, capi.BATCHED),
                            ({"stage_lanes": 32}, capi.BATCHED), ({"bin_stages": 1, "stage_lanes": 16, "stage_blocks": 7}, capi.BATCHED),
                            ({"bin_stages": 2}, capi.BATCHED), ({"bin_stages": 2, "stage_split": 2}, capi.BATCHED_GRAPH),
                            ({"stage_split": 2}, capi.BATCHED_GRAPH), ({"stage_split": 3, "bin_stages": 1, "stage_lanes": 32}, capi.BATCHED_GRAPH),
                            ({"stage_split": 4}, capi.BATCHED_GRAPH), ({"stage_waves": 3}, capi.BATCHED),
                            ({"stage_waves": 3, "bin_stages": 1, "stage_split": 2}, capi.BATCHED_GRAPH)):
            if chains != 4096 and not sched.get("stage_split") and sched:
                continue
            ctx = _ctx(cfg2_small, num_init_samples=8192, num_chains=chains, seed=5)
            ctx.prepass()
            ctx.set_schedule(**sched)
            say(f"{chains} chains, schedule {ctx.schedule()}, form {mode}: 2 iterations")
            ctx.run(2, mode)
            res.append(ctx.stats())
            films.append(ctx.film(1).sum())
            ctx.close()
        say(res)
        assert all(r == res[0] for r in res)
        assert np.allclose(films, films[0], rtol=1e-4)
    ctx = _ctx(cfg2_small, num_init_samples=8192, num_chains=4096, seed=5)
    ctx.prepass()
    say("autotune cfg2_small 4096 chains")
    best, ms = ctx.autotune(steps=1)
    say("autotune", best, ms)
    assert ms > 0 and best == ctx.schedule() and best["bin_stages"] in (0, 1, 2)
    before = ctx.stats()
    ctx.run(3, capi.BATCHED_GRAPH)
    after = ctx.stats()
    assert after["stage1"] + after["large_step"] - before["stage1"] - before["large_step"] == 3 * 4096
    ctx.close()


@pytest.mark.parametrize("which,kind", [("cfg1", "green"), ("cfg2_small", "green"), ("cfg1", "mira"), ("cfg2_small", "mira"),
                                        ("cfg4_small", "green"), ("cfg5_small", "green")])
def test_loop_vs_oracle(which, kind, request):
    """The DR loop (H2MC second stage; Green-Mira or Tierney-Mira acceptance)
    against the oracle's loop from the oracle's own seeds: acceptance rates per
    stage and the film mean agree within Monte Carlo error (cfg1: Lambertian
    room; cfg2_small: Phong and rough dielectric surfaces; cfg4_small and
    cfg5_small: Phong n in [600, 1000] and dielectric alpha in [0.02, 0.03)
    at maximum depths 10 and 12, the relative parameterisation of smooth
    vertices)."""
    cfg1 = request.getfixturevalue(which)
    O = orc.Scene(cfg1)
    maxd = int(cfg1["options"]["maxdepth"])
    chains, iters = (2048, 32) if which in ("cfg1", "cfg2_small") else (8192, 32)
    b, lf, seeds, c8, _ = O.prepass(20000, maxd, chains, seed=13)
    from drmlt import capi

    opt = capi.options_from_dpt(dict(cfg1["options"], type=kind), num_chains=chains, seed=17)
    ctx = capi.Context(cfg1, opt)
    say(f"loop {which} {kind} {chains} chains x {iters} iterations")
    ctx.seed_chains(seeds, c8, b, lf)
    ctx.run(iters, 2)
    gfilm, gs = ctx.film(1), ctx.stats()
    # config 5 renders the acceptance map (counts of accepted moves) in both
    amap = 1 if cfg1["options"].get("acceptanceMap") else 0
    ofilm, ost, _ = O.render(seeds, c8, lf, b, iters, seed=19, opts={"type": 1 if kind == "mira" else 0, "acceptancemap": amap})
    names = ["large_step", "large_step_accepted", "stage1", "stage1_accepted", "stage2_iso", "stage2_iso_accepted",
             "stage2_h2mc", "stage2_h2mc_accepted"]
    os_ = dict(zip(names, [int(v) for v in ost]))
    say("gpu stats", gs)
    say("oracle stats", os_)
    for a, t in (("large_step_accepted", "large_step"), ("stage1_accepted", "stage1"), ("stage2_iso_accepted", "stage2_iso"),
                 ("stage2_h2mc_accepted", "stage2_h2mc")):
        rg, ro = gs[a] / max(gs[t], 1), os_[a] / max(os_[t], 1)
        tol = 4 * np.sqrt(ro * (1 - ro) / max(os_[t], 1)) + 0.01
        assert abs(rg - ro) < tol, (a, rg, ro, tol)
    assert abs(gs["large_step"] / (chains * iters) - os_["large_step"] / (chains * iters)) < 0.02
    gm, om = gfilm.mean(axis=(0, 1)), ofilm.mean(axis=(0, 1))
    say("film mean gpu", gm, "oracle", om)
    assert np.allclose(gm, om, rtol=0.05)
    ctx.close()


def test_count_work_and_timing(cfg1):
    """Work counters and per-stage timing of one batched step are populated."""
    ctx = _ctx(cfg1, num_init_samples=8192, num_chains=1024, seed=2)
    ctx.prepass()
    say("count_work cfg1")
    w = ctx.count_work()
    assert w.shape == (7, 4)
    assert w[0, 0] > 0 and w[0, 1] > 0 and w[0, 2] >= 1024 and w[0, 3] > 0
    ctx.run(2, 1)
    t = ctx.timing()
    assert t[0] > 0 and (t[1:8] >= 0).all()
    ctx.close()


@pytest.mark.parametrize("which", ["cfg1", "cfg2_small"])
def test_mcmc_integrator(which, request):
    """integrator=mcmc (MCMCIntegrator, mcmc.cpp:125-160: the first stage
    alone) on the same kernels. The plain, batched and graph forms advance
    identically; one iteration from the same seeds counts exactly what the
    DR loop's first stage counts (also with the graph step split into chain
    ranges); no second-stage moves are made; and against
    the oracle's mcmc loop from the oracle's seeds the acceptance rates and
    the film mean agree within Monte Carlo error."""
    from drmlt import capi

    sc = request.getfixturevalue(which)
    res = {}
    for integ, mode in (("mcmc", capi.PLAIN), ("mcmc", capi.BATCHED), ("mcmc", capi.BATCHED_GRAPH), ("mcmc", "split"),
                        ("drmlt", capi.BATCHED)):
        ctx = _ctx(sc, num_init_samples=8192, num_chains=4096, seed=5, integrator=capi.INTEGRATORS[integ])
        say(f"{which} {integ} form {mode}: prepass + 1 iteration")
        ctx.prepass()
        if mode == "split":  # the graph step in three chain ranges on concurrent streams
            ctx.set_schedule(stage_split=3)
        ctx.run(1, capi.BATCHED_GRAPH if mode == "split" else mode)
        res[(integ, mode)] = ctx.stats()
        ctx.close()
    m = res[("mcmc", capi.PLAIN)]
    say("mcmc", m, "\ndrmlt", res[("drmlt", capi.BATCHED)])
    assert m == res[("mcmc", capi.BATCHED)] == res[("mcmc", capi.BATCHED_GRAPH)] == res[("mcmc", "split")]
    d = res[("drmlt", capi.BATCHED)]
    for k in ("large_step", "large_step_accepted", "stage1", "stage1_accepted"):
        assert m[k] == d[k], k
    assert m["stage2_iso"] == m["stage2_h2mc"] == 0 and d["stage2_iso"] + d["stage2_h2mc"] > 0

    O = orc.Scene(sc)
    maxd = int(sc["options"]["maxdepth"])
    chains, iters = 2048, 32
    b, lf, seeds, c8, _ = O.prepass(20000, maxd, chains, seed=13)
    opt = capi.options_from_dpt(dict(sc["options"], integrator="mcmc"), num_chains=chains, seed=17)
    ctx = capi.Context(sc, opt)
    say(f"mcmc loop {which} {chains} chains x {iters} iterations vs the oracle")
    ctx.seed_chains(seeds, c8, b, lf)
    ctx.run(iters, capi.BATCHED_GRAPH)
    gfilm, gs = ctx.film(1), ctx.stats()
    ctx.close()
    ofilm, ost, _ = O.render(seeds, c8, lf, b, iters, seed=19, opts={"secondstage": 0})
    names = ["large_step", "large_step_accepted", "stage1", "stage1_accepted", "stage2_iso", "stage2_iso_accepted",
             "stage2_h2mc", "stage2_h2mc_accepted"]
    os_ = dict(zip(names, [int(v) for v in ost]))
    say("gpu", gs, "\noracle", os_)
    assert os_["stage2_iso"] == os_["stage2_h2mc"] == gs["stage2_iso"] == gs["stage2_h2mc"] == 0
    for a, t in (("large_step_accepted", "large_step"), ("stage1_accepted", "stage1")):
        rg, ro = gs[a] / max(gs[t], 1), os_[a] / max(os_[t], 1)
        tol = 4 * np.sqrt(ro * (1 - ro) / max(os_[t], 1)) + 0.01
        assert abs(rg - ro) < tol, (a, rg, ro, tol)
    gm, om = gfilm.mean(axis=(0, 1)), ofilm.mean(axis=(0, 1))
    say("film mean gpu", gm, "oracle", om)
    assert np.allclose(gm, om, rtol=0.05)


@pytest.mark.parametrize("which", ["cfg1", "cfg2_small"])
def test_h2mc_integrator(which, request):
    """integrator=h2mc without lens moves (H2MCIntegrator, h2mc.cpp:196-600
    with lensperturbprob 0) on the same kernels: an isotropic small step with
    probability 0.2, else one H2MC move accepted with min(1, ratio). The
    plain, batched and graph forms (also the graph step split into chain
    ranges) advance identically; large steps count as
    in the DR loop from the same seeds; and against the oracle's h2mc loop
    from the oracle's seeds the acceptance rates of large steps, isotropic
    moves and H2MC moves and the film mean agree within Monte Carlo error."""
    from drmlt import capi

    sc = request.getfixturevalue(which)
    dpt = dict(sc["options"], integrator="h2mc", lensperturbprob=0.0)
    res = {}
    for mode in (capi.PLAIN, capi.BATCHED, capi.BATCHED_GRAPH, "split"):
        ctx = capi.Context(sc, capi.options_from_dpt(dpt, num_init_samples=8192, num_chains=4096, seed=5))
        say(f"{which} h2mc form {mode}: prepass + 2 iterations")
        ctx.prepass()
        if mode == "split":  # the graph step in three chain ranges on concurrent streams
            ctx.set_schedule(stage_split=3)
        ctx.run(2, capi.BATCHED_GRAPH if mode == "split" else mode)
        res[mode] = ctx.stats()
        ctx.close()
    say("h2mc", res[capi.PLAIN])
    assert res[capi.PLAIN] == res[capi.BATCHED] == res[capi.BATCHED_GRAPH] == res["split"]
    st = res[capi.PLAIN]
    assert st["stage2_iso"] == 0 and st["stage2_h2mc"] > 0 and st["stage1"] > 0
    assert st["large_step"] + st["stage1"] + st["stage2_h2mc"] == 2 * 4096

    O = orc.Scene(sc)
    maxd = int(sc["options"]["maxdepth"])
    chains, iters = 2048, 32
    b, lf, seeds, c8, _ = O.prepass(20000, maxd, chains, seed=13)
    ctx = capi.Context(sc, capi.options_from_dpt(dpt, num_chains=chains, seed=17))
    say(f"h2mc loop {which} {chains} chains x {iters} iterations vs the oracle")
    ctx.seed_chains(seeds, c8, b, lf)
    ctx.run(iters, capi.BATCHED_GRAPH)
    gfilm, gs = ctx.film(1), ctx.stats()
    ctx.close()
    ofilm, ost, _ = O.render(seeds, c8, lf, b, iters, seed=19, opts={"h2mc": 1})
    names = ["large_step", "large_step_accepted", "stage1", "stage1_accepted", "stage2_iso", "stage2_iso_accepted",
             "stage2_h2mc", "stage2_h2mc_accepted"]
    os_ = dict(zip(names, [int(v) for v in ost]))
    say("gpu", gs, "\noracle", os_)
    for a, t in (("large_step_accepted", "large_step"), ("stage1_accepted", "stage1"), ("stage2_h2mc_accepted", "stage2_h2mc")):
        rg, ro = gs[a] / max(gs[t], 1), os_[a] / max(os_[t], 1)
        tol = 4 * np.sqrt(ro * (1 - ro) / max(os_[t], 1)) + 0.01
        assert abs(rg - ro) < tol, (a, rg, ro, tol)
    gsmall, osmall = gs["stage1"] + gs["stage2_h2mc"], os_["stage1"] + os_["stage2_h2mc"]
    assert abs(gs["stage1"] / gsmall - 0.2) < 0.02 and abs(os_["stage1"] / osmall - 0.2) < 0.02
    gm, om = gfilm.mean(axis=(0, 1)), ofilm.mean(axis=(0, 1))
    say("film mean gpu", gm, "oracle", om)
    assert np.allclose(gm, om, rtol=0.05)

"""Per-state parity of the DR loop's own code against the oracle: one
iteration of drmlt_run from identical states, every chain's proposals
re-evaluated by the oracle from their exact primary samples; the proposals'
validity, triangles and f, and the loop's first- and second-stage acceptance
probabilities against the reference's formulas to 1e-5 absolute
(tests/step_parity.py describes the computation). The scenes include the reduced config-4 (Phong n in
[600, 1000]) and config-5 (dielectric alpha in [0.02, 0.03)) generators,
whose smooth vertices take the relative parameterisation (path.cpp:616-660),
at their maximum depths 10 and 12."""
import numpy as np
import pytest

import orc
import step_parity as sp

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


CASES = {
    "cfg1": dict(dpt={}, chains=4096, warm=3),
    "cfg2_small": dict(dpt={}, chains=8192, warm=3),
    "cfg4_small": dict(dpt={}, chains=8192, warm=3),
    "cfg5_small": dict(dpt={"acceptanceMap": False}, chains=8192, warm=3),
    "cfg1_mira": dict(scene="cfg1", dpt={"type": "mira"}, chains=4096, warm=3),
    # the mcmc integrator (the first stage alone) and the h2mc integrator
    # (isotropic moves with probability uniformmixing, else H2MC moves
    # accepted with min(1, ratio); lens mutations are covered below)
    "cfg1_mcmc": dict(scene="cfg1", dpt={"integrator": "mcmc"}, chains=4096, warm=3),
    "cfg2_small_h2mc": dict(scene="cfg2_small", dpt={"integrator": "h2mc", "lensperturbprob": 0.0}, chains=8192, warm=3),
}


def _run(request, name, mode):
    from drmlt import capi

    case = CASES[name]
    sc = request.getfixturevalue(case.get("scene", name))
    dpt = dict(sc["options"], **case["dpt"])
    opt = capi.options_from_dpt(dpt, num_init_samples=50000, num_chains=case["chains"], seed=77)
    ctx = capi.Context(sc, opt)
    say(f"step parity {name}: pre-pass, {case['warm']} + 1 iterations, mode {mode}")
    _, lf = ctx.prepass()
    before, after = sp.run_one_iteration(ctx, case["warm"], mode)
    O = orc.Scene(sc)
    say(f"step parity {name}: oracle re-evaluation and derivatives of the H2MC states")
    r = sp.analyse(O, ctx, lf, before, after, dr_type=1 if dpt.get("type") == "mira" else 0, sigma=float(opt.perturb_std_dev),
                   max_derv=int(opt.max_derivatives_depth), integrator=int(opt.integrator))
    ctx.close()
    s = sp.summary(r)
    ctx64 = capi.Context(sc, capi.options_from_dpt(dpt, num_chains=case["chains"], quirk_chad_mis=0))
    say(f"step parity {name}: rounding baselines (oracle float, device FP64)")
    e32, e64 = sp.rounding_baselines(O, ctx64, r)
    ctx64.close()
    f = np.concatenate([r["f_rel_y"], r["f_rel_z"], r["f_rel_s"]])
    s.update({"f_orc32_median": float(np.median(e32)), "f_orc32_p99": float(np.percentile(e32, 99)),
              "f_dev64_median": float(np.median(e64)), "f_dev64_p99": float(np.percentile(e64, 99)),
              "f_rel_p90": float(np.percentile(f, 90)), "f_orc32_p90": float(np.percentile(e32, 90))})
    say(f"step parity {name}: {s}")
    say(f"step parity {name}: worst second stages {sp.worst(r)}")
    return r, s


def _check(r, s, min_second=0.05):
    assert s["second"] >= min_second * s["chains"]
    # the loop and the oracle agree on which proposals are valid and on
    # the triangles of the valid ones
    assert s["valid_agree"] >= 0.995, s
    assert s["tri_equal"] >= 0.995, s
    # (A) f of the proposals: the loop's FP32 evaluation errs against the
    # oracle's FP64 as float arithmetic itself does -- the oracle's float
    # build (the reference as shipped computes in float) has the same error
    # distribution on the same states -- and the device's FP64 evaluation of
    # those states agrees with the oracle's FP64 at the 1e-6 level
    for q in ("median", "p90", "p99"):
        assert s[f"f_rel_{q}"] <= 1.5 * s[f"f_orc32_{q}"] + 1e-7, (q, s)
    assert s["f_dev64_median"] <= 2e-6, s
    # (B) the loop's acceptances from its f values within 1e-5 absolute
    assert s["a1_B_within_1e-5"] >= 0.999, s
    assert s["alpha2_B_within_1e-5"] >= 0.999, s


@pytest.mark.parametrize("name", ["cfg1", "cfg2_small", "cfg4_small", "cfg5_small", "cfg1_mira"])
def test_step_acceptances_per_state(name, request):
    from drmlt import capi

    r, s = _run(request, name, capi.BATCHED_GRAPH)
    _check(r, s)
    if name in ("cfg4_small", "cfg5_small"):
        assert s["aniso"] > 0.05 * s["chains"]  # H2MC moves were compared


def test_step_acceptances_per_state_integrators(request):
    """The mcmc integrator's first stage and the h2mc integrator's isotropic
    and H2MC moves, per state, to the same bars."""
    from drmlt import capi

    r, s = _run(request, "cfg1_mcmc", capi.BATCHED_GRAPH)
    assert s["second"] == 0
    _check(r, s, min_second=0.0)
    r, s = _run(request, "cfg2_small_h2mc", capi.BATCHED_GRAPH)
    assert s["aniso"] > 0.3 * s["chains"]  # H2MC moves (1 - uniformmixing of the small steps)
    _check(r, s)


def test_step_acceptances_per_state_plain_form(request):
    """The plain one-kernel form runs the same stage code."""
    from drmlt import capi

    r, s = _run(request, "cfg2_small", capi.PLAIN)
    _check(r, s)


@pytest.mark.parametrize("name", ["cfg1", "cfg2_small"])
def test_lens_step_acceptances_per_state(name, request):
    """The h2mc integrator's lens mutations (lensperturbprob 0.5), one
    iteration from identical states: each lens proposal against the oracle's
    FP64 PerturbLensBidir from the same state to the same screen position
    (validity, triangles, f, lens score), and the loop's acceptance against
    h2mc.cpp:516-523 computed from the device's lens scores and lens
    derivatives with the oracle's FP64 Gaussian, to 1e-5 absolute
    (tests/step_parity.py, analyse_lens). The acceptance with the oracle's
    own lens scores and central-difference derivatives is reported beside
    it."""
    from drmlt import capi

    sc = request.getfixturevalue(name)
    dpt = dict(sc["options"], integrator="h2mc", lensperturbprob=0.5)
    opt = capi.options_from_dpt(dpt, num_init_samples=50000, num_chains=8192, seed=78)
    ctx = capi.Context(sc, opt)
    say(f"lens step parity {name}: pre-pass, 3 + 1 iterations")
    ctx.prepass()
    before, after = sp.run_one_iteration(ctx, 3, capi.BATCHED_GRAPH)
    lens_after = ctx.read_lens()
    O = orc.Scene(sc)
    say(f"lens step parity {name}: oracle lens perturbation, scores and derivatives")
    r = sp.analyse_lens(O, ctx, before, after, lens_after, float(opt.lens_perturb_std_dev), max_derv=int(opt.max_derivatives_depth))
    ctx.close()
    s = sp.summary_lens(r)
    say(f"lens step parity {name}: {s}")
    say(f"lens step parity {name}: worst {sp.worst_lens(r)}")
    assert s["lens"] > 0.1 * s["chains"]
    assert s["score_known"] == 1.0, s
    assert s["valid_agree"] >= 0.995, s
    assert s["tri_equal"] >= 0.995, s
    assert s["f_rel_median"] <= 1e-5 and s["ls_rel_median"] <= 1e-5, s
    assert s["a1_B_within_1e-5"] >= 0.999, s

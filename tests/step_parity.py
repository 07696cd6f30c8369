"""Per-state parity of one DR iteration (delayed.cpp:396-606) against the
oracle: the helper behind tests/test_gpu_step_parity.py.

One iteration runs on the device from states that are read back first
(drmlt_read_step). For every chain, the proposals the loop made -- the
first-stage proposal y, the second-stage proposal z, Green's y* -- are
re-evaluated by the oracle from their exact primary samples
(PerturbPathBidir in double, path.cpp:1480-1628). Two things are checked:

(A) each proposal's f (ssScore) as the loop evaluated it (FP32) against the
    oracle's FP64 evaluation of the same primary samples, with the
    triangles the two traced;
(B) the loop's acceptances from the f values it evaluated, against the
    reference's formulas computed here: a1 (mutation.cpp:86-99 for a large
    step, mutation.cpp:197 for a small one), the stage-2 ratio
    (mutation.cpp:197 isotropic; drsmallstep.cpp:181-196 for H2MC, with the
    oracle's FP64 ComputeGaussian / GaussianLogPdf over the device's
    derivative values from drmlt_grad_hess) and alpha2
    (delayed.cpp:466-541, isInvalid included).

The same acceptances computed from the oracle's own f values (C) are
reported; they differ from (B) by the FP32 rounding of f that (A) measures."""
import numpy as np

import orc


def _iso_gaussian(D, sigma):
    # gaussian.cpp:4-16
    inv = np.zeros((32, 32))
    inv[np.arange(D), np.arange(D)] = 1.0 / sigma ** 2
    return np.zeros(32), inv, D * np.log(1.0 / sigma ** 2)


def has_derivative(c, l, max_derv):
    # path.cpp:3186-3206 with the derivative library's depth
    return c >= 1 and l >= 0 and c + l >= 3 and c + l - 1 <= max_derv


def run_one_iteration(ctx, warm, mode):
    """warm iterations, then the read-back before and after one more."""
    ctx.run(warm, mode)
    before = ctx.read_step()
    ctx.run(1, mode)
    after = ctx.read_step()
    return before, after


def well_formed(recs, nlights, ntris):
    """Records the loop could have evaluated (the checks of drmlt_perturb's
    argument validation): vertex counts within the fixed slots, light and
    triangle indices within the scene."""
    c, l = recs["c"], recs["l"]
    P = c + l - 1
    ok = (c >= 1) & (l >= 0) & (c <= 17) & (l <= 17) & (P >= 2) & (2 * np.maximum(P, 2) <= 32) & ~((c == 1) & (l < 2))
    ok &= (l < 1) | ((recs["light"] >= 0) & (recs["light"] < nlights) & (recs["lprim"] >= 0) & (recs["lprim"] < ntris))
    k = np.arange(16)[None]
    ok &= ~(((recs["cprim"] < -1) | (recs["cprim"] >= ntris)) & (k < (c - 1)[:, None])).any(1)
    ok &= ~(((recs["lgprim"] < -1) | (recs["lgprim"] >= ntris)) & (k < (l - 1)[:, None])).any(1)
    return ok


def oracle_eval(O, recs, nlights, ntris):
    """The oracle's PerturbPathBidir (double) of each record at its own
    primary samples; records that are not well formed evaluate as invalid."""
    ok = well_formed(recs, nlights, ntris)
    out = recs.copy()
    c8 = np.zeros((len(recs), 8), np.float32)
    if ok.any():
        z = np.zeros((int(ok.sum()), 32), np.float32)
        out[ok], c8[ok] = O.perturb(recs[ok], z, z, dbl=True)
    return out, c8


def _same_triangles(a, b):
    ok = np.ones(len(a), bool)
    for i in range(len(a)):
        nc, nl = a["c"][i] - 1, a["l"][i] - 1
        ok[i] = np.array_equal(a["cprim"][i, :nc], b["cprim"][i, :nc]) and np.array_equal(a["lgprim"][i, :nl], b["lgprim"][i, :nl])
    return ok


def _invalid(x):
    return ~np.isfinite(x) | (x < 0)  # delayed.cpp:252


def _clamp(x):
    return np.where(np.isnan(x), x, np.clip(x, 0.0, 1.0))  # utils.h:245-248 (NaN passes)


def _stage1(gen, cur_valid, y_valid, lx, ly, ls_x, ss_x, ls_y, ss_y, lf):
    lfd = np.asarray(lf, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        large = ((ly + 1) * ls_y * lfd[np.clip(lx, 0, len(lfd) - 1)]) / ((lx + 1) * ls_x * lfd[np.clip(ly, 0, len(lfd) - 1)])
        small = ss_y / ss_x
    a = np.where(gen, np.where(cur_valid, large, 0.0), small)
    a = np.where(y_valid, _clamp(a), 0.0)
    return np.where(_invalid(a), 0.0, a)  # delayed.cpp:417-419, 430-432


def _alpha2(dr_type, a2pre, astar, tr, a1):
    one_m_a1 = (np.float32(1) - a1.astype(np.float32)).astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        if dr_type == 0:
            astar = np.where(_invalid(astar), 0.0, astar)
            one_m_as = (np.float32(1) - astar.astype(np.float32)).astype(np.float64)
            a = np.where(astar.astype(np.float32) == 1, 0.0, np.minimum(1.0, a2pre * one_m_as / one_m_a1))
        else:
            astar = np.where(_invalid(astar), 1.0, astar)
            one_m_as = (np.float32(1) - astar.astype(np.float32)).astype(np.float64)
            a = np.where((astar.astype(np.float32) == 1) | _invalid(tr), 0.0, np.minimum(1.0, a2pre * tr * one_m_as / one_m_a1))
    return np.where(_invalid(a2pre), 0.0, a)


def analyse(O, ctx, lf, before, after, dr_type=0, sigma=0.01, max_derv=9, integrator=0):
    """Per-chain comparisons (A), (B), (C); returns a dict of arrays and the
    masks of the chains each comparison covers. integrator 1 (mcmc): the first
    stage alone; 2 (h2mc, lens mutations off): an iteration is a large step,
    an isotropic small step accepted with a1, or an H2MC move made by the
    second-stage kernels and accepted with min(1, ratio) (h2mc.cpp:584-587),
    which proposes no y."""
    n = len(after["perm"])
    idx = np.arange(n)
    pb = before["perm"]
    pc, pf, ps = pb & 3, (pb >> 2) & 3, (pb >> 4) & 3
    X, x8 = before["slots"][idx, pc], before["contrib8"][idx, pc]
    Y, Z, S = after["slots"][idx, pf], after["slots"][idx, ps], after["slots"][idx, 3]
    y8, z8, s8 = after["contrib8"][idx, pf], after["contrib8"][idx, ps], after["contrib8"][idx, 3]
    ps_valid_before = before["contrib8"][idx, ps, 0] > 0
    flags, a1g, a2g = after["flags"], after["a1"].astype(np.float64), after["alpha2"]
    off = after["off"].astype(np.float64)
    cur_valid = x8[:, 0] > 0
    gen = ((flags & 8) != 0) | ~cur_valid
    second = (flags & 2) != 0
    aniso = second & ((flags & 4) != 0)
    y_gpu, z_gpu, s_gpu = (flags & 32) != 0, (flags & 16) != 0, (flags & 64) != 0
    y_prop = ~second if integrator == 2 else np.ones(n, bool)  # chains whose iteration proposed a y

    nl, nt = len(O.sc["light_tris"]), len(O.sc["tri_p"])
    Yo, y8o = oracle_eval(O, Y, nl, nt)
    Zo, z8o = oracle_eval(O, Z, nl, nt)
    So, s8o = oracle_eval(O, S, nl, nt)
    y_o, z_o, s_o = y8o[:, 0] > 0, z8o[:, 0] > 0, s8o[:, 0] > 0

    # (A) f of the proposals, loop (FP32) against oracle (FP64)
    def frel(g8, o8, both):
        both = both & (o8[:, 5] > 0)
        return np.abs(g8[both, 5].astype(np.float64) - o8[both, 5]) / o8[both, 5]

    y_o = y_o & y_prop  # the slot of an H2MC move's chain holds no new y
    y_both, z_both = y_gpu & y_o, second & z_gpu & z_o
    lx, ly = X["c"] + X["l"] - 1, Y["c"] + Y["l"] - 1
    ss_x, ls_x = x8[:, 5].astype(np.float64), x8[:, 4].astype(np.float64)

    # (B) and (C) first stage
    a1B = _stage1(gen, cur_valid, y_gpu, lx, ly, ls_x, ss_x, y8[:, 4].astype(np.float64), y8[:, 5].astype(np.float64), lf)
    a1C = _stage1(gen, cur_valid, y_o, lx, ly, ls_x, ss_x, y8o[:, 4].astype(np.float64), y8o[:, 5].astype(np.float64), lf)

    # second stage: the stage-2 ratio before the delayed-rejection correction
    def ratio(ss_z, z_valid, log_gauss):
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            iso = _clamp(ss_z / ss_x)
            h2 = np.exp(log_gauss) * ss_z / ss_x
        return np.where(z_valid, np.where(aniso, h2, iso), 0.0)

    log_gauss = np.zeros(n)
    h = np.nonzero(aniso & (z_gpu | z_o) & cur_valid)[0]
    if len(h):
        D = 2 * np.maximum(X["c"][h] + X["l"][h] - 1, 2)
        gx, hx, okx = ctx.grad_hess(X[h], fp64=False)
        gz, hz, okz = ctx.grad_hess(np.where(well_formed(Z[h], nl, nt), Z[h], X[h]), fp64=False)
        for k, i in enumerate(h):
            d = int(D[k])
            lps = []
            for g_, h_, ok_, sc_, sign in ((gx[k], hx[k], okx[k], ss_x[i], 1.0), (gz[k], hz[k], okz[k], float(z8[i, 5]), -1.0)):
                use = (has_derivative(int(X["c"][i]), int(X["l"][i]), max_derv) and d <= 24 and sc_ > 1e-15 and ok_
                       and np.isfinite(g_[:d]).all() and np.isfinite(h_[:d, :d]).all())
                if use:
                    m, _, inv, ld = orc.compute_gaussian(np.array([d]), np.array([sc_]), sigma, g_[None], h_[None])
                    m, inv, ld = m[0], inv[0], float(ld[0])
                else:
                    m, inv, ld = _iso_gaussian(d, sigma)
                lps.append(orc.gaussian_logpdf(d, sign * off[i, :d], m[:d], inv[:32, :32], ld))
            log_gauss[i] = lps[1] - lps[0]
    a2preB = ratio(z8[:, 5].astype(np.float64), z_gpu, log_gauss)
    a2preC = ratio(z8o[:, 5].astype(np.float64), z_o, log_gauss)
    # y* is traced where the reference traces it: Green & Mira, a valid
    # ratio that is not zero, a second proposal object marked valid
    traceB = second & ~_invalid(a2preB) & (a2preB != 0) & ps_valid_before & (dr_type == 0) & (integrator == 0)
    traceC = second & ~_invalid(a2preC) & (a2preC != 0) & ps_valid_before & (dr_type == 0) & (integrator == 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        if dr_type == 0:
            astarB = np.where(traceB & s_gpu, _clamp(s8[:, 5].astype(np.float64) / z8[:, 5]), 0.0)
            astarC = np.where(traceC & s_o, _clamp(s8o[:, 5].astype(np.float64) / z8o[:, 5]), 0.0)
            tr = np.ones(n)
        else:
            # Tierney & Mira: the stored scores of the two proposal objects (a
            # stale score where a proposal was invalid, as in the reference)
            astarB = np.minimum(1.0, y8[:, 5].astype(np.float64) / z8[:, 5])
            astarC = np.minimum(1.0, np.where(y_o, y8o[:, 5], y8[:, 5]).astype(np.float64) / np.where(z_o, z8o[:, 5], z8[:, 5]))
            dzy = Z["u"].astype(np.float64) - Y["u"]
            dxy = X["u"].astype(np.float64) - Y["u"]
            D = 2 * np.maximum(X["c"] + X["l"] - 1, 2)
            mask = np.arange(32)[None] < D[:, None]
            s2 = np.float64(np.float32(sigma)) ** 2
            with np.errstate(over="ignore"):
                tr = np.exp(-0.5 * (dzy ** 2 * mask).sum(1) / s2 + 0.5 * (dxy ** 2 * mask).sum(1) / s2)
    if integrator == 2:
        # the h2mc integrator's H2MC move: min(1, ratio), 0 for an invalid ratio
        with np.errstate(invalid="ignore"):
            a2B = np.where(second & ~_invalid(a2preB), np.minimum(1.0, a2preB), 0.0)
            a2C = np.where(second & ~_invalid(a2preC), np.minimum(1.0, a2preC), 0.0)
    else:
        a2B = np.where(second, _alpha2(dr_type, a2preB, astarB, tr, after["a1"]), 0.0)
        a2C = np.where(second, _alpha2(dr_type, a2preC, astarC, tr, after["a1"]), 0.0)
    s_both = traceB & s_gpu & s_o
    return {"n": n, "second": second, "aniso": aniso, "large": gen, "flags": flags,
            "y_valid_agree": y_gpu == y_o, "z_valid_agree": ~second | (z_gpu == z_o), "s_valid_agree": ~traceB | (s_gpu == s_o),
            "y_tri": _same_triangles(Y, Yo)[y_both], "z_tri": _same_triangles(Z, Zo)[z_both], "s_tri": _same_triangles(S, So)[s_both],
            "f_rel_y": frel(y8, y8o, y_both), "f_rel_z": frel(z8, z8o, z_both), "f_rel_s": frel(s8, s8o, s_both),
            "a1_gpu": a1g, "a1_B": a1B, "a1_C": a1C, "alpha2_gpu": a2g, "alpha2_B": a2B, "alpha2_C": a2C,
            "alpha2_pre_B": a2preB, "astar_B": astarB, "star_traced": traceB, "Y": Y, "Z": Z, "S": S, "Yo": Yo, "Zo": Zo, "So": So,
            "y8o": y8o, "z8o": z8o, "s8o": s8o, "y_both": y_both, "z_both": z_both, "s_both": s_both}


def rounding_baselines(O, ctx64, r):
    """The f error the loop's FP32 evaluation has against the oracle's FP64
    set beside two baselines on the same valid states with the same
    triangles: the oracle's float build (the reference as shipped computes
    in float) against FP64, and the device's FP64 fixed-topology evaluation
    (drmlt_eval_logf; ctx64 must differentiate the primal, quirk_chad_mis 0)
    against the oracle's FP64. Returns the three relative-error arrays."""
    nl, nt = len(O.sc["light_tris"]), len(O.sc["tri_p"])
    rec, o64, g32 = [], [], []
    for key, okey, both in (("Y", "y8o", "y_both"), ("Z", "z8o", "z_both"), ("S", "s8o", "s_both")):
        m = r[both] & (r[okey][:, 5] > 0) & _same_triangles(r[key], r[key + "o"])
        rec.append(r[key][m])
        o64.append(r[okey][m, 5].astype(np.float64))
    recs, o64 = np.concatenate(rec), np.concatenate(o64)
    _, o32 = oracle_eval_float(O, recs, nl, nt)
    lf64, ok = ctx64.eval_logf(recs, fp64=True)
    keep = (o32[:, 0] > 0) & (ok == 1)
    e_orc32 = np.abs(o32[keep, 5].astype(np.float64) - o64[keep]) / o64[keep]
    e_dev64 = np.abs(np.exp(lf64[keep]) - o64[keep]) / o64[keep]
    return e_orc32, e_dev64


def oracle_eval_float(O, recs, nlights, ntris):
    ok = well_formed(recs, nlights, ntris)
    out = recs.copy()
    c8 = np.zeros((len(recs), 8), np.float32)
    if ok.any():
        z = np.zeros((int(ok.sum()), 32), np.float32)
        out[ok], c8[ok] = O.perturb(recs[ok], z, z, dbl=False)
    return out, c8


def summary(r):
    d1B, d2B = np.abs(r["a1_gpu"] - r["a1_B"]), np.abs(r["alpha2_gpu"] - r["alpha2_B"])[r["second"]]
    d1C, d2C = np.abs(r["a1_gpu"] - r["a1_C"]), np.abs(r["alpha2_gpu"] - r["alpha2_C"])[r["second"]]
    f = np.concatenate([r["f_rel_y"], r["f_rel_z"], r["f_rel_s"]])
    pct = lambda a, q: float(np.percentile(a, q)) if len(a) else 0.0
    return {"chains": r["n"], "large": int(r["large"].sum()), "second": int(r["second"].sum()), "aniso": int(r["aniso"].sum()),
            "star_traced": int(r["star_traced"].sum()),
            "valid_agree": float(np.mean(r["y_valid_agree"] & r["z_valid_agree"] & r["s_valid_agree"])),
            "tri_equal": float(np.mean(np.concatenate([r["y_tri"], r["z_tri"], r["s_tri"]]))),
            "f_compared": int(len(f)), "f_rel_median": pct(f, 50), "f_rel_p99": pct(f, 99), "f_within_1e-4": float(np.mean(f <= 1e-4)),
            "a1_B_within_1e-5": float(np.mean(d1B <= 1e-5)), "a1_B_max": float(d1B.max()),
            "alpha2_B_within_1e-5": float(np.mean(d2B <= 1e-5)) if len(d2B) else 1.0, "alpha2_B_max": float(d2B.max()) if len(d2B) else 0.0,
            "a1_C_within_1e-5": float(np.mean(d1C <= 1e-5)), "alpha2_C_within_1e-5": float(np.mean(d2C <= 1e-5)) if len(d2C) else 1.0}


def worst(r, k=5):
    """The chains with the largest (B) disagreement, for the log."""
    out = []
    for key, mask in (("a1", np.ones(r["n"], bool)), ("alpha2", r["second"])):
        d = np.where(mask, np.abs(r[key + "_gpu"] - r[key + "_B"]), -1.0)
        for i in np.argsort(-d)[:k]:
            out.append({"what": key, "chain": int(i), "err": float(d[i]), "flags": int(r["flags"][i]), "gpu": float(r[key + "_gpu"][i]),
                        "B": float(r[key + "_B"][i]), "a2pre": float(r["alpha2_pre_B"][i]), "astar": float(r["astar_B"][i])})
    return out


def _lens_gaussian(c, l, ls, ss, g, h, ok, sigma, max_derv):
    """initLensGaussian (h2mc.cpp:466-499) from given lens derivatives: the
    lens Gaussian exists for c + l > 3 within the derivative depth, with the
    derivatives zeroed where lensScore <= 1e-15 or they are not finite;
    isotropic otherwise (gaussian.cpp:4-16)."""
    if not (c >= 1 and l >= 0 and c + l > 3 and c + l - 1 <= max_derv):
        return _iso_gaussian(2, sigma)
    g32, h32 = np.zeros(32), np.zeros((32, 32))
    if ls > 1e-15 and ok and np.isfinite(g).all() and np.isfinite(h).all():
        g32[:2] = g
        h32[:2, :2] = h
    m, _, inv, ld = orc.compute_gaussian(np.array([2]), np.array([ss]), sigma, g32[None], h32[None])
    return m[0], inv[0], float(ld[0])


def _lens_accept(off, gx, gy, ls_x, ls_y):
    # h2mc.cpp:516-523: min(1, N_y(-off) / N_x(off) * lensScore_y / lensScore_x)
    py = orc.gaussian_logpdf(2, off, gx[0][:2], gx[1][:32, :32], gx[2])
    px = orc.gaussian_logpdf(2, -off, gy[0][:2], gy[1][:32, :32], gy[2])
    with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
        a = float(np.exp(px - py) * ls_y / ls_x)
    a = a if np.isnan(a) else min(max(a, 0.0), 1.0)
    return 0.0 if _invalid(np.array([a]))[0] else a


def lens_accept_fd(O, X, x8, off, ls_x, sigma, max_derv=9, fd_step=2e-6):
    """The lens mutation's acceptance (h2mc.cpp:516-523) of states X (their
    contribution rows x8, lens scores ls_x) moved by screen offsets off, from
    the oracle: FP64 PerturbLensBidir to x's screen position plus off (mod 1),
    lens Gaussians from central differences (step fd_step) of the FP64 lens
    function at x's and the proposal's screen positions, the FP64 Gaussian
    log densities. Returns (a, proposals, their c9 rows)."""
    scr = x8[:, 6:8] + off
    scr = (scr - np.floor(scr)).astype(np.float32)
    Yo, c9 = O.perturb_lens(X, scr, dbl=True)
    y_o = c9[:, 0] > 0
    gyf, hyf = np.zeros((len(X), 2)), np.zeros((len(X), 2, 2))
    _, gxf, hxf = O.lens_grad_hess_fd(X, x8[:, 6:8], h=fd_step)
    if y_o.any():
        _, gyf[y_o], hyf[y_o] = O.lens_grad_hess_fd(Yo[y_o], c9[y_o, 6:8], h=fd_step)
    a = np.zeros(len(X))
    for k in np.nonzero(y_o)[0]:
        cx, lx = int(X["c"][k]), int(X["l"][k])
        gx = _lens_gaussian(cx, lx, float(ls_x[k]), float(x8[k, 5]), gxf[k], hxf[k], True, sigma, max_derv)
        gy = _lens_gaussian(cx, lx, float(c9[k, 8]), float(c9[k, 5]), gyf[k], hyf[k], True, sigma, max_derv)
        a[k] = _lens_accept(off[k].astype(np.float64), gx, gy, float(ls_x[k]), float(c9[k, 8]))
    return a, Yo, c9


def analyse_lens(O, ctx, before, after, lens_after, sigma, max_derv=9, fd_step=2e-6):
    """Per-state parity of the h2mc integrator's lens mutations
    (h2mc.cpp:463-530) in one iteration. For each chain that made one (flag
    bit 7), the lens offset the loop drew (read back in off[0..1]) moves the
    current state x to the screen position the loop perturbed to; then
    (A) the oracle's PerturbLensBidir in FP64 from x to that screen position
        against the loop's proposal y: validity, triangles, f, lens score;
    (B) the loop's acceptance a1 against the reference's formula computed
        here with the oracle's FP64 ComputeGaussian / GaussianLogPdf over the
        lens scores and lensVertexPos the loop held for x and y
        (drmlt_read_lens after the iteration, lens_after) and the device's
        lens derivatives about those positions (drmlt_lens_grad_hess);
    (C) the same acceptance with y's lens score from the oracle's FP64
        PerturbLensBidir and both lens Gaussians from central differences
        (step fd_step) of the oracle's FP64 lens function, as the oracle's
        own loop computes them; x's lens score is the one the loop held (the
        score of the move that made x, which for a lens mutation's result is
        that mutation's tracked score, h2mc.cpp:520-530)."""
    L = np.nonzero((after["flags"] & 128) != 0)[0]
    pb = before["perm"][L]
    pc, pf = pb & 3, (pb >> 2) & 3
    X, x8 = before["slots"][L, pc], before["contrib8"][L, pc]
    Y, y8 = after["slots"][L, pf], after["contrib8"][L, pf]
    off = after["off"][L, :2].astype(np.float32)
    a1 = after["a1"][L].astype(np.float64)
    y_dev = (after["flags"][L] & 32) != 0
    ls_x, ls_y = lens_after[0][L, pc].astype(np.float64), lens_after[0][L, pf].astype(np.float64)
    aC, Yo, c9 = lens_accept_fd(O, X, x8, off, ls_x, sigma, max_derv, fd_step)
    y_o = c9[:, 0] > 0
    both = y_dev & y_o
    fb = both & (c9[:, 5] > 0)
    f_rel = np.abs(y8[fb, 5].astype(np.float64) - c9[fb, 5]) / c9[fb, 5]
    known = ((lens_after[1][L, pc] & 1) != 0) & (~y_dev | ((lens_after[1][L, pf] & 1) != 0))
    lp_x, lp_y = lens_after[2][L, pc], lens_after[2][L, pf]
    ox, okx = ctx.lens_grad_hess(X, x8[:, 6:8], lp_x)
    oy, oky = np.zeros((len(L), 9), np.float32), np.zeros(len(L), np.int32)
    if y_dev.any():
        oy[y_dev], oky[y_dev] = ctx.lens_grad_hess(Y[y_dev], y8[y_dev, 6:8], lp_y[y_dev])
    lb = both & (c9[:, 8] > 0)
    ls_rel = np.abs(ls_y[lb] - c9[lb, 8]) / c9[lb, 8]
    aB = np.zeros(len(L))
    for k in range(len(L)):
        cx, lx = int(X["c"][k]), int(X["l"][k])
        h2 = lambda o: np.array([[o[6], o[7]], [o[7], o[8]]], np.float64)
        if y_dev[k]:
            gx = _lens_gaussian(cx, lx, ls_x[k], float(x8[k, 5]), ox[k, 4:6].astype(np.float64), h2(ox[k]), okx[k], sigma, max_derv)
            gy = _lens_gaussian(cx, lx, ls_y[k], float(y8[k, 5]), oy[k, 4:6].astype(np.float64), h2(oy[k]), oky[k], sigma, max_derv)
            aB[k] = _lens_accept(off[k].astype(np.float64), gx, gy, ls_x[k], ls_y[k])
    return {"n": len(after["perm"]), "lens": len(L), "y_valid_agree": y_dev == y_o, "score_known": known,
            "tri_equal": _same_triangles(Y[both], Yo[both]), "f_rel": f_rel, "ls_rel": ls_rel,
            "a1_gpu": a1, "a1_B": aB, "a1_C": aC, "chains": L, "X": X, "ls_x": ls_x, "ls_y": ls_y,
            "ls_x_fixed": ox[:, 0].astype(np.float64), "ls_y_fixed": oy[:, 0].astype(np.float64), "ls_y_oracle": c9[:, 8].astype(np.float64)}


def summary_lens(r):
    dB, dC = np.abs(r["a1_gpu"] - r["a1_B"]), np.abs(r["a1_gpu"] - r["a1_C"])
    pct = lambda a, q: float(np.percentile(a, q)) if len(a) else 0.0
    return {"chains": r["n"], "lens": r["lens"], "valid_agree": float(np.mean(r["y_valid_agree"])), "score_known": float(np.mean(r["score_known"])),
            "tri_equal": float(np.mean(r["tri_equal"])) if len(r["tri_equal"]) else 1.0,
            "f_compared": int(len(r["f_rel"])), "f_rel_median": pct(r["f_rel"], 50), "f_rel_p99": pct(r["f_rel"], 99),
            "ls_rel_median": pct(r["ls_rel"], 50), "ls_rel_p99": pct(r["ls_rel"], 99),
            "a1_B_within_1e-5": float(np.mean(dB <= 1e-5)), "a1_B_max": float(dB.max()) if len(dB) else 0.0,
            "a1_C_within_1e-3": float(np.mean(dC <= 1e-3)), "a1_C_median": pct(dC, 50), "a1_C_p99": pct(dC, 99),
            "accept_mean_gpu": float(r["a1_gpu"].mean()) if len(dB) else 0.0}


def worst_lens(r, k=8):
    """The lens chains with the largest (B) disagreement, for the log."""
    d = np.abs(r["a1_gpu"] - r["a1_B"])
    out = []
    for i in np.argsort(-d)[:k]:
        out.append({"chain": int(r["chains"][i]), "err": float(d[i]), "c": int(r["X"]["c"][i]), "l": int(r["X"]["l"][i]),
                    "gpu": float(r["a1_gpu"][i]), "B": float(r["a1_B"][i]), "C": float(r["a1_C"][i]),
                    "ls_x": float(r["ls_x"][i]), "ls_x_fixed": float(r["ls_x_fixed"][i]), "ls_y": float(r["ls_y"][i]),
                    "ls_y_fixed": float(r["ls_y_fixed"][i]), "ls_y_oracle": float(r["ls_y_oracle"][i])})
    return out

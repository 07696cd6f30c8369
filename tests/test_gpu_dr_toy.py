"""Delayed-rejection chains on an analytic 2-D target (a periodic mixture on
the unit torus), with the second-stage acceptance computed by the device's
drmlt_dr_alpha2_f64 (delayed.cpp:466-541). Both the Green-Mira form (ghost
proposal y* = z + (x - y)) and the Tierney-Mira form (first-stage
transition ratio) keep the target invariant when the second-stage ratio
f(z)/f(x) enters unclamped, so the chains' histogram must match the
target's bin masses: a dropped (1 - a*) or (1 - a1) factor, a transition
ratio the wrong way round or a sign error would bias it visibly. The
reference clamps the isotropic second-stage ratio f(z)/f(x) at 1 before the
correction (delayed.cpp:490 with mutation.cpp:197), which breaks detailed
balance slightly; that variant's deviation is reported alongside."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CENTERS = np.array([[0.3, 0.35], [0.72, 0.68], [0.2, 0.8]])
WIDTHS = np.array([0.06, 0.1, 0.045])
WEIGHTS = np.array([1.0, 0.6, 0.8])


def _f(u):
    d = u[:, None, :] - CENTERS[None]
    d -= np.round(d)  # nearest image on the torus
    r2 = (d ** 2).sum(-1)
    return 0.05 + (WEIGHTS * np.exp(-r2 / (2 * WIDTHS ** 2))).sum(-1)


def _bin_masses(k):
    g = (np.arange(512) + 0.5) / 512
    X, Y = np.meshgrid(g, g, indexing="ij")
    f = _f(np.stack([X.ravel(), Y.ravel()], 1)).reshape(512, 512)
    m = f.reshape(k, 512 // k, k, 512 // k).sum(axis=(1, 3))
    return m / m.sum()


def _run(kind, clamp, chains=40000, iters=500, burn=60, s1=0.06, s2=0.02, seed=1):
    from drmlt import capi

    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, (chains, 2))
    fx = _f(x)
    k = 8
    hist = np.zeros((k, k))
    for it in range(iters):
        y = (x + s1 * rng.normal(size=x.shape)) % 1.0
        fy = _f(y)
        a1 = np.minimum(1.0, fy / fx).astype(np.float32).astype(np.float64)
        acc1 = rng.uniform(size=chains) < a1
        rej = ~acc1
        z = (x[rej] + s2 * rng.normal(size=(rej.sum(), 2))) % 1.0
        fz = _f(z)
        a2 = fz / fx[rej]
        if clamp:
            a2 = np.minimum(1.0, a2)
        if kind == "green":
            ys = (z + (x[rej] - y[rej])) % 1.0
            a_star = np.minimum(1.0, _f(ys) / fz)
            tr = np.ones(len(z))
        else:
            a_star = np.minimum(1.0, fy[rej] / fz)
            dz = y[rej] - z
            dz -= np.round(dz)
            dx = y[rej] - x[rej]
            dx -= np.round(dx)
            tr = np.exp(-((dz ** 2).sum(1) - (dx ** 2).sum(1)) / (2 * s1 * s1))  # q1(z -> y) / q1(x -> y)
        rows = np.stack([np.full(len(z), 0.0 if kind == "green" else 1.0), a1[rej], a2,
                         a_star.astype(np.float32).astype(np.float64), tr, np.ones(len(z))], 1)
        alpha2 = capi.dr_alpha2(rows)
        acc2 = rng.uniform(size=len(z)) < alpha2
        xn, fn = x.copy(), fx.copy()
        xn[acc1], fn[acc1] = y[acc1], fy[acc1]
        idx = np.nonzero(rej)[0][acc2]
        xn[idx], fn[idx] = z[acc2], fz[acc2]
        x, fx = xn, fn
        if it >= burn:
            b = np.minimum((x * k).astype(int), k - 1)
            np.add.at(hist, (b[:, 0], b[:, 1]), 1)
    return hist / hist.sum()


@pytest.mark.parametrize("kind", ["green", "mira"])
def test_dr_chain_keeps_the_target(kind):
    target = _bin_masses(8)
    h = _run(kind, clamp=False)
    rel = np.abs(h - target) / target
    l1 = np.abs(h - target).sum()
    hc = _run(kind, clamp=True, seed=2)
    l1c = np.abs(hc - target).sum()
    print(f"{kind}: unclamped max rel {rel.max():.3f} L1 {l1:.4f}; reference clamp (Q2) L1 {l1c:.4f}", flush=True)
    # with the exact formula (restated on the CPU) these chains reach L1 0.0025-0.0054 and max
    # relative bin error 0.013-0.018 over seeds; with a planted missing (1 - a1) factor L1 0.008-0.010,
    # without (1 - a*) 0.016, with the transition ratio inverted 0.035
    assert l1 < 0.007 and rel.max() < 0.04

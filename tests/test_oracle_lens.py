"""Pins the oracle's lens bookkeeping and lens mutation (the h2mc
integrator's lens branch, h2mc.cpp:463-530, PerturbLensBidir
path.cpp:1914-2067, lensContrib in path.cpp:433-976 and 1430-1478).

The lens score is the part of the path throughput that changes when the
first camera vertex moves with the second one held: f0 cos0 cos1 f1 / d^2
over the edge between them (the two BSDFs include their cosines). On the
Lambertian room that is Lum(rho0 * rho1) / pi^2 |cos0| |cos1| / d^2, computed
here from the scene's triangles, the camera ray and the state's primary
samples alone; mistakes planted in the oracle's bookkeeping break it."""
import ctypes as C

import numpy as np
import pytest

import orc
from drmlt import scenes

LUM = np.array([0.212671, 0.715160, 0.072169])


@pytest.fixture(scope="module")
def room():
    sc = scenes.config1()
    return sc, orc.Scene(sc)


def _states(O, n, seed, lengths=(3, 4, 5)):
    rng = np.random.default_rng(seed)
    cls = [(P - l + 1, l) for P in lengths for l in range(0, P + 1) if not (P - l + 1 == 1 and l < 2)]
    cl = np.array([cls[i % len(cls)] for i in range(n)], np.int32)
    uni = rng.uniform(0, 1, (n, 96)).astype(np.float32)
    st, c8 = O.generate_subpath(cl, uni, dbl=True)
    return cl, uni, st, c8


def _hit(orig, d, tri):
    """Moller-Trumbore distance along d to triangle tri (3 x 3), or inf."""
    e1, e2 = tri[1] - tri[0], tri[2] - tri[0]
    p = np.cross(d, e2)
    det = e1 @ p
    if abs(det) < 1e-14:
        return np.inf
    s = orig - tri[0]
    u = (s @ p) / det
    q = np.cross(s, e1)
    v = (d @ q) / det
    t = (e2 @ q) / det
    if u < -1e-9 or v < -1e-9 or u + v > 1 + 1e-9 or t <= 0:
        return np.inf
    return t


def _camera_ray(O, sx, sy):
    org = np.zeros(3); d = np.zeros(3); mm = np.zeros(2)
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    orc.lib().orc_camera_ray(C.c_void_p(O.h), C.c_double(sx), C.c_double(sy), P(org), P(d), P(mm))
    return org, d


def _lens_closed_form(sc, O, rec):
    """Lum(rho0 rho1) / pi^2 |cos0| |cos1| / d^2 of the first two camera
    vertices of a state with c >= 3 on the Lambertian room, from the
    triangles, the camera ray through the screen position and the spherical
    direction in vertex 0's primary samples (absolute parameterisation)."""
    c, l = int(rec["c"]), int(rec["l"])
    k = 4 + 2 * (l - 1) if l > 1 else 0  # GetPathPos: light dims first
    sx, sy = float(rec["u"][k]), float(rec["u"][k + 1])
    tris = sc["tri_p"].reshape(-1, 3, 3).astype(np.float64)
    p0, p1 = int(rec["cprim"][0]), int(rec["cprim"][1])
    org, d = _camera_ray(O, sx, sy)
    v0 = org + _hit(org, d, tris[p0]) * d
    phi, th = 2 * np.pi * float(rec["u"][k + 2]), np.pi * float(rec["u"][k + 3])
    w = np.array([np.cos(phi) * np.sin(th), np.sin(phi) * np.sin(th), np.cos(th)])
    v1 = v0 + _hit(v0, w, tris[p1]) * w
    n0 = np.cross(tris[p0][1] - tris[p0][0], tris[p0][2] - tris[p0][0])
    n1 = np.cross(tris[p1][1] - tris[p1][0], tris[p1][2] - tris[p1][0])
    n0 /= np.linalg.norm(n0); n1 /= np.linalg.norm(n1)
    e = v1 - v0
    d2 = e @ e
    e /= np.sqrt(d2)
    rho0 = sc["mats"][sc["tri_mat"][p0]][1:4].astype(np.float64)
    rho1 = sc["mats"][sc["tri_mat"][p1]][1:4].astype(np.float64)
    return LUM @ (rho0 * rho1) / np.pi ** 2 * abs(n0 @ e) * abs(n1 @ e) / d2


def _closed_form_errors(sc, O, cl, st, c8):
    ls = O.lens_score(st, dbl=True)[:, 0]
    rels = []
    for i in range(len(st)):
        c, l = cl[i]
        if c8[i, 0] == 0 or c < 3 or (c == 3 and l == 0):
            continue
        with np.errstate(invalid="ignore", divide="ignore"):
            ref = _lens_closed_form(sc, O, st[i])
        if not np.isfinite(ref) or ref <= 0:
            continue
        rels.append(abs(ls[i] - ref) / ref)
    return np.array(rels)


def test_lens_score_closed_form(room):
    sc, O = room
    cl, _, st, c8 = _states(O, 5000, 1)
    rel = _closed_form_errors(sc, O, cl, st, c8)
    assert len(rel) > 400
    assert np.median(rel) < 1e-6 and rel.max() < 1e-5, (np.median(rel), rel.max())


@pytest.mark.parametrize("plant", [20, 21])
def test_lens_closed_form_catches_planted_mistakes(room, plant):
    """20: the 1/d^2 of the second camera vertex dropped; 21: the second
    vertex's adjoint BSDF factor dropped."""
    sc, O = room
    cl, _, st, c8 = _states(O, 800, 1)
    with orc.planted(plant):
        rel = _closed_form_errors(sc, O, cl, st, c8)
    assert np.median(rel) > 0.05


def test_lens_score_zero_where_the_reference_has_none(room):
    """lensScore needs a moved first camera vertex with a fixed neighbour:
    0 for the path-length-2 strategies c = 3, l = 0 (the second camera
    vertex is the light: HandleHitLight at camera depth 1), c = 2, l = 1
    (DirectLighting at camera depth 0) and c = 1, l = 2 (ConnectToCamera at
    light depth 0); positive on the Lambertian room for every strategy of
    length >= 3 (the states the chains hold, which is where the reference
    has lens derivative functions, path.cpp:3195-3197)."""
    sc, O = room
    cl, uni, st, c8 = _states(O, 3000, 2, lengths=(2, 3, 4))
    tracked = O.generate_subpath_lens(cl, uni, dbl=True)
    ok = c8[:, 0] > 0
    none = ((cl[:, 0] == 3) & (cl[:, 1] == 0)) | ((cl[:, 0] == 2) & (cl[:, 1] == 1)) | ((cl[:, 0] == 1) & (cl[:, 1] == 2))
    assert (ok & none).sum() > 50 and (ok & ~none).sum() > 500
    assert np.all(tracked[ok & none] == 0)
    assert np.all(tracked[ok & ~none] > 0)


def test_generation_and_reevaluation_agree(room):
    """The lens score GenerateSubpath tracks (BSDF-sampled directions) equals
    the one a fixed-topology re-evaluation of the stored state gives
    (directions from the spherical primary samples), as the device computes
    it; the states carry float primary samples."""
    sc, O = room
    cl, uni, st, c8 = _states(O, 4000, 3, lengths=(3, 4, 5, 6))
    tracked = O.generate_subpath_lens(cl, uni, dbl=True)
    fixed = O.lens_score(st, dbl=True)[:, 0]
    ok = (c8[:, 0] > 0) & (tracked > 0)
    rel = np.abs(fixed[ok] - tracked[ok]) / tracked[ok]
    assert ok.sum() > 1000
    assert np.median(rel) < 1e-6 and rel.max() < 1e-3, (np.median(rel), rel.max())


def test_lens_perturbation_to_own_screen_reproduces_the_state(room):
    sc, O = room
    cl, uni, st, c8 = _states(O, 3000, 4)
    tracked = O.generate_subpath_lens(cl, uni, dbl=True)
    sel = np.where((c8[:, 0] > 0) & (tracked > 0))[0]
    out, c9 = O.perturb_lens(st[sel], c8[sel, 6:8], dbl=True)
    assert np.all(c9[:, 0] > 0)
    assert np.all(out["cprim"] == st[sel]["cprim"]) and np.all(out["lgprim"] == st[sel]["lgprim"])
    lens_rel = np.abs(c9[:, 8] - tracked[sel]) / tracked[sel]
    ls_rel = np.abs(c9[:, 4] - c8[sel, 4]) / c8[sel, 4]
    assert np.median(lens_rel) < 1e-6 and lens_rel.max() < 1e-3
    assert np.median(ls_rel) < 1e-6 and ls_rel.max() < 1e-3
    # the contribution's screen position comes back (c = 1: the projected point)
    assert np.abs(c9[:, 6:8] - c8[sel, 6:8]).max() < 1e-4


def test_lens_perturbation_moves_the_first_vertex(room):
    """A small screen offset moves the first camera vertex (c >= 2) along the
    new camera ray and keeps every other triangle; the last light vertex for
    c = 1 is the one the camera ray finds; the screen position of the
    contribution follows the offset."""
    sc, O = room
    cl, uni, st, c8 = _states(O, 3000, 5)
    tracked = O.generate_subpath_lens(cl, uni, dbl=True)
    sel = np.where((c8[:, 0] > 0) & (tracked > 0))[0]
    rng = np.random.default_rng(6)
    scr = np.mod(c8[sel, 6:8] + rng.normal(0, 2e-3, (len(sel), 2)), 1.0).astype(np.float32)
    out, c9 = O.perturb_lens(st[sel], scr, dbl=True)
    v = c9[:, 0] > 0
    assert v.mean() > 0.8
    assert np.abs(c9[v, 6:8] - scr[v]).max() < 1e-4
    cam = cl[sel, 0] >= 2
    same_rest = (out["cprim"][:, 1:] == st[sel]["cprim"][:, 1:]).all(1) & (out["lgprim"] == st[sel]["lgprim"]).all(1)
    assert same_rest[v & cam].all()


def test_lens_derivatives_are_converged(room):
    """Central differences of the lens function at two steps agree, and the
    Hessian is symmetric: the lens Gaussian's derivatives are converged."""
    sc, O = room
    cl, uni, st, c8 = _states(O, 600, 7)
    tracked = O.generate_subpath_lens(cl, uni, dbl=True)
    sel = np.where((c8[:, 0] > 0) & (tracked > 0))[0][:150]
    l1, g1, h1 = O.lens_grad_hess_fd(st[sel], c8[sel, 6:8], h=2e-6)
    l2, g2, h2 = O.lens_grad_hess_fd(st[sel], c8[sel, 6:8], h=1e-5)
    fin = np.isfinite(l1) & np.isfinite(g1).all(1) & np.isfinite(g2).all(1)
    assert fin.mean() > 0.9
    np.testing.assert_allclose(np.log(tracked[sel][fin]), l1[fin], rtol=1e-4, atol=1e-3)
    gs = np.abs(g1[fin]).max(1, keepdims=True) + 1.0
    assert np.median(np.abs(g1[fin] - g2[fin]) / gs) < 1e-4
    hs = np.abs(h2[fin]).reshape(-1, 4).max(1) + 1.0
    assert np.median(np.abs(h1[fin] - h2[fin]).reshape(-1, 4).max(1) / hs) < 1e-2
    assert np.all(h1[fin][:, 0, 1] == h1[fin][:, 1, 0])


def test_h2mc_with_lens_converges_to_bdpt():
    """The oracle's h2mc loop at lensperturbprob 0.3 and 1 converges to an
    independent bidirectional estimate, makes lens mutations at the expected
    share, and accepts them."""
    from test_oracle_dr import _blocks, _small_cfg1

    sc = _small_cfg1()
    s = orc.Scene(sc)
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
    for lp in (0.3, 1.0):
        film, st, _ = s.render(seeds, c8, lf, b, iters, threads=8, seed=7, opts={"h2mc": 1, "lensperturbprob": lp})
        large, _, iso, _, _, _, h2, _, lens, lens_acc = (int(x) for x in st)
        assert large + iso + h2 + lens == iters * 256
        assert abs(iso / (iso + h2 + lens) - 0.2) < 0.01
        # every small step not mixed to isotropic has a lens score on this scene
        # except the c = 3, l = 0 / c = 2, l = 1 / c = 1, l = 2 states
        assert lens / (h2 + lens) > 0.5 * lp and lens_acc > 0.5 * lens
        img = film / (iters * 256 / (W * H))
        rb, ib = _blocks(ref), _blocks(img)
        mask = rb > 0.05 * rb.max()
        rel = np.abs(ib[mask] - rb[mask]) / rb[mask]
        assert np.median(rel) < 0.08, (lp, np.median(rel))
        assert ib.sum() == pytest.approx(rb.sum(), rel=0.05)


@pytest.mark.parametrize("which", ["cfg1", "cfg2_small"])
def test_lens_acceptance_helper_matches_the_oracle_loop(which, request):
    """The lens-mutation acceptance that the GPU per-state parity test
    computes from the oracle's functions (step_parity.lens_accept_fd:
    PerturbLensBidir, central-difference lens Gaussians, Gaussian log
    densities) equals the oracle loop's own LensMutateAt for the same states
    and offsets (an independent restatement checked against the oracle's
    C++), and disagrees once the offset's sign is flipped."""
    import step_parity as sp

    sc = request.getfixturevalue(which)
    O = orc.Scene(sc)
    cl, uni, st, c8 = _states(O, 2000 if which == "cfg1" else 8000, 11)
    ls = O.lens_score(st, dbl=True)[:, 0]
    sel = np.where((c8[:, 0] > 0) & (ls > 0))[0][:200]
    assert len(sel) > 40
    X, x8, lsx = st[sel], c8[sel], ls[sel].astype(np.float64)
    rng = np.random.default_rng(5)
    sigma = 0.01
    off = (rng.normal(size=(len(sel), 2)) * sigma).astype(np.float32)
    a_helper, _, c9 = sp.lens_accept_fd(O, X, x8, off, lsx, sigma)
    a_oracle = O.lens_accept(X, x8, off, sigma)
    valid = c9[:, 0] > 0
    assert valid.mean() > 0.8 and (a_oracle[valid] > 0).mean() > 0.5
    err = np.abs(a_helper - a_oracle)
    print(which, "helper vs oracle loop: max", err.max(), "median", np.median(err))
    assert err.max() < 1e-4 and np.median(err) < 1e-6, (err.max(), np.median(err))
    # the offset's sign flipped
    a_planted, _, _ = sp.lens_accept_fd(O, X, x8, -off, lsx, sigma)
    assert np.abs(a_planted - a_oracle)[valid].max() > 1e-3

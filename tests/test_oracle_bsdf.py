"""Pins the oracle's BSDFs (src/lambertian.cpp, src/phong.cpp,
src/roughdielectric.cpp) with closed forms and identities that a dropped
term, a wrong sign or a swapped argument would break:
Lambertian closed form; pdf normalisation by Monte Carlo; sample weight ==
eval / pdf at the sampled direction; reverse pdf == pdf with arguments
swapped; reciprocity of f. The values of the glossy and refractive lobes are
pinned by what the mathematics fixes: the Phong albedo and pdf integral at
normal incidence (exactly Kd + Ks and 1), the rough dielectric's smooth
limit (Fresnel F(0), 1 - F(0), the radiance scaling 1 / eta'^2), energy
conservation at grazing incidence and the generalized reciprocity of
refraction; each of these fails on a mistake planted in the oracle."""
import numpy as np
import pytest

import orc
from drmlt import scenes

RNG = np.random.default_rng(1234)


def rand_dir_hemi(n, normal):
    d = RNG.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    s = np.sign(d @ normal)
    return d * s[:, None]


MATS = {
    "lambert": scenes.lambert((0.5, 0.3, 0.7)),
    "phong": scenes.phong((0.2, 0.3, 0.1), (0.4, 0.4, 0.4), 30.0),
    "phong_sharp": scenes.phong((0.05, 0.05, 0.05), (0.5, 0.5, 0.5), 800.0),
    "diel": scenes.rough_dielectric(0.3, 1.5, 1.0),
}


def test_lambert_closed_form():
    m = MATS["lambert"]
    n = np.array([0.0, 0.0, 1.0])
    for _ in range(200):
        wi, wo = rand_dir_hemi(2, n)
        out = orc.bsdf_eval(m, wi, n, wo)
        cwo, cwi = wo[2], wi[2]
        if cwo < 1e-4 or cwi < 1e-4:
            continue
        np.testing.assert_allclose(out[:3], m[1:4].astype(np.float64) * cwo / np.pi, rtol=1e-12)
        np.testing.assert_allclose([out[3], out[4], out[5]], [cwo, cwo / np.pi, cwi / np.pi], rtol=1e-12)


@pytest.mark.parametrize("name", ["lambert", "phong", "phong_sharp", "diel"])
def test_sample_weight_is_eval_over_pdf(name):
    m = MATS[name]
    n = np.array([0.3, -0.2, 0.93]); n /= np.linalg.norm(n)
    checked = 0
    for k in range(400):
        wi = rand_dir_hemi(1, n)[0]
        r0, r1, disc = RNG.uniform(size=3)
        ok, wo, s = orc.bsdf_sample(m, wi, n, r0, r1, disc)
        if not ok:
            continue
        e = orc.bsdf_eval(m, wi, n, wo)
        if e[4] <= 0 or not np.any(e[:3] > 0):
            continue
        if name == "diel":
            # the sampled lobe is chosen with probability F (reflection) or 1-F
            # (refraction): the sample weight omits that factor, eval/pdf includes it
            # (eval rebuilds the half vector from wi, wo; the sampled micronormal
            # carries sin = sqrt(1 - cos^2) rounding, hence 1e-5)
            np.testing.assert_allclose(s[4], e[4], rtol=1e-5)  # same pdf
            np.testing.assert_allclose(s[3], e[3], rtol=1e-9)  # same cosWo
            np.testing.assert_allclose(s[5], e[5], rtol=1e-5)  # same reverse pdf
        else:
            np.testing.assert_allclose(s[:3], e[:3] / e[4], rtol=1e-6)
            np.testing.assert_allclose(s[4], e[4], rtol=1e-6)
            np.testing.assert_allclose(s[5], e[5], rtol=1e-6)
        checked += 1
    assert checked > 100


@pytest.mark.parametrize("name", ["lambert", "phong"])
def test_reverse_pdf_and_reciprocity(name):
    m = MATS[name]
    n = np.array([0.0, 0.0, 1.0])
    for _ in range(300):
        wi, wo = rand_dir_hemi(2, n)
        a = orc.bsdf_eval(m, wi, n, wo)
        b = orc.bsdf_eval(m, wo, n, wi)
        if a[4] <= 0 or b[4] <= 0:
            continue
        np.testing.assert_allclose(a[5], b[4], rtol=1e-10)
        # f(wi, wo) = contrib / cosWo is symmetric
        np.testing.assert_allclose(a[:3] / a[3], b[:3] / b[3], rtol=1e-10)


@pytest.mark.parametrize("name,expect", [("lambert", 1.0), ("phong", None), ("diel", None)])
def test_pdf_integrates_to_at_most_one(name, expect):
    """Monte Carlo over the sphere: E[pdf(wo) / (1/4pi)] = integral of pdf <= 1."""
    m = MATS[name]
    n = np.array([0.0, 0.0, 1.0])
    wi = np.array([0.3, 0.1, 0.95]); wi /= np.linalg.norm(wi)
    N = 2_000_000  # own generator: the estimate does not depend on test order
    d = np.random.default_rng(77).normal(size=(N, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    integral = orc.bsdf_eval_batch(m, wi, n, d)[:, 4].sum() / N * 4 * np.pi
    if expect is not None:
        assert abs(integral - expect) < 0.03
    else:
        assert 0.5 < integral < 1.03


def test_dielectric_refraction_obeys_snell():
    m = MATS["diel"]
    n = np.array([0.0, 0.0, 1.0])
    seen = 0
    for _ in range(500):
        wi = rand_dir_hemi(1, n)[0]
        ok, wo, s = orc.bsdf_sample(m, wi, n, *RNG.uniform(size=2), 0.999)  # uDisc > F: refraction
        if not ok or wo[2] >= 0:
            continue
        # the half vector of a refraction pair satisfies wi + eta * wo parallel to -h
        h = wi + 1.5 * wo
        h /= np.linalg.norm(h)
        # tangential components of wi and wo around h obey sin_i = eta sin_t
        si = np.linalg.norm(np.cross(wi, h)); st = np.linalg.norm(np.cross(wo, h))
        np.testing.assert_allclose(si, 1.5 * st, rtol=1e-6, atol=1e-9)
        seen += 1
    assert seen > 50


def test_phong_roughness_threshold():
    # src/phong.cpp:148-155: sqrt(2 / (2 + 4 n)) on the specular lobe, 1 on the diffuse lobe
    m = MATS["phong_sharp"]
    ks = m[11]
    assert orc.lib().orc_roughness(orc.P(np.ascontiguousarray(m)), orc.C.c_double(ks * 0.5)) == pytest.approx(np.sqrt(2 / (2 + 4 * 800.0)))
    assert orc.lib().orc_roughness(orc.P(np.ascontiguousarray(m)), orc.C.c_double(min(1.0, ks + 0.01))) == 1.0


# ---------------------------------------------------------------- values the mathematics fixes
# The checks below pin the values of the glossy and refractive lobes, not only
# their consistency: each is shown to fail on a mistake planted in the oracle
# (tests/orc.py planted; the mistakes are listed in oracle/orc_core.h).

def _gl_panels(a, b, first, ratio, npt):
    """Gauss-Legendre nodes and weights on [a, b] over panels that grow
    geometrically from width `first` at a."""
    edges = [a]
    h = first
    while edges[-1] + h < b:
        edges.append(edges[-1] + h)
        h *= ratio
    edges.append(b)
    x, w = np.polynomial.legendre.leggauss(npt)
    e = np.asarray(edges)
    lo, hi = e[:-1], e[1:]
    return (0.5 * (hi - lo)[:, None] * (x[None] + 1) + lo[:, None]).ravel(), (0.5 * (hi - lo)[:, None] * w[None]).ravel()


N_UP = np.array([0.0, 0.0, 1.0])


def phong_normal_incidence(m):
    """Directional albedo (RGB) and pdf integral at wi = n. The mirror
    direction is the normal, so the integrands depend on mu = cos(theta_o)
    only: 2 pi x Gauss-Legendre in mu (exact for the cos^n polynomials up to
    the reference's 1e-10 cut-offs)."""
    x, w = np.polynomial.legendre.leggauss(3000)
    mu, w = 0.5 * (x + 1), 0.5 * w
    wo = np.stack([np.sqrt(1 - mu ** 2), np.zeros_like(mu), mu], -1)
    out = orc.bsdf_eval_batch(m, N_UP, N_UP, wo)
    return 2 * np.pi * (out[:, :3] * w[:, None]).sum(0), 2 * np.pi * (out[:, 4] * w).sum()


def phong_normal_incidence_ok(m):
    alb, pint = phong_normal_incidence(m)
    expect = m[1:4].astype(np.float64) + m[4:7].astype(np.float64)
    return bool(np.allclose(alb, expect, rtol=1e-6, atol=0) and abs(pint - 1.0) < 1e-6)


@pytest.mark.parametrize("expo", [10.0, 100.0, 555.0, 1000.0])
@pytest.mark.parametrize("kd", [(0.0, 0.0, 0.0), (0.2, 0.1, 0.3)])
def test_phong_albedo_and_pdf_at_normal_incidence(expo, kd):
    """The modified Phong lobe (n + 2)/(2 pi) cos^n(alpha) x cos(theta_o)
    integrates to exactly 1 when the mirror direction is the normal, and its
    sampling density (n + 1)/(2 pi) cos^n(alpha) to exactly 1 (Lafortune &
    Willems 1994): at wi = n the albedo is Kd + Ks and the pdf (lobe mixed with
    the cosine lobe by KsWeight) integrates to 1."""
    assert phong_normal_incidence_ok(scenes.phong(kd, (0.4, 0.5, 0.6), expo))


@pytest.mark.parametrize("theta_deg", [30.0, 60.0, 80.0])
def test_phong_albedo_at_most_kd_plus_ks(theta_deg):
    """Away from normal incidence the horizon clips the lobe: the albedo
    stays at or below Kd + Ks (2-D quadrature over the hemisphere)."""
    m = scenes.phong((0.2, 0.1, 0.3), (0.4, 0.5, 0.6), 50.0)
    th = np.radians(theta_deg)
    wi = np.array([np.sin(th), 0.0, np.cos(th)])
    t, wt = _gl_panels(0.0, np.pi / 2, 1e-3, 1.2, 12)
    nphi = 720
    phi = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    T, Ph = np.meshgrid(t, phi, indexing="ij")
    wo = np.stack([np.sin(T) * np.cos(Ph), np.sin(T) * np.sin(Ph), np.cos(T)], -1).reshape(-1, 3)
    f = orc.bsdf_eval_batch(m, wi, N_UP, wo)[:, :3].reshape(T.shape + (3,))
    alb = (f * (np.sin(T) * wt[:, None])[..., None]).sum((0, 1)) * 2 * np.pi / nphi
    assert np.all(alb <= m[1:4].astype(np.float64) + m[4:7] + 1e-6)
    assert np.all(alb >= m[1:4] * (1 - 1e-6))  # the cosine lobe alone integrates to Kd


@pytest.mark.parametrize("plant", [1, 2])
def test_phong_check_fails_on_planted_normalisation(plant):
    """A lobe normalised by (n + 1) instead of (n + 2), or a pdf by (n + 2)
    instead of (n + 1), fails the normal-incidence check."""
    m = scenes.phong((0.2, 0.1, 0.3), (0.4, 0.5, 0.6), 1000.0)
    assert phong_normal_incidence_ok(m)
    with orc.planted(plant):
        assert not phong_normal_incidence_ok(m)


def diel_normal_incidence(m, inside, adjoint=True):
    """Reflectance R and transmittance T at wi = +n (outside) or -n (inside):
    the integrands depend on theta_o only; 2 pi x Gauss-Legendre panels in
    theta refined towards the specular directions."""
    s = -1.0 if inside else 1.0
    wi = np.array([0.0, 0.0, s])
    t, w = _gl_panels(0.0, np.pi / 2, 1e-7, 1.5, 16)
    wr = np.stack([np.sin(t), np.zeros_like(t), s * np.cos(t)], -1)
    wt = np.stack([np.sin(t), np.zeros_like(t), -s * np.cos(t)], -1)
    R = 2 * np.pi * (orc.bsdf_eval_batch(m, wi, N_UP, wr, adjoint)[:, 0] * np.sin(t) * w).sum()
    T = 2 * np.pi * (orc.bsdf_eval_batch(m, wi, N_UP, wt, adjoint)[:, 0] * np.sin(t) * w).sum()
    return R, T


def diel_normal_incidence_ok(m, inside):
    """As the roughness goes to 0 the rough dielectric becomes the smooth
    interface: reflectance F(0) = ((eta - 1)/(eta + 1))^2 (Fresnel at normal
    incidence, the same from either side), transmittance 1 - F(0) for
    importance (the adjoint lobe) and (1 - F(0)) / eta'^2 for radiance, where
    eta' is the index ratio across the interface seen from wi (Veach 1997,
    5.2: radiance scales by the squared index ratio)."""
    eta = float(m[9])
    f0 = ((eta - 1) / (eta + 1)) ** 2
    R, T = diel_normal_incidence(m, inside)
    Rn, Tn = diel_normal_incidence(m, inside, adjoint=False)
    eta_ = 1.0 / eta if inside else eta
    return abs(R - f0) < 1e-5 and abs(T - (1 - f0)) < 1e-5 and abs(Rn - R) < 1e-12 and abs(Tn - T / eta_ ** 2) < 1e-6 * Tn


@pytest.mark.parametrize("alpha", [0.005, 0.01])
@pytest.mark.parametrize("inside", [False, True])
def test_dielectric_smooth_limit(alpha, inside):
    assert diel_normal_incidence_ok(scenes.rough_dielectric(alpha, 1.5, 1.0), inside)


def diel_furnace(m, theta_deg):
    """R + T (importance) at incidence theta over the whole sphere (2-D
    quadrature: Gauss-Legendre panels in theta_o, midpoint rule in phi)."""
    th = np.radians(theta_deg)
    wi = np.array([np.sin(th), 0.0, np.cos(th)])
    t1, w1 = _gl_panels(0.0, np.pi / 2, 1e-4, 1.3, 12)
    t, w = np.concatenate([t1, np.pi - t1[::-1]]), np.concatenate([w1, w1[::-1]])
    nphi = 512
    phi = (np.arange(nphi) + 0.5) * 2 * np.pi / nphi
    T, Ph = np.meshgrid(t, phi, indexing="ij")
    wo = np.stack([np.sin(T) * np.cos(Ph), np.sin(T) * np.sin(Ph), np.cos(T)], -1).reshape(-1, 3)
    f = orc.bsdf_eval_batch(m, wi, N_UP, wo, True)[:, 0].reshape(T.shape)
    return float((f * np.sin(T) * w[:, None]).sum() * 2 * np.pi / nphi)


@pytest.mark.parametrize("alpha", [0.3, 0.5])
@pytest.mark.parametrize("theta_deg", [80.0, 85.0])
def test_dielectric_white_furnace_at_grazing_incidence(alpha, theta_deg):
    """A microfacet interface loses energy to shadowing and masking, never
    gains it: R + T <= 1 for importance at every incidence. Without the
    shadowing-masking term the integral exceeds 1 at grazing incidence (the
    projected microfacet area grows as 1 / cos(theta_i))."""
    a = diel_furnace(scenes.rough_dielectric(alpha, 1.5, 1.0), theta_deg)
    assert 0.9 < a <= 1.0 + 1e-3


def test_dielectric_reciprocity_and_reverse_pdf():
    """Reflection: f(wi, wo) = f(wo, wi). Transmission: f(wi, wo) / eta_o^2 =
    f(wo, wi) / eta_i^2 for the importance lobe (Walter et al. 2007, eq. 21;
    Veach 1997, 5.2), i.e. f(wi, wo) = eta'^2 f(wo, wi) with eta' the index
    ratio seen from wi. The reverse pdf of (wi, wo) is the pdf of (wo, wi).
    IOR 2, exact in float with its inverse 0.5: with 1.5 the material's
    float 1/eta differs from the exact inverse by 3e-8, which next to the
    critical angle (where 1 - F has unbounded slope) moves 1 - F by up to 1%
    between the two directions."""
    m = scenes.rough_dielectric(0.3, 2.0, 1.0)
    d = RNG.normal(size=(4000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    wi, wo = d[:2000], d[2000:]
    a = orc.bsdf_eval_batch(m, wi, N_UP, wo, True)
    b = np.stack([orc.bsdf_eval(m, wo[k], N_UP, wi[k], adjoint=True) for k in range(2000)])
    live = (a[:, 0] > 1e-12) & (b[:, 0] > 1e-12)
    assert live.sum() > 500
    fa = a[live, 0] / np.abs(a[live, 3])
    fb = b[live, 0] / np.abs(b[live, 3])
    eta_i = np.where(wi[live, 2] > 0, 2.0, 0.5)
    trans = wi[live, 2] * wo[live, 2] < 0
    ratio = np.where(trans, eta_i ** 2, 1.0)
    np.testing.assert_allclose(fa, ratio * fb, rtol=1e-5)
    np.testing.assert_allclose(a[live, 5], b[live, 4], rtol=1e-5)
    assert trans.sum() > 100 and (~trans).sum() > 100


@pytest.mark.parametrize("plant", [3, 4, 5, 6, 7])
def test_dielectric_checks_fail_on_planted_mistakes(plant):
    """Each planted mistake -- the Fresnel factor dropped from reflection (3),
    G dropped (4), eta'^2 dropped from transmission (5), the radiance
    scaling dropped (6), F in place of 1 - F (7) -- fails the smooth-limit or
    the grazing furnace check."""
    m = scenes.rough_dielectric(0.01, 1.5, 1.0)
    m_rough = scenes.rough_dielectric(0.3, 1.5, 1.0)
    with orc.planted(plant):
        caught = not (diel_normal_incidence_ok(m, False) and diel_normal_incidence_ok(m, True))
        if plant == 4:
            caught = caught or diel_furnace(m_rough, 80.0) > 1.0 + 1e-3
    assert caught

"""Pins the oracle's BSDFs (src/lambertian.cpp, src/phong.cpp,
src/roughdielectric.cpp) with closed forms and identities that a dropped
term, a wrong sign or a swapped argument would break:
Lambertian closed form; pdf normalisation by Monte Carlo; sample weight ==
eval / pdf at the sampled direction; reverse pdf == pdf with arguments
swapped; reciprocity of f."""
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
    N = 60000
    d = RNG.normal(size=(N, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    tot = 0.0
    for k in range(N):
        tot += orc.bsdf_eval(m, wi, n, d[k])[4]
    integral = tot / N * 4 * np.pi
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

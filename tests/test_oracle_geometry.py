"""Pins the oracle's ray queries, camera and area-light sampling
(src/scene.cpp:80-111, src/trianglemesh.cpp:30-79, 236-363,
src/camera.cpp:38-74) with brute force and round trips."""
import numpy as np

import orc


def _rays(n, rng, center, spread):
    org = rng.uniform(-spread, spread, (n, 3)) + np.asarray(center)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return org.astype(np.float32), d.astype(np.float32)


def test_bvh_equals_brute_force(cfg2_small):
    rng = np.random.default_rng(5)
    bvh = orc.Scene(cfg2_small, use_bvh=True)
    brute = orc.Scene(cfg2_small, use_bvh=False)
    org, d = _rays(4000, rng, (0, 0.4, 0), 2.0)
    tmin = np.full(len(org), 1e-3, np.float32)
    tmax = np.full(len(org), np.inf, np.float32)
    p1, t1 = bvh.trace(org, d, tmin, tmax)
    p2, t2 = brute.trace(org, d, tmin, tmax)
    assert (p1 >= 0).sum() > 1000
    np.testing.assert_array_equal(p1, p2)
    np.testing.assert_array_equal(t1[p1 >= 0], t2[p2 >= 0])
    tmax2 = rng.uniform(0.05, 3.0, len(org)).astype(np.float32)
    np.testing.assert_array_equal(bvh.occluded(org, d, tmin, tmax2), brute.occluded(org, d, tmin, tmax2))


def test_hit_matches_ray_plane(cfg1):
    """The t of the closest hit is the ray-plane distance of the named triangle."""
    rng = np.random.default_rng(6)
    s = orc.Scene(cfg1)
    org, d = _rays(3000, rng, (0, 1, 0), 0.9)
    prim, t = s.trace(org, d, np.full(len(org), 1e-3, np.float32), np.full(len(org), np.inf, np.float32))
    P = cfg1["tri_p"].astype(np.float64).reshape(-1, 3, 3)
    ok = prim >= 0
    assert ok.mean() > 0.7  # room open only at the front
    p0 = P[prim[ok], 0]
    nrm = np.cross(P[prim[ok], 1] - p0, P[prim[ok], 2] - p0)
    tp = np.einsum("ij,ij->i", p0 - org[ok], nrm) / np.einsum("ij,ij->i", d[ok], nrm)
    np.testing.assert_allclose(t[ok], tp, rtol=2e-4, atol=1e-5)


def test_camera_rays_and_projection(cfg1):
    """SamplePrimary / ProjectPoint (camera.cpp:38-74): the centre ray looks at the
    target, the horizontal field of view between screen x = 0 and 1 is fov, and
    projecting a point on any primary ray returns that ray's screen position."""
    L = orc.lib()
    s = orc.Scene(cfg1)
    cam = cfg1["cam"].astype(np.float64)
    pos, tgt, fov = cam[0:3], cam[3:6], cam[9]
    o = np.zeros(3); d = np.zeros(3); mm = np.zeros(2)
    L.orc_camera_ray(orc.C.c_void_p(s.h), orc.C.c_double(0.5), orc.C.c_double(0.5), orc.P(o), orc.P(d), orc.P(mm))
    np.testing.assert_allclose(o, pos, atol=1e-12)
    np.testing.assert_allclose(d, (tgt - pos) / np.linalg.norm(tgt - pos), atol=1e-12)
    dl = np.zeros(3); dr = np.zeros(3)
    L.orc_camera_ray(orc.C.c_void_p(s.h), orc.C.c_double(0.0), orc.C.c_double(0.5), orc.P(o), orc.P(dl), orc.P(mm))
    L.orc_camera_ray(orc.C.c_void_p(s.h), orc.C.c_double(1.0), orc.C.c_double(0.5), orc.P(o), orc.P(dr), orc.P(mm))
    np.testing.assert_allclose(np.degrees(np.arccos(dl @ dr)), fov, rtol=1e-9)
    rng = np.random.default_rng(8)
    for _ in range(200):
        sx, sy = rng.uniform(0.01, 0.99, 2)
        L.orc_camera_ray(orc.C.c_void_p(s.h), orc.C.c_double(sx), orc.C.c_double(sy), orc.P(o), orc.P(d), orc.P(mm))
        p = o + d * rng.uniform(0.5, 10.0)
        out = np.zeros(2)
        assert L.orc_project(orc.C.c_void_p(s.h), orc.P(p), orc.P(out)) == 1
        np.testing.assert_allclose(out, [sx, sy], atol=1e-9)


def test_area_sampling_round_trip(cfg1):
    """Triangle area sampling and its inverse (trianglemesh.cpp:236-363): the
    parameters of a sampled point are recovered, samples stay on the triangle,
    their mean is the centroid and the pdf is 1 / light area."""
    L = orc.lib()
    s = orc.Scene(cfg1)
    first, cnt = cfg1["light_tris"][0]
    P = cfg1["tri_p"][first:first + cnt].astype(np.float64).reshape(-1, 3, 3)
    area = sum(0.5 * np.linalg.norm(np.cross(t[1] - t[0], t[2] - t[0])) for t in P)
    rng = np.random.default_rng(7)
    pts = []
    for _ in range(4000):
        u0, u1 = rng.uniform(size=2)
        pos = np.zeros(3); nrm = np.zeros(3); pdf = np.zeros(1); back = np.zeros(2)
        L.orc_sample_tri(orc.C.c_void_p(s.h), 0, int(first), orc.C.c_double(u0), orc.C.c_double(u1), orc.P(pos), orc.P(nrm), orc.P(pdf))
        L.orc_sample_param(orc.C.c_void_p(s.h), int(first), orc.P(pos), orc.P(back))
        np.testing.assert_allclose(back, [u0, u1], atol=1e-7)
        np.testing.assert_allclose(pdf[0], 1.0 / area, rtol=1e-6)
        pts.append(pos)
    pts = np.array(pts)
    t = P[0]
    np.testing.assert_allclose(pts.mean(0), t.mean(0), atol=0.01)

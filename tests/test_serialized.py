"""The reference's serialized subpath layout (the arguments of its derivative
functions, PathFuncDerv, src/path.h:145-149): the oracle's restatement of
Serialize (src/path.cpp:2303-2391) and the library's host-side reader
(drmlt_serialized_to_rec, the decoder behind drmlt_path_func_derv). CPU only:
no device call is made."""
import numpy as np
import pytest

import orc
from drmlt import capi

# the (camera, light) vertex counts of every strategy up to path length 6
CLS = [(c, l) for c in range(1, 8) for l in range(0, 7) if c + l >= 3 and c + l - 1 <= 6 and not (c == 1 and l < 2)]


def states(sc, n_per=300, seed=1):
    """Valid oracle states of every (c, l) of CLS: (c, l) -> records."""
    O = orc.Scene(sc)
    rng = np.random.default_rng(seed)
    out = {}
    for c, l in CLS:
        cl = np.tile(np.array([[c, l]], np.int32), (n_per, 1))
        st, c8 = O.generate_subpath(cl, rng.uniform(0, 1, (n_per, 160)).astype(np.float32), dbl=False)
        keep = st[c8[:, 0] > 0]
        if len(keep):
            out[(c, l)] = keep
    return O, out


def test_layout_sizes():
    """Record sizes read off the reference: triangle-mesh shape 1 + 1 +
    2 (3 + 3 + 3) + 2 (3 + 3 + 3) + 1 + 6 + 1 = 46 (trianglemesh.cpp:3-11),
    BSDF max(4, 9, 10) = 10, light max(point 7, area 1 + 46 + 3 = 50, env 56)
    = 56; vertParams for maxDervDepth 9 (drsmallstep.cpp:104-105) 18 x 46 +
    18 x 10 + 56 + 36 + 18 + 5 = 1123; primary 2 max(c + l - 1, 2) + 1."""
    O = orc.Scene(__import__("drmlt.scenes", fromlist=["config1"]).config1())
    p, v = O.serialize(np.zeros(1, capi.PATH_REC), max_depth=9)
    assert p.shape[1] == 2 * 17 + 1 and v.shape[1] == 1123
    assert len(O.serialize_scene()) == 38


@pytest.mark.parametrize("which", ["cfg1", "cfg2_small"])
def test_round_trip(which, request):
    """Serialized by the oracle, read back by the library: every field the
    layout carries comes back bit-exact, the others take the record defaults."""
    sc = request.getfixturevalue(which)
    O, by = states(sc)
    assert len(by) >= 12, sorted(by)
    for (c, l), st in by.items():
        primary, vert = O.serialize(st)
        got = capi.serialized_to_rec(sc, c, l, st["screen"], primary, vert)
        D = 2 * max(c + l - 1, 2)
        assert (got["c"] == c).all() and (got["l"] == l).all()
        np.testing.assert_array_equal(got["u"][:, :D], st["u"][:, :D])
        np.testing.assert_array_equal(got["screen"], st["screen"])
        np.testing.assert_array_equal(got["cprim"][:, :c - 1], st["cprim"][:, :c - 1])
        np.testing.assert_array_equal(got["lgprim"][:, :max(l - 1, 0)], st["lgprim"][:, :max(l - 1, 0)])
        if l >= 1:
            np.testing.assert_array_equal(got["light"], st["light"])
            np.testing.assert_array_equal(got["lprim"], st["lprim"])
        else:
            assert (got["light"] == -1).all() and (got["lprim"] == -1).all()
        # bsdfDiscrete and roulette weights where serialized (not the last
        # camera vertex's discrete sample, not each subpath's last weight)
        np.testing.assert_array_equal(got["cdisc"][:, :c - 2], st["cdisc"][:, :c - 2])
        np.testing.assert_array_equal(got["crr"][:, :c - 2], st["crr"][:, :c - 2])
        np.testing.assert_array_equal(got["ldisc"][:, :max(l - 1, 0)], st["ldisc"][:, :max(l - 1, 0)])
        np.testing.assert_array_equal(got["lrr"][:, :max(l - 2, 0)], st["lrr"][:, :max(l - 2, 0)])
        assert (got["cprim"][:, c - 1:] == -1).all() and (got["crr"][:, max(c - 2, 0):] == 1).all()


def test_known_offsets(cfg2_small):
    """Positions in the layout read off Serialize for a (c, l) = (3, 3) path:
    lens vertex 0-2 (the camera position), light pick probability 3, the
    light record 4-59 (type 1 = area, its shape at 5-50: type 0, isMoving 0,
    p0, p1 - p0, p2 - p0, ...; the emission at 51-53), the first light surface vertex at 60 (shape),
    106 (bsdfDiscrete), 107 (useAbsoluteParam), 108 (BSDF type), then its
    roulette weight at 118; the second (last) light vertex at 119 with no
    weight after its BSDF record (119 + 46 + 2 + 10 = 177); the camera
    vertices from 177."""
    sc = cfg2_small
    O, by = states(sc, n_per=6000)
    st = by[(3, 3)][:20]
    assert len(st) >= 5
    primary, vert = O.serialize(st)
    tp = sc["tri_p"].reshape(-1, 9)
    mats = sc["mats"].reshape(-1, 16)
    emit = sc["light_emit"].reshape(-1, 4)
    cam = sc["cam"]
    for i in range(len(st)):
        v = vert[i]
        np.testing.assert_allclose(v[0:3], cam[0:3], rtol=1e-6)
        np.testing.assert_allclose(v[3], emit[st["light"][i], 3] / emit[:, 3].sum(), rtol=1e-6)
        assert v[4] == 1.0 and v[5] == 0.0 and v[6] == 0.0
        lp = st["lprim"][i]
        np.testing.assert_array_equal(v[7:10], tp[lp, 0:3])
        np.testing.assert_array_equal(v[10:13], tp[lp, 3:6] - tp[lp, 0:3])
        np.testing.assert_array_equal(v[51:54], emit[st["light"][i], 0:3])
        g0 = st["lgprim"][i, 0]
        assert v[60] == 0.0
        np.testing.assert_array_equal(v[62:65], tp[g0, 0:3])
        assert v[106] == st["ldisc"][i, 0]
        assert v[108] == mats[sc["tri_mat"][g0], 0]
        assert v[118] == st["lrr"][i, 0]
        g1 = st["lgprim"][i, 1]
        np.testing.assert_array_equal(v[121:124], tp[g1, 0:3])
        assert v[165] == st["ldisc"][i, 1]
        c0 = st["cprim"][i, 0]
        np.testing.assert_array_equal(v[179:182], tp[c0, 0:3])
        assert v[177 + 46] == st["cdisc"][i, 0]
        assert primary[i, 0] == 0.0
        np.testing.assert_array_equal(primary[i, 1:11], st["u"][i, :10])


def test_mismatches_are_rejected(cfg2_small):
    """A triangle the scene does not hold, a BSDF record of the wrong type, a
    buffer shorter than the path and a light record on a non-emissive
    triangle are each an argument error."""
    sc = cfg2_small
    O, by = states(sc, n_per=6000)
    st = by[(3, 3)][:4]
    primary, vert = O.serialize(st)
    ok = capi.serialized_to_rec(sc, 3, 3, st["screen"], primary, vert)
    assert len(ok) == 4
    bad = vert.copy()
    bad[0, 62] = np.nextafter(bad[0, 62], np.float32(1e9))  # one bit of a vertex position
    with pytest.raises(capi.DrmltError, match="not in the context's scene"):
        capi.serialized_to_rec(sc, 3, 3, st["screen"], primary, bad)
    bad = vert.copy()
    bad[1, 108] = (bad[1, 108] + 1) % 3
    with pytest.raises(capi.DrmltError, match="BSDF"):
        capi.serialized_to_rec(sc, 3, 3, st["screen"], primary, bad)
    with pytest.raises(capi.DrmltError, match="shorter"):
        capi.serialized_to_rec(sc, 3, 3, st["screen"], primary, vert[:, :150])
    bad = vert.copy()
    bad[2, 5:51] = bad[2, 60:106]  # the light record names a surface triangle
    with pytest.raises(capi.DrmltError, match="not emissive"):
        capi.serialized_to_rec(sc, 3, 3, st["screen"], primary, bad)
    with pytest.raises(capi.DrmltError):
        capi.serialized_to_rec(sc, 1, 1, st["screen"], primary, vert)  # no such strategy


def test_material_from_serialized(cfg2_small):
    """The BSDF records Serialize writes (the oracle's restatement of
    lambertian.cpp:10-13, phong.cpp:13-19, roughdielectric.cpp:13-20) convert
    back to the scene's own material records, for every BSDF type the scene
    has; an unknown type is an argument error."""
    sc = cfg2_small
    O, by = states(sc, n_per=6000)
    st = by[(3, 3)]
    _, vert = O.serialize(st)
    mats = sc["mats"].reshape(-1, 16)
    used = {0: [0, 1, 2, 3, 12], 1: [0, 1, 2, 3, 4, 5, 6, 7, 11, 12], 2: [0, 1, 2, 3, 4, 5, 6, 8, 9, 10, 12]}
    seen = set()
    for i in range(len(st)):
        for off, prim in ((108, st["lgprim"][i, 0]), (167, st["lgprim"][i, 1])):
            want = mats[sc["tri_mat"][prim]]
            got = capi.material_from_serialized(vert[i, off:off + 10], two_sided=bool(want[12]))
            k = used[int(want[0])]
            np.testing.assert_array_equal(got[k], want[k])
            seen.add(int(want[0]))
    assert len(seen) >= 2, seen
    with pytest.raises(capi.DrmltError):
        capi.material_from_serialized(np.array([7] + [0] * 9, np.float32))

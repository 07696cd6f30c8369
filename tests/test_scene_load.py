"""Scene files in the reference's XML format (drmlt_scene_load: ParseScene,
src/parsescene.cpp:136-613, with the OBJ and .serialized loaders), on the
host without a GPU.

Pinned by: round trips of the procedural benchmark scenes through exported
XML + OBJ files (the arrays come back as exported); transforms against
matrices built here from their definitions (transform.cpp:4-90 and the
composition order of ParseTransform, parsescene.cpp:83-134); the
reference's angle-weighted vertex normals on meshes without normals,
including its obtuse-corner weight; Mitsuba .serialized files (v3 and v4,
float and double, several shapes) written here from the published format;
<dpt> keys against options_from_dpt; and the rejection of what the device
path cannot render."""
import numpy as np
import pytest

from drmlt import capi, scenes
from scene_files import export_scene, write_obj, write_serialized


def test_config1_round_trip(tmp_path):
    sc = scenes.config1()
    p = export_scene(sc, tmp_path, film_name="out.exr")
    ld = capi.load_scene(p)
    assert np.array_equal(ld["tri_p"], sc["tri_p"])
    np.testing.assert_allclose(ld["tri_n"], sc["tri_n"], atol=1e-6)  # vn are normalised on read
    assert np.array_equal(ld["tri_mat"], sc["tri_mat"]) and np.array_equal(ld["tri_light"], sc["tri_light"])
    np.testing.assert_allclose(ld["mats"], sc["mats"], rtol=1e-6, atol=1e-7)
    assert np.array_equal(ld["light_tris"], sc["light_tris"]) and np.array_equal(ld["light_emit"], sc["light_emit"])
    cam, c0 = ld["cam"], np.asarray(sc["cam"], np.float32)
    assert np.array_equal(cam[0:3], c0[0:3]) and np.array_equal(cam[9:14], c0[9:14])
    d0 = (c0[3:6] - c0[0:3]) / np.linalg.norm(c0[3:6] - c0[0:3])
    np.testing.assert_allclose(cam[3:6] - cam[0:3], d0, atol=1e-6)  # the look direction
    assert ld["output"] == str(tmp_path / "out.exr") and ld["integrator"] == "drmlt"
    want = capi.options_from_dpt(sc["options"])
    for name, _ in capi.Options._fields_:
        assert getattr(ld["opt"], name) == pytest.approx(getattr(want, name)), name


def test_config5_materials_round_trip(tmp_path):
    """Rough dielectric and Lambertian records with sphere meshes (shading
    normals per vertex), and the acceptanceMap / type options."""
    sc = scenes.config5(grid_res=16)
    ld = capi.load_scene(export_scene(sc, tmp_path))
    assert np.array_equal(ld["tri_p"], sc["tri_p"])
    np.testing.assert_allclose(ld["mats"], sc["mats"], rtol=1e-6, atol=1e-7)
    assert ld["opt"].acceptance_map == 1 and ld["opt"].max_depth == 12 and ld["opt"].num_chains == 1048576
    assert ld["output"] == str(tmp_path / "image.exr")


def _ref_transform(ops):
    """ParseTransform's composition: each listed transform applies after the earlier ones."""
    M = np.eye(4)
    for op, a in ops:
        if op == "scale":
            T = np.diag([*a, 1.0])
        elif op == "translate":
            T = np.eye(4)
            T[:3, 3] = a
        elif op == "rotate":
            ang, ax = np.radians(a[0]), np.asarray(a[1:], float) / np.linalg.norm(a[1:])
            K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
            T = np.eye(4)
            T[:3, :3] = np.cos(ang) * np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * np.outer(ax, ax)  # Rodrigues
        elif op == "matrix":
            T = np.asarray(a, float).reshape(4, 4)
        M = T @ M
    return M


def _one_shape_scene(folder, shape_xml, extra=""):
    (folder / "light.obj").write_text("v -0.1 1 -0.1\nv 0.1 1 -0.1\nv 0.1 1 0.1\nv -0.1 1 0.1\nf 1 2 3 4\n")
    xml = f"""<scene version="0.5.0">
  <sensor type="perspective"><float name="fov" value="40"/>
    <transform name="toWorld"><lookat origin="0 0.5 3" target="0,0.5,0" up="0 1 0"/></transform>
    <film type="hdrfilm"><integer name="width" value="64"/><integer name="height" value="48"/></film></sensor>
  <bsdf type="diffuse" id="white"><rgb name="reflectance" value="0.8"/></bsdf>
  {shape_xml}
  <shape type="obj"><string name="filename" value="light.obj"/>
    <emitter type="area"><rgb name="radiance" value="5, 6, 7"/></emitter></shape>
  {extra}
</scene>"""
    p = folder / "scene.xml"
    p.write_text(xml)
    return p


def test_shape_transforms_compose_as_the_reference(tmp_path):
    rng = np.random.default_rng(3)
    P = rng.uniform(-1, 1, (5, 9)).astype(np.float32)
    N = rng.normal(size=(5, 9)).astype(np.float32)
    for t in range(5):
        for k in range(3):
            N[t, 3 * k:3 * k + 3] /= np.linalg.norm(N[t, 3 * k:3 * k + 3])
    write_obj(tmp_path / "m.obj", P, N)
    mat = [1.0, 0.1, 0.0, 0.3, 0.0, 0.9, 0.2, -0.1, 0.05, 0.0, 1.1, 0.4, 0, 0, 0, 1]
    ops = [("scale", (2.0, 0.5, 1.5)), ("rotate", (30.0, 0.3, 1.0, -0.2)), ("translate", (0.1, -0.2, 0.7)), ("matrix", mat)]
    xml = ('<shape type="obj"><string name="filename" value="m.obj"/><transform name="toWorld">'
           '<scale x="2" y="0.5" z="1.5"/><rotate x="0.3" y="1" z="-0.2" angle="30"/><translate x="0.1" y="-0.2" z="0.7"/>'
           f'<matrix value="{" ".join(str(x) for x in mat)}"/></transform><ref id="white"/></shape>')
    ld = capi.load_scene(_one_shape_scene(tmp_path, xml))
    M = _ref_transform(ops)
    ph = np.concatenate([P.reshape(-1, 3), np.ones((15, 1))], 1) @ M.T
    want_p = (ph[:, :3] / ph[:, 3:]).reshape(5, 9)
    np.testing.assert_allclose(ld["tri_p"][:5], want_p, rtol=1e-5, atol=1e-5)
    # normals: the transposed inverse (not renormalised by the loader)
    want_n = (N.reshape(-1, 3) @ np.linalg.inv(M)[:3, :3]).reshape(5, 9)
    np.testing.assert_allclose(ld["tri_n"][:5], want_n, rtol=1e-4, atol=1e-5)
    # the quad light: two triangles (0 1 2) and (0 2 3), one area light over them
    assert ld["light_tris"].tolist() == [[5, 2]] and ld["light_emit"].tolist() == [[5, 6, 7, 1]]
    assert ld["tri_light"].tolist() == [-1] * 5 + [0, 0]
    q = ld["tri_p"][5:].reshape(2, 3, 3)
    assert np.array_equal(q[1, 0], q[0, 0]) and np.array_equal(q[1, 1], q[0, 2])
    assert ld["mats"][0].tolist()[:4] == pytest.approx([0, 0.8, 0.8, 0.8])
    assert ld["cam"][10:12].tolist() == [64, 48] and ld["cam"][9] == 40


def _angle_weighted_normals(V, F, reference_quirk=True):
    """Nelson Max's angle-weighted vertex normals; the reference's obtuse
    corner weight is (pi - 2) asin(|u + v| / 2) instead of the angle."""
    out = np.zeros_like(V)
    for f in F:
        n = np.cross(V[f[1]] - V[f[0]], V[f[2]] - V[f[0]])
        n /= np.linalg.norm(n)
        for i in range(3):
            a = V[f[(i + 1) % 3]] - V[f[i]]
            b = V[f[(i + 2) % 3]] - V[f[i]]
            a, b = a / np.linalg.norm(a), b / np.linalg.norm(b)
            if a @ b < 0:
                w = (np.pi - 2) * np.arcsin(0.5 * np.linalg.norm(a + b)) if reference_quirk else np.arccos(a @ b)
            else:
                w = 2 * np.arcsin(0.5 * np.linalg.norm(b - a))
            out[f[i]] += n * w
    return out / np.linalg.norm(out, axis=1, keepdims=True)


def test_normals_of_meshes_without_normals(tmp_path):
    """A folded strip whose shared vertices have obtuse corners on one side:
    the loader's normals follow the reference's weights (and differ from
    true angle weights there)."""
    # vertex 2 sits near the middle of edge 0-1: its corner in triangle 0 is obtuse (about 157 degrees)
    V = np.array([[0, 0, 0], [2, 0, 0], [1, 0.2, 0.1], [1.5, 1.0, -0.6], [-1.5, 0.2, 0.9]], np.float64)
    F = np.array([[0, 1, 2], [1, 3, 2], [0, 2, 4]])
    lines = ["v %.17g %.17g %.17g" % tuple(v) for v in V] + ["f %d %d %d" % tuple(f + 1) for f in F]
    (tmp_path / "fold.obj").write_text("\n".join(lines) + "\n")
    ld = capi.load_scene(_one_shape_scene(tmp_path, '<shape type="obj"><string name="filename" value="fold.obj"/><ref id="white"/></shape>'))
    got = ld["tri_n"][:3].reshape(3, 3, 3)
    want = _angle_weighted_normals(V, F)
    true_angles = _angle_weighted_normals(V, F, reference_quirk=False)
    for t, f in enumerate(F):
        np.testing.assert_allclose(got[t], want[f], atol=2e-6)
    assert np.abs(want - true_angles).max() > 1e-3  # the check can tell the two weightings apart
    # a planar fan: every vertex normal is the plane's
    Vp = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [-1, 0.5, 0]], np.float64)
    lines = ["v %.17g %.17g %.17g" % tuple(v) for v in Vp] + ["f 1 2 3 4", "f 1 4 5"]
    (tmp_path / "fan.obj").write_text("\n".join(lines) + "\n")
    ld = capi.load_scene(_one_shape_scene(tmp_path, '<shape type="obj"><string name="filename" value="fan.obj"/><ref id="white"/></shape>'))
    np.testing.assert_allclose(ld["tri_n"][:3].reshape(-1, 3), np.tile([0, 0, 1.0], (9, 1)), atol=1e-7)


@pytest.mark.parametrize("version,double", [(4, False), (4, True), (3, False)])
def test_serialized_meshes(tmp_path, version, double):
    rng = np.random.default_rng(version + 2 * double)
    meshes = []
    for k in range(3):
        V = rng.uniform(-1, 1, (6, 3))
        F = rng.integers(0, 6, (4, 3))
        m = {"v": V, "f": F, "name": f"part{k}"}
        if k == 1:
            Nn = rng.normal(size=(6, 3))
            m["n"] = Nn / np.linalg.norm(Nn, axis=1, keepdims=True)
        meshes.append(m)
    write_serialized(tmp_path / "m.serialized", meshes, version, double)
    for idx in range(3):
        xml = (f'<shape type="serialized"><string name="filename" value="m.serialized"/>'
               f'<integer name="shapeIndex" value="{idx}"/><transform name="toWorld"><translate x="0.5"/></transform>'
               '<ref id="white"/></shape>')
        ld = capi.load_scene(_one_shape_scene(tmp_path, xml))
        m = meshes[idx]
        V = np.asarray(m["v"], np.float64 if double else np.float32).astype(np.float32) + np.float32([0.5, 0, 0])
        np.testing.assert_allclose(ld["tri_p"][:4], V[m["f"]].reshape(4, 9), rtol=1e-6, atol=1e-6)
        if "n" in m:
            np.testing.assert_allclose(ld["tri_n"][:4], np.asarray(m["n"], np.float32)[m["f"]].reshape(4, 9), atol=1e-6)
        else:
            nonzero = np.linalg.norm(ld["tri_n"][:4].reshape(-1, 3), axis=1)
            assert np.all((np.abs(nonzero - 1) < 1e-5) | (nonzero == 0))


def test_dpt_options(tmp_path):
    dpt = {"integrator": "drmlt", "uniformmixingprobability": 0.35, "lensperturbprob": 0.0, "spp": 64, "numinitsamples": 5000, "largestepprob": 0.25, "mindepth": 2, "maxdepth": 7,
           "directspp": 4, "perturbstddev": 0.02, "roughnessthreshold": 0.05, "numchains": 2048, "seedoffset": 3,
           "uselightcoordinatesampling": False, "largestepmultiplexed": True, "type": "mira", "anisoperturbprob": 0.7,
           "anisoperturbstddev": 0.004, "acceptanceMap": True, "reportintervalseconds": 30}
    sc = scenes.config1()
    ld = capi.load_scene(export_scene(sc, tmp_path, dpt=dpt))
    want = capi.options_from_dpt({k: v for k, v in dpt.items() if k != "reportintervalseconds"})
    for name, _ in capi.Options._fields_:
        assert getattr(ld["opt"], name) == pytest.approx(getattr(want, name)), name
    assert ld["report_interval_seconds"] == 30 and ld["opt"].type == 1 and ld["opt"].use_light_coordinate_sampling == 0


@pytest.mark.parametrize("extra,shape,msg", [
    ('<emitter type="point"><point name="position" x="0" y="1" z="0"/></emitter>', "", "outside a shape"),
    ('<texture type="bitmap" id="t"><string name="filename" value="a.png"/></texture>', "", "textures"),
    ("", '<shape type="sphere"><ref id="white"/></shape>', "Invalid shape"),
    ("", '<shape type="obj"><string name="filename" value="missing.obj"/></shape>', "cannot open"),
    ("", '<shape type="obj"><string name="filename" value="light.obj"/><ref id="nope"/></shape>', "ref not found"),
    ("", '<shape type="obj"><string name="filename" value="light.obj"/><bsdf type="conductor"/></shape>', "Unknown BSDF"),
    ("", '<shape type="obj"><string name="filename" value="light.obj"/><bsdf type="diffuse">'
         '<texture name="reflectance" type="bitmap"/></bsdf></shape>', "textures"),
])
def test_unsupported_inputs_are_rejected(tmp_path, extra, shape, msg):
    p = _one_shape_scene(tmp_path, shape, extra)
    with pytest.raises(capi.DrmltError, match=msg):
        capi.load_scene(p)


def test_malformed_files_are_rejected(tmp_path):
    (tmp_path / "bad.xml").write_text("<scene><sensor></scene>")
    with pytest.raises(capi.DrmltError, match="XML parse error"):
        capi.load_scene(tmp_path / "bad.xml")
    with pytest.raises(capi.DrmltError, match="cannot open"):
        capi.load_scene(tmp_path / "none.xml")
    (tmp_path / "flip.xml").write_text(_one_shape_scene(tmp_path, "").read_text().replace(
        '<lookat origin="0 0.5 3"', '<scale x="-1"/><lookat origin="0 0.5 3"'))
    with pytest.raises(capi.DrmltError, match="look-at"):
        capi.load_scene(tmp_path / "flip.xml")


@pytest.mark.parametrize("name,code", [("drmlt", 0), ("mcmc", 1), ("h2mc", 2), ("mc", -1)])
def test_integrator_option(tmp_path, name, code):
    ld = capi.load_scene(export_scene(scenes.config1(), tmp_path, dpt={"integrator": name}))
    assert ld["integrator"] == name and ld["opt"].integrator == code
    assert capi.options_from_dpt({"integrator": name}).integrator == code


def test_corrupt_serialized_is_rejected(tmp_path):
    write_serialized(tmp_path / "m.serialized", [{"v": np.eye(3), "f": [[0, 1, 2]]}] * 2)
    raw = bytearray((tmp_path / "m.serialized").read_bytes())
    raw[-4:] = (10 ** 9).to_bytes(4, "little")  # a mesh count far beyond the file
    (tmp_path / "bad.serialized").write_bytes(bytes(raw))
    xml = ('<shape type="serialized"><string name="filename" value="bad.serialized"/>'
           '<integer name="shapeIndex" value="1"/><ref id="white"/></shape>')
    with pytest.raises(capi.DrmltError, match="offset table"):
        capi.load_scene(_one_shape_scene(tmp_path, xml))
    raw = bytearray((tmp_path / "m.serialized").read_bytes())
    raw[10:14] = b"\xff\xff\xff\xff"  # inside the first zlib stream
    (tmp_path / "bad2.serialized").write_bytes(bytes(raw))
    with pytest.raises(capi.DrmltError):
        capi.load_scene(_one_shape_scene(tmp_path, xml.replace("bad.serialized", "bad2.serialized").replace('value="1"', 'value="0"')))


def test_serialized_header_beyond_the_stream_is_rejected(tmp_path):
    """Vertex and triangle counts that a stream of this length cannot hold
    (deflate expands at most 1032:1) are refused before any allocation."""
    import struct
    import zlib
    body = struct.pack("<I", 0x1000) + b"mesh\0" + struct.pack("<QQ", 1 << 30, 1 << 30) + bytes(64)
    (tmp_path / "huge.serialized").write_bytes(struct.pack("<HH", 0x041C, 4) + zlib.compress(body) + struct.pack("<QI", 0, 1))
    xml = ('<shape type="serialized"><string name="filename" value="huge.serialized"/>'
           '<integer name="shapeIndex" value="0"/><ref id="white"/></shape>')
    with pytest.raises(capi.DrmltError, match="exceeds the stream"):
        capi.load_scene(_one_shape_scene(tmp_path, xml))

"""Writes scene dicts (drmlt.scenes) as scene files in the reference's XML
format with OBJ meshes, and Mitsuba .serialized meshes, for the loader
tests. Test infrastructure."""
import struct
import zlib
from pathlib import Path

import numpy as np


def _f(x):
    return "%.9g" % float(np.float32(x))


def _rgb(v):
    return ", ".join(_f(x) for x in v)


def bsdf_xml(rec, ident=None):
    rec = np.asarray(rec, np.float32)
    t = int(rec[0])
    idattr = f' id="{ident}"' if ident else ""
    if t == 0:
        inner = f'<bsdf type="diffuse"><rgb name="reflectance" value="{_rgb(rec[1:4])}"/></bsdf>'
    elif t == 1:
        inner = (f'<bsdf type="phong"><rgb name="diffuseReflectance" value="{_rgb(rec[1:4])}"/>'
                 f'<rgb name="specularReflectance" value="{_rgb(rec[4:7])}"/>'
                 f'<float name="exponent" value="{_f(rec[7])}"/></bsdf>')
    else:
        inner = (f'<bsdf type="roughdielectric"><float name="alpha" value="{_f(rec[8])}"/>'
                 f'<float name="intIOR" value="{_f(rec[9])}"/><float name="extIOR" value="1"/>'
                 f'<rgb name="specularReflectance" value="{_rgb(rec[1:4])}"/>'
                 f'<rgb name="specularTransmittance" value="{_rgb(rec[4:7])}"/></bsdf>')
    if rec[12] > 0:
        return f'<bsdf type="twosided"{idattr}>{inner}</bsdf>'
    return inner.replace("<bsdf ", f"<bsdf{idattr} ", 1) if ident else inner


def write_obj(path, P, N=None):
    """Triangles (n x 9 positions, optional n x 9 normals) as an OBJ file with
    three vertices per triangle."""
    lines = []
    for p in np.asarray(P, np.float32).reshape(-1, 3):
        lines.append("v " + " ".join(_f(x) for x in p))
    if N is not None:
        for n in np.asarray(N, np.float32).reshape(-1, 3):
            lines.append("vn " + " ".join(_f(x) for x in n))
    for t in range(len(P)):
        a, b, c = 3 * t + 1, 3 * t + 2, 3 * t + 3
        lines.append(f"f {a}//{a} {b}//{b} {c}//{c}" if N is not None else f"f {a} {b} {c}")
    Path(path).write_text("\n".join(lines) + "\n")


def export_scene(sc, folder, dpt=None, film_name=None):
    """The scene dict as <folder>/scene.xml plus one OBJ per run of triangles
    sharing material and light. Returns the XML path."""
    folder = Path(folder)
    folder.mkdir(parents=True, exist_ok=True)
    cam = np.asarray(sc["cam"], np.float32)
    out = ['<?xml version="1.0" encoding="utf-8"?>', "<!-- exported for the loader tests -->", '<scene version="0.5.0">']
    opts = dict(sc.get("options", {}))
    if dpt:
        opts.update(dpt)
    out.append("  <dpt>")
    for k, v in opts.items():
        if isinstance(v, bool):
            v = "true" if v else "false"
        kind = "string" if isinstance(v, str) else ("integer" if isinstance(v, int) else "float")
        out.append(f'    <{kind} name="{k}" value="{v}"/>')
    out.append("  </dpt>")
    film = f'<string name="filename" value="{film_name}"/>' if film_name else ""
    out.append(f'  <sensor type="perspective"><float name="fov" value="{_f(cam[9])}"/>'
               f'<float name="nearClip" value="{_f(cam[12])}"/><float name="farClip" value="{_f(cam[13])}"/>'
               f'<transform name="toWorld"><lookat origin="{_rgb(cam[0:3])}" target="{_rgb(cam[3:6])}" up="{_rgb(cam[6:9])}"/>'
               f'</transform><sampler type="independent"/><film type="hdrfilm"><integer name="width" value="{int(cam[10])}"/>'
               f'<integer name="height" value="{int(cam[11])}"/>{film}</film></sensor>')
    for k, rec in enumerate(sc["mats"]):
        out.append("  " + bsdf_xml(rec, f"mat{k}"))
    mat, light = np.asarray(sc["tri_mat"]), np.asarray(sc["tri_light"])
    t, n, s = 0, len(mat), 0
    while t < n:
        e = t
        while e < n and mat[e] == mat[t] and light[e] == light[t]:
            e += 1
        name = f"shape{s}.obj"
        write_obj(folder / name, sc["tri_p"][t:e], sc["tri_n"][t:e])
        em = ""
        if light[t] >= 0:
            em = f'<emitter type="area"><rgb name="radiance" value="{_rgb(sc["light_emit"][light[t]][:3])}"/></emitter>'
        out.append(f'  <shape type="obj"><string name="filename" value="{name}"/><ref id="mat{mat[t]}"/>{em}</shape>')
        t, s = e, s + 1
    out.append("</scene>")
    p = folder / "scene.xml"
    p.write_text("\n".join(out) + "\n")
    return p


def write_serialized(path, meshes, version=4, double=False):
    """Mitsuba's serialized format: per mesh the format id 0x041C, the
    version, a zlib stream (flags, [v4: name], vertex and triangle counts,
    positions, [normals], indices), then the offset table and mesh count."""
    blob, offsets = b"", []
    for m in meshes:
        V = np.asarray(m["v"], np.float64 if double else np.float32)
        F = np.asarray(m["f"], np.uint32)
        flags = (0x2000 if double else 0x1000) | (0x0001 if "n" in m else 0)
        body = struct.pack("<I", flags)
        if version == 4:
            body += m.get("name", "mesh").encode() + b"\0"
        body += struct.pack("<QQ", len(V), len(F)) + V.tobytes()
        if "n" in m:
            body += np.asarray(m["n"], V.dtype).tobytes()
        body += F.tobytes()
        offsets.append(len(blob))
        blob += struct.pack("<HH", 0x041C, version) + zlib.compress(body)
    if version == 4:
        tail = b"".join(struct.pack("<Q", o) for o in offsets)
    else:
        tail = b"".join(struct.pack("<I", o) for o in offsets)
    Path(path).write_bytes(blob + tail + struct.pack("<I", len(meshes)))

"""A reader for single-part scanline OpenEXR files with HALF/FLOAT channels and
NO or ZIP compression, written from the published file layout (magic
20000630, version 2, attribute list, chunk offset table, chunks of
(y, size, data); ZIP: zlib-deflated bytes after delta coding (each byte minus
its predecessor, plus 128) of the even-indexed bytes followed by the
odd-indexed ones). Test infrastructure: it checks the library's EXR writer."""
import struct
import zlib

import numpy as np

PIXEL = {0: np.uint32, 1: np.float16, 2: np.float32}
LINES = {0: 1, 2: 1, 3: 16}


def _cstr(b, i):
    j = b.index(b"\0", i)
    return b[i:j].decode(), j + 1


def read_header(b):
    magic, version = struct.unpack_from("<II", b, 0)
    assert magic == 20000630, hex(magic)
    assert version & 0xFF == 2 and not version & 0x200, hex(version)  # version 2, scanline single part
    i, attrs = 8, {}
    while b[i] != 0:
        name, i = _cstr(b, i)
        typ, i = _cstr(b, i)
        (size,) = struct.unpack_from("<i", b, i)
        i += 4
        attrs[name] = (typ, b[i:i + size])
        i += size
    return attrs, i + 1


def parse_channels(v):
    out, i = [], 0
    while v[i] != 0:
        name, i = _cstr(v, i)
        ptype, plin, xs, ys = struct.unpack_from("<iB3xii", v, i)
        i += 16
        out.append((name, ptype, plin, xs, ys))
    return out


def read_exr(path):
    """Returns (channels dict name -> H x W array, header attributes)."""
    b = open(path, "rb").read()
    attrs, pos = read_header(b)
    chans = parse_channels(attrs["channels"][1])
    comp = attrs["compression"][1][0]
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    W, H = x1 - x0 + 1, y1 - y0 + 1
    lines = LINES[comp]
    n = (H + lines - 1) // lines
    offs = struct.unpack_from(f"<{n}Q", b, pos)
    out = {c[0]: np.zeros((H, W), PIXEL[c[1]]) for c in chans}
    for k, off in enumerate(offs):
        y, size = struct.unpack_from("<ii", b, off)
        data = b[off + 8:off + 8 + size]
        nl = min(lines, y1 + 1 - y)
        raw_size = sum(np.dtype(PIXEL[c[1]]).itemsize for c in chans) * W * nl
        if comp in (2, 3) and size < raw_size:
            t = np.frombuffer(zlib.decompress(data), np.uint8).astype(np.int64)
            t = (np.cumsum(t - 128) + 128) % 256  # undo the delta coding (t[0] is stored as is)
            t = t.astype(np.uint8)
            half = (len(t) + 1) // 2
            raw = np.empty(len(t), np.uint8)
            raw[0::2], raw[1::2] = t[:half], t[half:]
            data = raw.tobytes()
        assert len(data) == raw_size, (len(data), raw_size)
        i = 0
        for r in range(nl):
            for name, ptype, *_ in chans:
                dt = np.dtype(PIXEL[ptype])
                out[name][y - y0 + r] = np.frombuffer(data, dt, W, i)
                i += dt.itemsize * W
    return out, attrs

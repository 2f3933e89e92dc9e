"""Minimal independent OpenEXR reader for the writer's output (uncompressed
scanline FLOAT/HALF), following the OpenEXR file-layout document: magic,
version word, NUL-terminated attributes (name, type, u32 size, payload), an
empty name ending the header, the u64 line-offset table, then per scanline
(i32 y, u32 size, channels in the order of the channel list)."""
import struct

import numpy as np


def parse(data: bytes):
    magic, version = struct.unpack_from("<II", data, 0)
    assert magic == 20000630, hex(magic)
    off = 8
    attrs = {}

    def cstr():
        nonlocal off
        end = data.index(b"\0", off)
        s = data[off:end].decode()
        off = end + 1
        return s

    while True:
        name = cstr()
        if not name:
            break
        typ = cstr()
        (size,) = struct.unpack_from("<I", data, off)
        off += 4
        attrs[name] = (typ, data[off:off + size])
        off += size
    chans = []
    payload = attrs["channels"][1]
    p = 0
    while payload[p] != 0:
        end = payload.index(b"\0", p)
        cname = payload[p:end].decode()
        ptype = struct.unpack_from("<I", payload, end + 1)[0]
        chans.append((cname, ptype))
        p = end + 1 + 16
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    assert attrs["compression"][1] == b"\0"
    offsets = struct.unpack_from(f"<{h}Q", data, off)
    planes = {c: np.zeros((h, w), np.float32) for c, _ in chans}
    for row, o in enumerate(offsets):
        y, size = struct.unpack_from("<iI", data, o)
        assert y == row + y0
        q = o + 8
        for cname, ptype in chans:
            if ptype == 1:
                vals = np.frombuffer(data, dtype="<f2", count=w, offset=q).astype(np.float32)
                q += 2 * w
            else:
                assert ptype == 2
                vals = np.frombuffer(data, dtype="<f4", count=w, offset=q)
                q += 4 * w
            planes[cname][row] = vals
        assert q == o + 8 + size
    strings = {k: v[1].decode() for k, v in attrs.items() if v[0] == "string"}
    return {"version": version, "width": w, "height": h, "channels": chans, "planes": planes,
            "strings": strings, "attrs": attrs}


def read_rgb(path):
    e = parse(open(path, "rb").read())
    return np.stack([e["planes"]["R"], e["planes"]["G"], e["planes"]["B"]], axis=-1), e

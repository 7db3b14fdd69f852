"""OpenEXR output of the film (Image::writeEXR, Utils/Image.cpp:49-106):
single-part scanline file, no compression, three HALF channels stored in
B, G, R order, increasing-Y line order.  `read_exr` reads such files back
(uncompressed scanline, HALF or FLOAT channels) for checks and tools.

The layout follows the OpenEXR file format specification: magic number and
version, a header of (name, type, size, value) attributes ending with an
empty name, a table of 64-bit scanline offsets, then one block per scanline
(y, byte count, each channel's samples in header channel order).
"""
from __future__ import annotations

import struct
from pathlib import Path

import numpy as np

MAGIC = 20000630
HALF, FLOAT = 1, 2


def _attr(name: str, typ: str, value: bytes) -> bytes:
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(value)) + value


def write_exr(path, rgb: np.ndarray) -> None:
    """rgb: (H, W, 3) linear floats; stored as HALF in B, G, R channel order."""
    path = Path(path)
    if path.suffix != ".exr":
        raise ValueError("Output file must be .exr")
    img = np.asarray(rgb, np.float32)
    if img.ndim != 3 or img.shape[2] < 3:
        raise ValueError("expected an (H, W, 3) image")
    h, w = img.shape[:2]
    chans = [("B", 2), ("G", 1), ("R", 0)]
    chlist = b"".join(c.encode() + b"\0" + struct.pack("<iB3xii", HALF, 0, 1, 1)
                      for c, _ in chans) + b"\0"
    box = struct.pack("<iiii", 0, 0, w - 1, h - 1)
    header = b"".join([
        _attr("channels", "chlist", chlist),
        _attr("compression", "compression", b"\0"),
        _attr("dataWindow", "box2i", box),
        _attr("displayWindow", "box2i", box),
        _attr("lineOrder", "lineOrder", b"\0"),
        _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0)),
        _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0)),
        _attr("screenWindowWidth", "float", struct.pack("<f", 1.0)),
    ]) + b"\0"
    head = struct.pack("<ii", MAGIC, 2) + header
    line_bytes = w * 2 * len(chans)
    table_at = len(head)
    first = table_at + 8 * h
    offsets = struct.pack(f"<{h}Q", *[first + y * (8 + line_bytes) for y in range(h)])
    with np.errstate(over="ignore"):   # beyond the half range -> inf, as the format stores it
        half = img[..., :3].astype(np.float16)
    blocks = []
    for y in range(h):
        data = b"".join(half[y, :, c].astype("<f2").tobytes() for _, c in chans)
        blocks.append(struct.pack("<ii", y, len(data)) + data)
    path.write_bytes(head + offsets + b"".join(blocks))


def read_exr(path) -> tuple[np.ndarray, dict]:
    """Returns ((H, W, C) float32 in the file's channel order, header dict
    with 'channels' = list of (name, pixel_type))."""
    buf = Path(path).read_bytes()
    magic, version = struct.unpack_from("<ii", buf, 0)
    if magic != MAGIC:
        raise ValueError("not an OpenEXR file")
    if version & 0xFFFFFE00:
        raise ValueError("only single-part scanline files are supported")
    pos, hdr = 8, {}
    while buf[pos] != 0:
        end = buf.index(b"\0", pos)
        name = buf[pos:end].decode()
        tend = buf.index(b"\0", end + 1)
        typ = buf[end + 1:tend].decode()
        (size,) = struct.unpack_from("<i", buf, tend + 1)
        val = buf[tend + 5:tend + 5 + size]
        pos = tend + 5 + size
        if typ == "chlist":
            chans, q = [], 0
            while val[q] != 0:
                e = val.index(b"\0", q)
                (pt,) = struct.unpack_from("<i", val, e + 1)
                chans.append((val[q:e].decode(), pt))
                q = e + 1 + 16
            hdr["channels"] = chans
        elif typ == "box2i":
            hdr[name] = struct.unpack("<iiii", val)
        elif typ in ("compression", "lineOrder"):
            hdr[name] = val[0]
        else:
            hdr[name] = val
    pos += 1
    if hdr.get("compression", 0) != 0:
        raise ValueError("only uncompressed files are supported")
    x0, y0, x1, y1 = hdr["dataWindow"]
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack_from(f"<{h}Q", buf, pos)
    chans = hdr["channels"]
    out = np.zeros((h, w, len(chans)), np.float32)
    for off in offsets:
        y, _ = struct.unpack_from("<ii", buf, off)
        q = off + 8
        for c, (_, pt) in enumerate(chans):
            dt, nb = ("<f2", 2) if pt == HALF else ("<f4", 4)
            out[y - y0, :, c] = np.frombuffer(buf, dt, w, q).astype(np.float32)
            q += w * nb
    return out, hdr

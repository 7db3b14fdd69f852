"""Image files of the reference's Image class (Utils/Image.cpp):
* write_exr (Image::writeEXR, :49-106): single-part scanline file, no
  compression, three HALF channels stored in B, G, R order, increasing-Y
  line order;
* read_exr: scanline files with NONE, ZIPS or ZIP compression and HALF,
  FLOAT or UINT channels; load_exr_image maps them as Image::readEXR does
  (:108-180): R, G, B (first letter of the name) to channels 0-2, the
  file's channel count kept, other channels zero;
* read_png / write_png (Image::readMisc / writePNG, :182-259): 8-bit grey,
  RGB or RGBA PNG, read bottom-up (stb's vertical flip), /255 and ^2.2 when
  converting to linear; written with a v^(1/2.2) gamma and clamp;
* load_image (Image::load): .exr via load_exr_image, else read_png.

The layout follows the OpenEXR file format specification: magic number and
version, a header of (name, type, size, value) attributes ending with an
empty name, a table of 64-bit scanline offsets, then one block per scanline
(y, byte count, each channel's samples in header channel order).
"""
from __future__ import annotations

import struct
import zlib
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
    comp = hdr.get("compression", 0)
    if comp not in (0, 2, 3):
        raise ValueError(f"EXR compression {comp} is not read (NONE, ZIPS, ZIP)")
    x0, y0, x1, y1 = hdr["dataWindow"]
    w, h = x1 - x0 + 1, y1 - y0 + 1
    lines = 16 if comp == 3 else 1
    nchunk = (h + lines - 1) // lines
    offsets = struct.unpack_from(f"<{nchunk}Q", buf, pos)
    chans = hdr["channels"]
    types = {0: ("<u4", 4), HALF: ("<f2", 2), FLOAT: ("<f4", 4)}
    line_bytes = w * sum(types[pt][1] for _, pt in chans)
    out = np.zeros((h, w, len(chans)), np.float32)
    for off in offsets:
        y, size = struct.unpack_from("<ii", buf, off)
        y -= y0
        n = min(lines, h - y)
        raw = line_bytes * n
        data = buf[off + 8:off + 8 + size]
        if comp != 0 and size != raw:       # zlib, then the predictor and byte split undone
            t = np.frombuffer(zlib.decompress(data), np.uint8).astype(np.int64)
            t = (np.cumsum(t - 128) + 128) & 0xFF   # t[i] += t[i - 1] - 128 (mod 256)
            half = (raw + 1) // 2
            d = np.empty(raw, np.uint8)
            d[0::2] = t[:half]
            d[1::2] = t[half:]
            data = d.tobytes()
        q = 0
        for ln in range(n):
            for c, (_, pt) in enumerate(chans):
                dt, nb = types[pt]
                out[y + ln, :, c] = np.frombuffer(data, dt, w, q).astype(np.float32)
                q += w * nb
    return out, hdr


def load_exr_image(path) -> np.ndarray:
    """(H, W, C) as Image::readEXR leaves it: R, G, B in channels 0-2, the
    file's channel count, any further channel zero."""
    data, hdr = read_exr(path)
    names = [n for n, _ in hdr["channels"]]
    out = np.zeros_like(data)
    for c, letter in enumerate("RGB"):
        idx = [j for j, n in enumerate(names) if n[:1] == letter]
        if not idx:
            raise ValueError("EXR file without R, G and B channels")
        out[..., c] = data[..., idx[0]]
    return out


def _png_chunks(buf: bytes):
    if buf[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("unsupported image format (PNG and EXR are read)")
    p = 8
    while p + 12 <= len(buf):
        (n,) = struct.unpack_from(">I", buf, p)
        yield buf[p + 4:p + 8], buf[p + 8:p + 8 + n]
        p += 12 + n


def read_png(path, to_linear: bool = True) -> np.ndarray:
    """(H, W, C) float32 of an 8-bit PNG, rows bottom-up, /255 (^2.2 if
    to_linear)."""
    buf = Path(path).read_bytes()
    idat, w = [], 0
    for typ, d in _png_chunks(buf):
        if typ == b"IHDR":
            w, h, depth, ctype, _, _, inter = struct.unpack(">IIBBBBB", d)
        elif typ == b"IDAT":
            idat.append(d)
        elif typ == b"IEND":
            break
    ch = {0: 1, 2: 3, 4: 2, 6: 4}.get(ctype, 0)
    if depth != 8 or ch == 0 or inter != 0:
        raise ValueError("PNG must be 8-bit grey, grey + alpha, RGB or RGBA, not interlaced")
    if ch == 2:
        raise ValueError("Unsupported color type: 2")
    raw = np.frombuffer(zlib.decompress(b"".join(idat)), np.uint8).reshape(h, w * ch + 1)
    stride = w * ch
    pix = np.zeros((h, stride), np.uint8)
    prev = np.zeros(stride, np.int64)
    for y in range(h):
        f, row = raw[y, 0], raw[y, 1:].astype(np.int64)
        cur = np.zeros(stride, np.int64)
        if f in (0, 2):
            cur = (row + (prev if f == 2 else 0)) & 0xFF
        else:
            for i in range(stride):
                a = cur[i - ch] if i >= ch else 0
                b = prev[i]
                c = prev[i - ch] if i >= ch else 0
                if f == 1:
                    pr = a
                elif f == 3:
                    pr = (a + b) // 2
                elif f == 4:
                    pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                    pr = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
                else:
                    raise ValueError("corrupt PNG filter")
                cur[i] = (row[i] + pr) & 0xFF
        pix[h - 1 - y] = cur
        prev = cur
    img = pix.reshape(h, w, ch).astype(np.float32) / np.float32(255.0)
    if to_linear:
        img = np.power(img.astype(np.float64), 2.2).astype(np.float32)
    return img


def write_png(path, img: np.ndarray, gamma_correct: bool = True) -> None:
    """8-bit PNG of an (H, W, C) linear image (row 0 first in the file)."""
    path = Path(path)
    if path.suffix != ".png":
        raise ValueError("Output file must be .png")
    a = np.asarray(img, np.float32)
    if a.ndim == 2:
        a = a[..., None]
    h, w, ch = a.shape
    if gamma_correct:
        a = np.clip(np.power(a, np.float32(1 / 2.2)), 0.0, 1.0)
    b = (a * np.float32(255.0)).astype(np.uint8)
    raw = b"".join(b"\0" + b[y].tobytes() for y in range(h))

    def chunk(typ, data):
        return (struct.pack(">I", len(data)) + typ + data +
                struct.pack(">I", zlib.crc32(typ + data) & 0xFFFFFFFF))
    ihdr = struct.pack(">IIBBBBB", w, h, 8, {1: 0, 2: 4, 3: 2, 4: 6}[ch], 0, 0, 0)
    path.write_bytes(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr) +
                     chunk(b"IDAT", zlib.compress(raw, 6)) + chunk(b"IEND", b""))


def load_image(path, to_linear: bool = True) -> np.ndarray:
    """Image::load: .exr through load_exr_image, anything else as PNG."""
    path = Path(path)
    return load_exr_image(path) if path.suffix == ".exr" else read_png(path, to_linear)

"""Writers for the asset-ingest fixtures: PNG (any filter per row, gAMA, sBIT,
grey / RGB / RGBA, 8 and 16 bits), scanline OpenEXR (NONE, RLE, ZIPS, ZIP;
HALF and FLOAT channels; any data window) and OBJ / MTL text. They follow the
published file formats (PNG specification, OpenEXR file layout), not any
decoder, so a decoder that reads them back to the values written is right on
its own terms."""
from __future__ import annotations

import struct
import zlib

import numpy as np


# ---- PNG ----------------------------------------------------------------------
def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def _chunk(typ: bytes, body: bytes) -> bytes:
    return struct.pack(">I", len(body)) + typ + body + struct.pack(">I", zlib.crc32(typ + body) & 0xFFFFFFFF)


def write_png(path, samples: np.ndarray, filters=None, gamma: int | None = None, sbit=None,
              split_idat: int = 0) -> None:
    """samples: uint8 or uint16, [h, w] (grey), [h, w, 3] (RGB) or [h, w, 4]
    (RGBA). filters: one filter type (0..4) per row, cycled. gamma: the gAMA
    value in 1e-5 units. split_idat: bytes per IDAT chunk (0: one chunk)."""
    s = np.asarray(samples)
    depth = 16 if s.dtype == np.uint16 else 8
    if s.ndim == 2:
        s = s[..., None]
    h, w, ch = s.shape
    ctype = {1: 0, 3: 2, 4: 6}[ch]
    raw = s.astype(">u2").tobytes() if depth == 16 else s.astype(np.uint8).tobytes()
    bpp = ch * depth // 8
    row_bytes = w * bpp
    filters = filters or [0]
    out = bytearray()
    prev = bytes(row_bytes)
    for y in range(h):
        cur = raw[y * row_bytes:(y + 1) * row_bytes]
        ft = filters[y % len(filters)]
        out.append(ft)
        for i in range(row_bytes):
            a = cur[i - bpp] if i >= bpp else 0
            b = prev[i] if y else 0
            c = prev[i - bpp] if (y and i >= bpp) else 0
            pred = [0, a, b, (a + b) >> 1, _paeth(a, b, c)][ft]
            out.append((cur[i] - pred) & 0xFF)
        prev = cur
    z = zlib.compress(bytes(out), 9)
    body = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 0))
    if gamma is not None:
        body += _chunk(b"gAMA", struct.pack(">I", gamma))
    if sbit is not None:
        body += _chunk(b"sBIT", bytes(sbit))
    body += _chunk(b"tEXt", b"Comment\0ancillary chunks are skipped")
    step = split_idat or len(z)
    for k in range(0, len(z), step):
        body += _chunk(b"IDAT", z[k:k + step])
    body += _chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(body)


# ---- OpenEXR --------------------------------------------------------------------
NONE, RLE, ZIPS, ZIP = 0, 1, 2, 3


def _attr(name: str, typ: str, body: bytes) -> bytes:
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<I", len(body)) + body


def _predict_interleave(raw: bytes) -> bytes:
    t = raw[0::2] + raw[1::2]
    out = bytearray(t)
    for i in range(len(t) - 1, 0, -1):
        out[i] = (t[i] - t[i - 1] + 128) & 0xFF
    return bytes(out)


def _rle(data: bytes) -> bytes:
    out, i, n = bytearray(), 0, len(data)
    while i < n:
        run = 1
        while i + run < n and run < 128 and data[i + run] == data[i]:
            run += 1
        if run >= 3:
            out += struct.pack("b", run - 1) + data[i:i + 1]
            i += run
            continue
        j = i
        while j < n and j - i < 127:
            if j + 2 < n and data[j] == data[j + 1] == data[j + 2]:
                break
            j += 1
        out += struct.pack("b", -(j - i)) + data[i:j]
        i = j
    return bytes(out)


def write_exr(path, channels: dict, compression: int = NONE, origin=(0, 0)) -> list[bool]:
    """channels: name -> [h, w] float16 (HALF) or float32 (FLOAT) array.
    Returns, per chunk, whether it was stored compressed (a chunk that would
    not shrink is stored as is, as OpenEXR does)."""
    names = sorted(channels)
    h, w = next(iter(channels.values())).shape
    x0, y0 = origin
    chl = b""
    for nm in names:
        t = 1 if channels[nm].dtype == np.float16 else 2
        chl += nm.encode() + b"\0" + struct.pack("<iB3xii", t, 0, 1, 1)
    chl += b"\0"
    hdr = (_attr("channels", "chlist", chl) + _attr("compression", "compression", bytes([compression]))
           + _attr("dataWindow", "box2i", struct.pack("<iiii", x0, y0, x0 + w - 1, y0 + h - 1))
           + _attr("displayWindow", "box2i", struct.pack("<iiii", x0, y0, x0 + w - 1, y0 + h - 1))
           + _attr("lineOrder", "lineOrder", b"\0") + _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
           + _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0, 0))
           + _attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\0")
    lpc = 16 if compression == ZIP else 1
    chunks, packed = [], []
    for cy in range(0, h, lpc):
        raw = b""
        for y in range(cy, min(h, cy + lpc)):
            for nm in names:
                a = channels[nm][y]
                raw += a.astype("<f2" if a.dtype == np.float16 else "<f4").tobytes()
        data = raw
        if compression == RLE:
            data = _rle(_predict_interleave(raw))
        elif compression in (ZIPS, ZIP):
            data = zlib.compress(_predict_interleave(raw), 9)
        if compression == NONE or len(data) >= len(raw):
            data = raw
        packed.append(data is not raw)
        chunks.append(struct.pack("<iI", y0 + cy, len(data)) + data)
    head = struct.pack("<II", 20000630, 2) + hdr
    table, off = b"", len(head) + 8 * len(chunks)
    for c in chunks:
        table += struct.pack("<Q", off)
        off += len(c)
    with open(path, "wb") as f:
        f.write(head + table + b"".join(chunks))
    return packed


# ---- OBJ / MTL ------------------------------------------------------------------
OBJ_A = """# two objects, five groups
mtllib a.mtl
o first
v 0 0 0
v 1 0 0
v 0 1 0
v 1 1 0
vt 0 0
vt 1 0
vt 0 0.25
vn 0 0 1
usemtl red
f 1/1/1 2/2/1 3/3/1
f 2/2/1 4/2/1 3/3/1 1/1/1
f 1//1 2//1 4//1

usemtl glass
f 1//1 2//1 4//1
usemtl missing
f 1//1 3//1 4//1
　 usemtl tex
f 1/1/1 2/2/1 4/3/1
o second
v 0 0 1
v 1 0 1
v 0 1 1
vn 0 1 0
g grp
usemtl red
f 5//2 6//2 7//2
"""
MTL_A = """newmtl red
Kd 0.8 0.1 0.1
Ns 10
newmtl glass
Kd 0.2 0.2 0.2
Ni 1.5
  newmtl tex
map_Kd tex.png
"""


def write_obj_fixture(directory) -> None:
    """a.obj (two objects, five face groups in both face forms, a quad, a
    face in the wrong form, an unknown material, a line with leading
    ideographic and plain spaces), a.mtl (matte, glass with Ni 1.5, and an
    image-textured material) and tex.png (6x4 RGBA with one transparent
    texel) into `directory`."""
    from pathlib import Path

    d = Path(directory)
    (d / "a.obj").write_text(OBJ_A, encoding="utf-8")
    (d / "a.mtl").write_text(MTL_A, encoding="utf-8")
    tex = np.random.default_rng(7).integers(0, 256, (4, 6, 4), dtype=np.uint8)
    tex[0, 0, 3] = 0
    write_png(d / "tex.png", tex)

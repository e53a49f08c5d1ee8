"""Test infrastructure (the checker, never the product): a pure-Python
restatement of the reference's asset ingest, used to check the library's
loaders (include/clear_amd_loaders.hpp) on the same files.

- decode_png: the image the reference gets from libpng with the settings of
  loadPNG (OpenCL_PathTracer/ImageLoader.cpp:72-201): RGB padded to 4 bytes
  with 0xFF (png_set_filler), 16-bit samples stripped to the high byte
  (png_set_strip_16), sBIT shifts (png_set_shift), and png_set_gamma(screen,
  file) with the file gamma from gAMA or 0.45455, the screen gamma 1.0, or 2.2
  (SCREEN_GAMMA) with gamma correction. libpng's gamma table arithmetic is
  restated from its published source (png.c: png_product2, png_reciprocal2,
  png_gamma_significant, png_gamma_8bit_correct / _16bit_correct).
- decode_exr: Imf::RgbaInputFile's view of a scanline OpenEXR file
  (ImageLoader.cpp:27-56): R, G, B, A as half floats (missing colour 0,
  missing alpha 1), NONE / RLE / ZIPS / ZIP chunks (OpenEXR's ImfRle / ImfZip
  predictor and byte reordering).
- parse_obj: the objects, groups and materials OBJ::load and loadMtl build
  (ModelLoader.cpp:36-361) before they are handed to the Scene.
"""
from __future__ import annotations

import math
import os
import re
import struct
import zlib

import numpy as np


# ---- PNG --------------------------------------------------------------------
def _png_fixed(v: float) -> int:
    return int(math.floor(v * 100000.0 + 0.5))


def _gamma_significant(g: int) -> bool:
    return g < 95000 or g > 105000


def _product2(a: int, b: int) -> int:
    return int(math.floor(a * 1e-5 * b + 0.5))


def _reciprocal2(a: int, b: int) -> int:
    return int(math.floor(1e15 / a / b + 0.5))


def gamma8(v: int, g: int) -> int:
    """png_gamma_8bit_correct: floor(255 (v/255)^(g 1e-5) + 0.5), 0 and 255 fixed."""
    if 0 < v < 255:
        return int(math.floor(255.0 * math.pow(v / 255.0, g * 0.00001) + 0.5))
    return v


def gamma16(v: int, g: int) -> int:
    if 0 < v < 65535:
        return int(math.floor(65535.0 * math.pow(v / 65535.0, g * 0.00001) + 0.5))
    return v


def gamma_setting(file_gamma: int | None, gamma_correction: bool) -> int | None:
    """The table exponent (1e-5 units) libpng uses, or None when it leaves samples alone."""
    screen = _png_fixed(float(os.environ.get("SCREEN_GAMMA", "2.2")) if gamma_correction else 1.0)
    fg = file_gamma if file_gamma is not None else _png_fixed(0.45455)
    if fg <= 0 or not _gamma_significant(_product2(fg, screen)):
        return None
    g = _reciprocal2(fg, screen)
    return g if _gamma_significant(g) else None


def _paeth(a: int, b: int, c: int) -> int:
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def decode_png(path, gamma_correction: bool = False):
    """(pixels uint8 [h, w, 4] or [h, w], color type 0 RGB8 / 1 RGBA8 / 3 gray8)."""
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, gam, sbit = 8, b"", None, None
    while True:
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        typ, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        if typ == b"IHDR":
            w, h, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", body)
        elif typ == b"gAMA":
            gam = struct.unpack(">I", body)[0]
        elif typ == b"sBIT":
            sbit = list(body)
        elif typ == b"IDAT":
            idat += body
        elif typ == b"IEND":
            break
        pos += 12 + n
    assert interlace == 0 and depth in (8, 16) and ctype in (0, 2, 6)
    ch = {0: 1, 2: 3, 6: 4}[ctype]
    bps = depth // 8
    bpp, row_bytes = ch * bps, w * ch * bps
    raw = zlib.decompress(idat)
    rows, prev = [], bytearray(row_bytes)
    for y in range(h):
        ft = raw[y * (row_bytes + 1)]
        src = raw[y * (row_bytes + 1) + 1:(y + 1) * (row_bytes + 1)]
        cur = bytearray(row_bytes)
        for i in range(row_bytes):
            a = cur[i - bpp] if i >= bpp else 0
            b = prev[i] if y else 0
            c = prev[i - bpp] if (y and i >= bpp) else 0
            pred = [0, a, b, (a + b) >> 1, _paeth(a, b, c)][ft]
            cur[i] = (src[i] + pred) & 0xFF
        rows.append(cur)
        prev = cur
    samples = np.frombuffer(b"".join(rows), dtype=">u2" if depth == 16 else np.uint8).reshape(h, w, ch).astype(np.int64)
    g = gamma_setting(gam, gamma_correction)
    out = np.zeros((h, w, ch), np.int64)
    for c in range(ch):
        v = samples[..., c]
        alpha = ch == 4 and c == 3
        if g is not None and not alpha:
            f = gamma8 if depth == 8 else gamma16
            lut = {int(x): f(int(x), g) for x in np.unique(v)}
            v = np.vectorize(lut.get)(v)
        if depth == 16:
            v = v >> 8
        shift = 8 - sbit[c] if (sbit is not None and 0 < sbit[c] < 8) else 0
        out[..., c] = v >> shift
    if ch == 1:
        return out[..., 0].astype(np.uint8), 3
    if ch == 3:
        out = np.concatenate([out, np.full((h, w, 1), 255, np.int64)], axis=2)
        return out.astype(np.uint8), 0
    return out.astype(np.uint8), 1


# ---- OpenEXR ------------------------------------------------------------------
def _unpredict_interleave(t: bytes) -> bytes:
    t = bytearray(t)
    for i in range(1, len(t)):
        t[i] = (t[i - 1] + t[i] - 128) & 0xFF
    n = len(t)
    half = (n + 1) // 2
    out = bytearray(n)
    out[0::2] = t[:half]
    out[1::2] = t[half:]
    return bytes(out)


def _rle_decode(d: bytes) -> bytes:
    out, i = bytearray(), 0
    while i < len(d):
        n = struct.unpack("b", d[i:i + 1])[0]
        i += 1
        if n < 0:
            out += d[i:i - n]
            i -= n
        else:
            out += d[i:i + 1] * (n + 1)
            i += 1
    return bytes(out)


def decode_exr(path):
    """float16 [h, w, 4] (R, G, B, A)."""
    data = open(path, "rb").read()
    assert struct.unpack("<I", data[:4])[0] == 20000630
    pos, attrs = 8, {}
    while data[pos] != 0:
        e = data.index(b"\0", pos)
        name = data[pos:e].decode()
        e2 = data.index(b"\0", e + 1)
        size = struct.unpack("<I", data[e2 + 1:e2 + 5])[0]
        attrs[name] = data[e2 + 5:e2 + 5 + size]
        pos = e2 + 5 + size
    pos += 1
    chans, q, cl = [], 0, attrs["channels"]
    while cl[q] != 0:
        e = cl.index(b"\0", q)
        chans.append((cl[q:e].decode(), struct.unpack("<i", cl[e + 1:e + 5])[0]))
        q = e + 17
    comp = attrs["compression"][0]
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    lpc = 16 if comp == 3 else 1
    nch = (h + lpc - 1) // lpc
    offsets = struct.unpack(f"<{nch}Q", data[pos:pos + 8 * nch])
    out = np.zeros((h, w, 4), np.float16)
    out[..., 3] = 1.0
    line_bytes = sum(w * (2 if t == 1 else 4) for _, t in chans)
    for off in offsets:
        y, size = struct.unpack("<iI", data[off:off + 8])
        lines = min(lpc, y1 - y + 1)
        body = data[off + 8:off + 8 + size]
        if size != line_bytes * lines:
            raw = _rle_decode(body) if comp == 1 else zlib.decompress(body)
            body = _unpredict_interleave(raw)
        q = 0
        for li in range(lines):
            for name, t in chans:
                nb = w * (2 if t == 1 else 4)
                vals = np.frombuffer(body[q:q + nb], "<f2" if t == 1 else "<f4")
                q += nb
                slot = "RGBA".find(name) if name in ("R", "G", "B", "A") else -1
                if slot >= 0:
                    out[y - y0 + li, :, slot] = vals.astype(np.float16)
    return out


# ---- OBJ / MTL -------------------------------------------------------------------
_LEAD = "　 ¥t"  # the reference's "　 ¥t" (its bytes, decoded as UTF-8)


def _strip(line: str) -> str:
    i = 0
    while i < len(line) and line[i] in _LEAD:
        i += 1
    return line[i:]


_F = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"


def _floats(line: str, key: str, n: int):
    m = re.match(r"^" + re.escape(key) + r"\s+" + r"\s+".join([f"({_F})"] * n), line)
    return [float(np.float32(float(x))) for x in m.groups()] if m else None


def _word(line: str, key: str):
    m = re.match(r"^" + re.escape(key) + r"\s+(\S+)", line)
    return m.group(1) if m else None


def parse_mtl(path):
    mats, cur = [], None
    directory = os.path.dirname(str(path))
    directory = directory + "/" if directory else ""
    for line in open(path, encoding="utf-8", errors="surrogateescape"):
        line = _strip(line.rstrip("\r\n"))
        if not line:
            continue
        c = line[0]
        if c == "K" and len(line) >= 2 and cur is not None:
            v = _floats(line, "K" + line[1], 3) or [0.0, 0.0, 0.0]
            if line[1] in "dsa":
                cur["K" + line[1]] = v
        elif c == "N" and cur is not None:
            v = _floats(line, "Ni", 1)
            if v:
                cur["Ni"] = v[0]
        elif c == "m" and cur is not None:
            for key, dst in (("map_Kd", "mapKd"), ("map_Ks", "mapKs"), ("map_Ka", "mapKa"), ("map_bump", "mapKd")):
                t = _word(line, key)
                if t:
                    cur[dst] = directory + t
                    break
        elif c == "n":
            nm = _word(line, "newmtl")
            if nm:
                cur = {"name": nm, "Kd": [0.0] * 3, "Ks": [0.0] * 3, "Ka": [0.0] * 3, "Ni": 1.0, "mapKd": ""}
                mats.append(cur)
    return mats


def parse_obj(path):
    """{'objects': [{'positions', 'uvs', 'normals', 'groups': [{'material', 'format',
    'pos', 'uv', 'nrm'}]}], 'materials': [...]} as ModelLoader builds them."""
    directory = os.path.dirname(str(path))
    directory = directory + "/" if directory else ""
    objects, materials, mat_names = [], [], {}
    cur_group, group_name, mat_name, phase_f = None, "", "", False

    def obj():
        if not objects:
            objects.append({"positions": [], "uvs": [], "normals": [], "groups": []})
        return objects[-1]

    vtn = re.compile(r"^f\s+(\d+)/(\d+)/(\d+)\s+(\d+)/(\d+)/(\d+)\s+(\d+)/(\d+)/(\d+)")
    vn = re.compile(r"^f\s+(\d+)//(\d+)\s+(\d+)//(\d+)\s+(\d+)//(\d+)")
    for line in open(path, encoding="utf-8", errors="surrogateescape"):
        line = _strip(line.rstrip("\r\n"))
        if not line:
            phase_f = False
            continue
        if phase_f and line[0] != "f":
            phase_f = False
        c = line[0]
        if c == "o":
            nm = _word(line, "o")
            if nm:
                objects.append({"positions": [], "uvs": [], "normals": [], "groups": []})
        elif c == "g":
            nm = _word(line, "g")
            if nm:
                group_name = nm
        elif c == "v" and len(line) > 1:
            if line[1] == " ":
                obj()["positions"] += _floats(line, "v", 3) or [0.0, 0.0, 0.0]
            elif line[1] == "t":
                u, v = _floats(line, "vt", 2) or [0.0, 0.0]
                obj()["uvs"] += [u, float(np.float32(1.0) - np.float32(v))]
            elif line[1] == "n":
                obj()["normals"] += _floats(line, "vn", 3) or [0.0, 0.0, 0.0]
        elif c == "f":
            if not phase_f:
                phase_f = True
                cur_group = {"name": group_name, "material": mat_name, "format": None, "pos": [], "uv": [], "nrm": []}
                obj()["groups"].append(cur_group)
                if vtn.match(line):
                    cur_group["format"] = "VTN"
                elif vn.match(line):
                    cur_group["format"] = "VN"
            m = vtn.match(line) if cur_group["format"] == "VTN" else vn.match(line) if cur_group["format"] == "VN" else None
            if m:
                g = [int(x) - 1 for x in m.groups()]
                if cur_group["format"] == "VTN":
                    for k in range(3):
                        cur_group["pos"].append(g[3 * k])
                        cur_group["uv"].append(g[3 * k + 1])
                        cur_group["nrm"].append(g[3 * k + 2])
                else:
                    for k in range(3):
                        cur_group["pos"].append(g[2 * k])
                        cur_group["nrm"].append(g[2 * k + 1])
        elif c == "m" and line.startswith("mtllib"):
            nm = _word(line, "mtllib")
            if nm:
                for m in parse_mtl(directory + nm):
                    mat_names.setdefault(m["name"], len(materials))
                    materials.append(m)
        elif c == "u" and line.startswith("usemtl"):
            nm = _word(line, "usemtl")
            if nm:
                mat_name = nm
    return {"objects": objects, "materials": materials, "material_index": mat_names}

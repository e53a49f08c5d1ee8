"""Asset ingest (clr_image_load, clr_scene_build_files): PNG and OpenEXR
decoding and OBJ / MTL loading, on fixtures written by tests/asset_files.py.

Each decoder is checked two ways: against values the file format fixes
(samples written = samples read wherever the reference's libpng setup leaves
them alone, libpng's gamma law where it does not, IEEE half conversion for
FLOAT channels, hand-worked face lists for the OBJ fixture), and against the
oracle's independent restatement (oracle/assets_oracle.py) on the same files.
"""
import math
import struct
import os

import numpy as np
import pytest

import clear_amd as ca
from oracle import assets_oracle as ao
from tests import asset_files as af


def _rng(seed):
    return np.random.default_rng(seed)


# ---- PNG ----------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(13, 7), (9, 11, 3), (6, 5, 4)])
@pytest.mark.parametrize("filters", [[0], [1], [2], [3], [4], [0, 1, 2, 3, 4]])
def test_png_samples_as_stored(tmp_path, shape, filters):
    """gamma_correction on, no gAMA: file gamma 0.45455 times screen gamma 2.2
    is 1.00001, which libpng treats as no gamma, so samples come back as
    written (RGB padded with a 0xFF byte)."""
    s = _rng(len(shape) * 10 + filters[0]).integers(0, 256, shape, dtype=np.uint8)
    p = tmp_path / "a.png"
    af.write_png(p, s, filters=filters, split_idat=37)
    px, w, h, ct = ca.load_image(p, gamma_correction=True)
    assert (w, h) == (shape[1], shape[0])
    if len(shape) == 2:
        assert ct == ca.COLOR_GRAY8
        np.testing.assert_array_equal(px, s)
    elif shape[2] == 3:
        assert ct == ca.COLOR_RGB8
        np.testing.assert_array_equal(px[..., :3], s)
        assert (px[..., 3] == 255).all()
    else:
        assert ct == ca.COLOR_RGBA8
        np.testing.assert_array_equal(px, s)
    o, oct_ = ao.decode_png(p, True)
    np.testing.assert_array_equal(px, o)
    assert oct_ == {ca.COLOR_GRAY8: 3, ca.COLOR_RGB8: 0, ca.COLOR_RGBA8: 1}[ct]


def test_png_default_gamma_law(tmp_path):
    """gamma_correction off, no gAMA: libpng maps v to
    floor(255 (v/255)^g + 0.5) with g = 1 / (0.45455 * 1.0) rounded to five
    decimals (2.19998); 0 and 255 are fixed; alpha is never gamma-corrected."""
    v = np.arange(256, dtype=np.uint8).reshape(16, 16)
    s = np.stack([v, v[::-1], v.T, v], axis=2)
    p = tmp_path / "g.png"
    af.write_png(p, s, filters=[4, 1])
    px, *_ = ca.load_image(p, gamma_correction=False)
    g = round(1.0 / 0.45455, 5)
    assert g == 2.19998
    law = np.array([math.floor(255.0 * (x / 255.0) ** g + 0.5) for x in range(256)], np.uint8)
    assert law[0] == 0 and law[255] == 255 and law[128] == 56
    np.testing.assert_array_equal(px[..., :3], law[s[..., :3]])
    np.testing.assert_array_equal(px[..., 3], s[..., 3])
    np.testing.assert_array_equal(px, ao.decode_png(p, False)[0])


@pytest.mark.parametrize("gama, correction, exponent", [
    (100000, False, None),      # file 1.0 x screen 1.0: no gamma
    (100000, True, 0.45455),    # file 1.0, screen 2.2: v^(1/2.2)
    (45455, True, None),        # the usual sRGB-ish file shown on a 2.2 screen: no gamma
    (50000, False, 2.0),        # file 0.5, screen 1.0: v^2
])
def test_png_gama_chunk(tmp_path, gama, correction, exponent):
    v = np.arange(256, dtype=np.uint8).reshape(16, 16)
    p = tmp_path / "c.png"
    af.write_png(p, v, gamma=gama)
    px, *_ = ca.load_image(p, gamma_correction=correction)
    if exponent is None:
        np.testing.assert_array_equal(px, v)
    else:
        law = np.array([math.floor(255.0 * (x / 255.0) ** exponent + 0.5) for x in range(256)], np.uint8)
        np.testing.assert_array_equal(px, law[v])
    np.testing.assert_array_equal(px, ao.decode_png(p, correction)[0])


def test_png_screen_gamma_env(tmp_path, monkeypatch):
    """SCREEN_GAMMA replaces 2.2 when gamma correction is on: a file gamma of
    1.0 on a screen of 2.0 gives v^0.5."""
    v = np.arange(256, dtype=np.uint8).reshape(16, 16)
    p = tmp_path / "e.png"
    af.write_png(p, v, gamma=100000)
    monkeypatch.setenv("SCREEN_GAMMA", "2.0")
    px, *_ = ca.load_image(p, gamma_correction=True)
    law = np.array([math.floor(255.0 * (x / 255.0) ** 0.5 + 0.5) for x in range(256)], np.uint8)
    np.testing.assert_array_equal(px, law[v])
    np.testing.assert_array_equal(px, ao.decode_png(p, True)[0])


def test_png_16bit_high_byte(tmp_path):
    s = _rng(3).integers(0, 65536, (7, 9, 4), dtype=np.uint16)
    p = tmp_path / "d.png"
    af.write_png(p, s, filters=[0, 1, 2, 3, 4])
    px, w, h, ct = ca.load_image(p, gamma_correction=True)
    assert ct == ca.COLOR_RGBA8 and (w, h) == (9, 7)
    np.testing.assert_array_equal(px, (s >> 8).astype(np.uint8))
    np.testing.assert_array_equal(px, ao.decode_png(p, True)[0])


def test_png_16bit_gamma_against_oracle(tmp_path):
    """16-bit samples with a gamma table: the library and the oracle share the
    restatement (16-bit gamma, then the high byte); libpng builds a combined
    16-to-8 table instead, which can differ by one in the last bit, so this
    case is compared with the oracle only."""
    s = _rng(4).integers(0, 65536, (5, 6, 3), dtype=np.uint16)
    p = tmp_path / "d16.png"
    af.write_png(p, s)
    px, *_ = ca.load_image(p, gamma_correction=False)
    np.testing.assert_array_equal(px, ao.decode_png(p, False)[0])
    law = np.floor(255.0 * (s / 65535.0) ** 2.19998 + 0.5)
    assert np.abs(px[..., :3].astype(int) - law).max() <= 1


def test_png_sbit_shift(tmp_path):
    """sBIT: samples are shifted down to their significant bits (png_set_shift)."""
    s = _rng(5).integers(0, 256, (8, 8, 3), dtype=np.uint8)
    p = tmp_path / "s.png"
    af.write_png(p, s, sbit=[5, 6, 8])
    px, *_ = ca.load_image(p, gamma_correction=True)
    np.testing.assert_array_equal(px[..., 0], s[..., 0] >> 3)
    np.testing.assert_array_equal(px[..., 1], s[..., 1] >> 2)
    np.testing.assert_array_equal(px[..., 2], s[..., 2])
    np.testing.assert_array_equal(px, ao.decode_png(p, True)[0])


def test_png_rejects(tmp_path):
    p = tmp_path / "bad.png"
    af.write_png(p, np.zeros((4, 4, 3), np.uint8))
    b = bytearray(p.read_bytes())
    b[29] ^= 1  # inside IHDR's CRC
    p.write_bytes(bytes(b))
    with pytest.raises(ca.ClrError, match="CRC"):
        ca.load_image(p)
    q = tmp_path / "x.jpg"
    q.write_bytes(b"\xff\xd8")
    with pytest.raises(ca.ClrError, match="unsupported image format"):
        ca.load_image(q)
    with pytest.raises(ca.ClrError, match="cannot read"):
        ca.load_image(tmp_path / "missing.png")
    r = tmp_path / "big.png"
    af.write_png(r, np.zeros((1, 5121), np.uint8))
    with pytest.raises(ca.ClrError, match="5120"):
        ca.load_image(r)


# ---- OpenEXR ----------------------------------------------------------------------
def _smooth(h, w, k):
    y, x = np.mgrid[0:h, 0:w]
    v = 0.25 * k + np.sin(0.3 * x + 0.1 * k) * np.cos(0.2 * y) * 4.0
    return (np.round(v * 8.0) / 8.0).astype(np.float32)  # short mantissas: RLE finds runs


@pytest.mark.parametrize("comp", [af.NONE, af.RLE, af.ZIPS, af.ZIP])
@pytest.mark.parametrize("dtype", [np.float16, np.float32])
def test_exr_channels(tmp_path, comp, dtype):
    h, w = 37, 23  # ZIP: two full 16-line chunks and a ragged one
    ch = {c: _smooth(h, w, k).astype(dtype) for k, c in enumerate("RGBA")}
    ch["Z"] = _smooth(h, w, 9).astype(dtype)  # not a colour channel: skipped
    p = tmp_path / "a.exr"
    packed = af.write_exr(p, ch, comp, origin=(-3, 5))
    if comp != af.NONE:
        assert any(packed)
    px, W, H, ct = ca.load_image(p)
    assert ct == ca.COLOR_RGBA16F and (W, H) == (w, h)
    for k, c in enumerate("RGBA"):
        np.testing.assert_array_equal(px[..., k].view(np.uint16), ch[c].astype(np.float16).view(np.uint16))
    np.testing.assert_array_equal(px.view(np.uint16), ao.decode_exr(p).view(np.uint16))


def test_exr_float_to_half_rounding(tmp_path):
    """FLOAT channels become halves by round-to-nearest-even (numpy's float16
    cast is IEEE): ties, overflow to infinity, subnormals, signed zeros, NaN."""
    vals = np.array([0.0, -0.0, 1.0, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, 65504.0, 65519.99, 65520.0, 1e9,
                     -1e9, 2.0 ** -24, 2.0 ** -25, 2.0 ** -25 * 1.0001, 3 * 2.0 ** -26, 6.1e-5, 1e-30,
                     np.inf, -np.inf, 0.1, -0.3], np.float32)
    vals = np.concatenate([vals, _rng(6).standard_normal(44).astype(np.float32) * 100])
    r = vals.reshape(8, 8)
    p = tmp_path / "r.exr"
    af.write_exr(p, {"R": r, "G": -r, "B": r * 1e-4}, af.NONE)
    px, *_ = ca.load_image(p)
    for k, a in enumerate([r, -r, r * np.float32(1e-4)]):
        with np.errstate(over="ignore"):
            want = a.astype(np.float16)
        np.testing.assert_array_equal(px[..., k].view(np.uint16), want.view(np.uint16))
    assert (px[..., 3] == 1.0).all()  # no alpha channel: 1
    nan = np.full((2, 2), np.nan, np.float32)
    q = tmp_path / "n.exr"
    af.write_exr(q, {"R": nan}, af.NONE)
    px, *_ = ca.load_image(q)
    assert np.isnan(px[..., 0].astype(np.float32)).all()
    assert (px[..., 1:3] == 0).all()  # missing colour channels: 0


def test_exr_rejects(tmp_path):
    p = tmp_path / "piz.exr"
    af.write_exr(p, {"R": np.zeros((2, 2), np.float16)}, af.NONE)
    b = bytearray(p.read_bytes())
    at = b.index(b"compression\0compression\0") + len(b"compression\0compression\0") + 4
    b[at] = 4  # PIZ
    p.write_bytes(bytes(b))
    with pytest.raises(ca.ClrError, match="compression"):
        ca.load_image(p)


def test_exr_rejects_malformed_offsets_and_windows(tmp_path):
    """A chunk offset near 2^64 (where off + 8 wraps) and a data window whose
    int32 corners are 2^32 - 1 apart are rejected, not read."""
    p = tmp_path / "ok.exr"
    af.write_exr(p, {"R": np.zeros((2, 2), np.float16)}, af.NONE)
    good = p.read_bytes()
    table = good.index(b"\0", good.rindex(b"screenWindowWidth")) + 1 + len("float") + 1 + 4 + 4 + 1
    q = tmp_path / "offset.exr"
    b = bytearray(good)
    b[table:table + 8] = struct.pack("<Q", 2 ** 64 - 4)
    q.write_bytes(bytes(b))
    with pytest.raises(ca.ClrError, match="chunk offset"):
        ca.load_image(q)
    q = tmp_path / "window.exr"
    b = bytearray(good)
    at = b.index(b"dataWindow\0box2i\0") + len(b"dataWindow\0box2i\0") + 4
    b[at:at + 16] = struct.pack("<iiii", -(2 ** 31), 0, 2 ** 31 - 1, 1)
    q.write_bytes(bytes(b))
    with pytest.raises(ca.ClrError, match="image size"):
        ca.load_image(q)


# ---- OBJ / MTL ------------------------------------------------------------------------
# (p, n, uv or None, material name) per face, in file order; "missing" falls
# back to the first material; the quad keeps its first three vertices and the
# p//n face inside the p/t/n group is skipped.
FACES_A = [((0, 1, 2), (0, 0, 0), (0, 1, 2), "red"),
           ((1, 3, 2), (0, 0, 0), (1, 1, 2), "red"),
           ((0, 1, 3), (0, 0, 0), None, "glass"),
           ((0, 2, 3), (0, 0, 0), None, "red"),
           ((0, 1, 3), (0, 0, 0), (0, 1, 2), "tex"),
           ((4, 5, 6), (1, 1, 1), None, "red")]


@pytest.fixture
def obj_dir(tmp_path):
    af.write_obj_fixture(tmp_path)
    return tmp_path


def _oracle_faces(parsed, vbase=0, nbase=0, tbase=0):
    out = []
    for o in parsed["objects"]:
        for g in o["groups"]:
            m = parsed["materials"][parsed["material_index"].get(g["material"], 0)]["name"]
            for k in range(0, len(g["pos"]) - 2, 3):
                uv = tuple(t + tbase for t in g["uv"][k:k + 3]) if g["format"] == "VTN" else None
                out.append((tuple(x + vbase for x in g["pos"][k:k + 3]),
                            tuple(x + nbase for x in g["nrm"][k:k + 3]), uv, m))
    return out


def _mat_kind(host, ptr):
    md = host.arrays["materials"]
    n = md[ptr]
    ids = [int(md[ptr + 8])] if n == 1 else [int(md[ptr + 8]), int(md[ptr + 20])]
    return n, ids


def _tex(host, off):
    return host.arrays["textures"][off:]


def test_obj_faces_and_materials(obj_dir):
    parsed = ao.parse_obj(obj_dir / "a.obj")
    assert _oracle_faces(parsed) == FACES_A
    host = ca.HostScene.from_files([obj_dir / "a.obj"], width=32, height=24)
    f = host.faces
    assert host.num_faces == len(FACES_A) and (host.width, host.height) == (32, 24)
    none = 0xFFFFFFFF
    mat_of = {}
    for i, (p, n, uv, m) in enumerate(FACES_A):
        assert tuple(f["p"][i]) == p and tuple(f["vn"][i]) == n
        assert tuple(f["uv"][i]) == (uv if uv else (none,) * 3)
        assert tuple(f["vt"][i]) == (none,) * 3 and f["lightPtr"][i] == 0xFFFF
        mat_of.setdefault(m, int(f["matPtr"][i]))
        assert int(f["matPtr"][i]) == mat_of[m]
    assert mat_of["red"] < mat_of["glass"] < mat_of["tex"]  # created in file order
    verts = host.vertices[:7, :3]
    np.testing.assert_array_equal(verts, [[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1]])
    np.testing.assert_array_equal(host.arrays["uvs"].view(np.float32)[:6], [0, 1, 1, 1, 0, 0.75])
    np.testing.assert_array_equal(host.arrays["normals"].view(np.float32).reshape(-1, 4)[:2, :3], [[0, 0, 1], [0, 1, 0]])
    # red: one Oren-Nayar element whose reflectance is the Kd constant
    assert _mat_kind(host, mat_of["red"]) == (1, [0])
    md = host.arrays["materials"]
    r_off = int(md[mat_of["red"] + 12:mat_of["red"] + 16].view(np.uint32)[0])
    t = _tex(host, r_off)
    assert t[0] == 0 and np.allclose(t[16:28].view(np.float32), [0.8, 0.1, 0.1])
    # glass (Ni 1.5): specular reflection + transmission, etaExt = Ni, etaInt = 1
    assert _mat_kind(host, mat_of["glass"]) == (2, [1, 2])
    g = mat_of["glass"]
    assert md[g + 28:g + 36].view(np.float32).tolist() == [1.5, 1.0]
    # tex: the Kd image (RGBA, read without gamma correction), used as the alpha texture
    assert _mat_kind(host, mat_of["tex"]) == (1, [0])
    k_off = int(md[mat_of["tex"] + 12:mat_of["tex"] + 16].view(np.uint32)[0])
    t = _tex(host, k_off)
    assert t[0] == 2 and t[4:12].view(np.uint32).tolist() == [6, 4]
    data = t[int(t[12:16].view(np.int32)[0]):][:6 * 4 * 4].reshape(4, 6, 4)
    np.testing.assert_array_equal(data, ca.load_image(obj_dir / "tex.png", False)[0])
    alpha = f["alphaTexPtr"]
    assert alpha[4] == k_off and all(alpha[i] == 0xFFFFFFFF for i in (0, 1, 2, 3, 5))


def test_two_models_share_texture_files(obj_dir):
    """A second model indexes its own vertices and gets its own materials
    (names are prefixed with the model number), while an image file loaded
    once is shared; emission turns a model's faces into lights."""
    one = ca.HostScene.from_files([obj_dir / "a.obj"])
    two = ca.HostScene.from_files([obj_dir / "a.obj", obj_dir / "a.obj"], emissions=[None, (2.0, 1.0, 0.5)])
    n = len(FACES_A)
    assert two.num_faces == 2 * n and two.num_lights == n
    f = two.faces
    parsed = ao.parse_obj(obj_dir / "a.obj")
    want = _oracle_faces(parsed, vbase=7, nbase=2, tbase=3)
    for i, (p, nn, uv, _) in enumerate(want):
        assert tuple(f["p"][n + i]) == p and tuple(f["vn"][n + i]) == nn
        assert tuple(f["uv"][n + i]) == (uv if uv else (0xFFFFFFFF,) * 3)
    assert (f["lightPtr"][:n] == 0xFFFF).all() and (f["lightPtr"][n:] != 0xFFFF).all()
    assert len(set(f["matPtr"][n:].tolist()) & set(f["matPtr"][:n].tolist())) == 0
    assert f["alphaTexPtr"][4] == f["alphaTexPtr"][n + 4] != 0xFFFFFFFF
    np.testing.assert_array_equal(one.faces, f[:n])


def test_obj_without_materials_and_errors(tmp_path):
    (tmp_path / "m.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1//1 2//1 3//1\nvn 0 0 1\n")
    host = ca.HostScene.from_files([tmp_path / "m.obj"])
    assert host.num_faces == 1
    assert _mat_kind(host, int(host.faces["matPtr"][0])) == (1, [0])
    with pytest.raises(ca.ClrError, match="not an .obj"):
        ca.HostScene.from_files([tmp_path / "m.ply"])
    with pytest.raises(ca.ClrError, match="cannot read"):
        ca.HostScene.from_files([tmp_path / "none.obj"])
    (tmp_path / "e.obj").write_text("# nothing\n")
    with pytest.raises(ca.ClrError, match="no vertices"):
        ca.HostScene.from_files([tmp_path / "e.obj"])


def _env_distribution(host):
    """(multiplier, marginal pdf, marginal cdf, conditional pdfs, conditional
    cdfs) of the scene's image-based environment: others[1] -> environment
    head -> light property -> ImageBasedEnvLElem -> ContinuousConsts2D_H."""
    od, md = host.arrays["others"], host.arrays["materials"]
    u32 = lambda b, o: int(b[o:o + 4].view(np.uint32)[0])
    i32 = lambda b, o: int(b[o:o + 4].view(np.int32)[0])
    f32 = lambda b, o, n=1: b[o:o + 4 * n].view(np.float32)
    lp = u32(od, u32(od, 4))
    assert md[lp] == 1  # one emission element
    e = (lp + 1 + 3) // 4 * 4
    assert md[e] == 0  # image based
    mult, d = float(f32(md, e + 8)[0]), u32(md, e + 12)
    assert od[d] == 2 and od[d + 8] == 1  # ContinuousConsts2D_H over a ContinuousConsts1D marginal
    sy = u32(od, d + 12)
    mpdf, mcdf = f32(od, d + i32(od, d + 28), sy), f32(od, d + i32(od, d + 32), sy + 1)
    kids = d + i32(od, d + 4)
    cp, cc = [], []
    for i in range(sy):
        c = kids + 28 * i
        sx = u32(od, c + 4)
        assert f32(od, c + 16)[0] == np.float32(1.0) / np.float32(sx)
        cp.append(f32(od, c + i32(od, c + 20), sx))
        cc.append(f32(od, c + i32(od, c + 24), sx + 1))
    return mult, mpdf, mcdf, np.array(cp), np.array(cc)


def test_environment_distribution_from_exr(obj_dir):
    """An OpenEXR environment adds one light at infinity whose 2-D importance
    distribution is, per row, the pixels' luminance (0.2126 R + 0.7152 G +
    0.0722 B) normalised over the row, and over rows the row means times
    sin(pi (i + 0.5) / H) (the solid angle of a latitude band) normalised;
    every CDF rises to exactly 1."""
    h, w = 12, 20
    rng = _rng(8)
    env = {c: (0.05 + rng.random((h, w)) * (1 + 3 * k)).astype(np.float16) for k, c in enumerate("RGB")}
    af.write_exr(obj_dir / "env.exr", env, af.ZIP)
    host = ca.HostScene.from_files([obj_dir / "a.obj"], env_path=obj_dir / "env.exr", env_multiplier=2.0)
    base = ca.HostScene.from_files([obj_dir / "a.obj"])
    assert host.num_lights == base.num_lights + 1
    mult, mpdf, mcdf, cpdf, ccdf = _env_distribution(host)
    assert mult == 2.0 and mpdf.size == h and cpdf.shape == (h, w)
    R, G, B = (env[c].astype(np.float64) for c in "RGB")
    lum = 0.2126 * R + 0.7152 * G + 0.0722 * B
    np.testing.assert_allclose(cpdf, lum / lum.mean(axis=1, keepdims=True), rtol=1e-5)
    band = lum.mean(axis=1) * np.sin(np.pi * (np.arange(h) + 0.5) / h)
    np.testing.assert_allclose(mpdf, band / band.mean(), rtol=1e-5)
    for cdf, pdf in [(mcdf, mpdf)] + list(zip(ccdf, cpdf)):
        assert cdf[0] == 0.0 and cdf[-1] == 1.0 and (np.diff(cdf) >= 0).all()
        np.testing.assert_allclose(np.diff(cdf), pdf / pdf.size, rtol=1e-4, atol=1e-7)


def test_header_only_cpp_api(obj_dir, tmp_path):
    """The C++ drop-in (include/clear_amd_loaders.hpp) compiles on its own
    with g++ and zlib and loads the fixture as the library does."""
    import shutil
    import subprocess

    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "probe.cpp"
    src.write_text("""
#include <cstdio>
#include "clear_amd_loaders.hpp"
using namespace clr::host;
int main(int argc, char** argv) {
    Scene scene;
    std::string err;
    if (!loadModel(argv[1], &scene, UINT64_MAX, UINT64_MAX, &err)) { std::printf("error %s\\n", err.c_str()); return 1; }
    std::vector<uint8_t> px; uint32_t w = 0, h = 0; ColorChannel::Value c;
    if (!loadImage(argv[2], &px, &w, &h, &c, false, &err)) { std::printf("error %s\\n", err.c_str()); return 1; }
    std::printf("%llu %llu %u %u %d %zu\\n", (unsigned long long)scene.numFaces(),
                (unsigned long long)scene.numVertices(), w, h, (int)c, px.size());
    return 0;
}
""")
    exe = tmp_path / "probe"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(root, "include"), str(src), "-o", str(exe),
                    "-lz"], check=True, capture_output=True)
    out = subprocess.run([str(exe), str(obj_dir / "a.obj"), str(obj_dir / "tex.png")], check=True,
                         capture_output=True, text=True).stdout.split()
    assert out == [str(len(FACES_A)), "7", "6", "4", "1", str(6 * 4 * 4)]

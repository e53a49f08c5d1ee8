"""Pins the post-processing oracle (oracle/post_oracle.cpp, the reference's
scatter form of post_processing.cl) by closed forms and layout invariants,
and the host BMP writer (clr_save_bmp) by its header fields."""
import math
import struct

import numpy as np
import pytest

import clear_amd as ca
import oracle as orc

GAUSS = np.array([
    0.000763447, 0.001831415, 0.003421534, 0.004978302, 0.005641155, 0.004978302, 0.003421534, 0.001831415,
    0.000763447, 0.001831415, 0.004393336, 0.008207832, 0.011942326, 0.013532428, 0.011942326, 0.008207832,
    0.004393336, 0.001831415, 0.003421534, 0.008207832, 0.01533425, 0.022311201, 0.025281903, 0.022311201,
    0.01533425, 0.008207832, 0.003421534, 0.004978302, 0.011942326, 0.022311201, 0.032462606, 0.036784952,
    0.032462606, 0.022311201, 0.011942326, 0.004978302, 0.005641155, 0.013532428, 0.025281903, 0.036784952,
    0.041682812, 0.036784952, 0.025281903, 0.013532428, 0.005641155, 0.004978302, 0.011942326, 0.022311201,
    0.032462606, 0.036784952, 0.032462606, 0.022311201, 0.011942326, 0.004978302, 0.003421534, 0.008207832,
    0.01533425, 0.022311201, 0.025281903, 0.022311201, 0.01533425, 0.008207832, 0.003421534, 0.001831415,
    0.004393336, 0.008207832, 0.011942326, 0.013532428, 0.011942326, 0.008207832, 0.004393336, 0.001831415,
    0.000763447, 0.001831415, 0.003421534, 0.004978302, 0.005641155, 0.004978302, 0.003421534, 0.001831415,
    0.000763447], np.float32).reshape(9, 9)


def impulse(h, w, y, x, v):
    img = np.zeros((h, w, 4), np.float32)
    img[y, x, :3] = v
    return img


def test_gaussian_table_is_a_sampled_isotropic_gaussian():
    # symmetric, separable (rank 1), peaked at the centre, sums to 1 within rounding
    assert np.allclose(GAUSS, GAUSS.T) and np.allclose(GAUSS, GAUSS[::-1, ::-1])
    s = np.linalg.svd(GAUSS.astype(np.float64), compute_uv=False)
    assert s[1] < 1e-6 * s[0]
    assert abs(GAUSS.astype(np.float64).sum() - 1.0) < 1e-4
    assert GAUSS.argmax() == 40


def test_scaling():
    rng = np.random.default_rng(1)
    img = rng.uniform(0, 100, (5, 7, 4)).astype(np.float32)
    out = orc.post_scaling(img, 16)
    np.testing.assert_array_equal(out[..., :3], img[..., :3] * np.float32(1.0 / 16))


def test_gaussian_interior_impulse_and_border_mass():
    out = orc.post_gaussian(impulse(20, 21, 10, 9, 8.0), 2)
    expect = np.zeros((20, 21), np.float32)
    expect[6:15, 5:14] = (GAUSS * np.float32(1.0 / 2)) * np.float32(8.0)
    np.testing.assert_array_equal(out[..., 0], expect)
    # at a corner every tap is clamped into the image: the mass is preserved
    # and the corner pixel receives all taps with both offsets <= 0
    out = orc.post_gaussian(impulse(12, 12, 0, 0, 1.0), 1)
    assert abs(float(out[..., 1].sum()) - float(GAUSS.astype(np.float64).sum())) < 1e-5
    assert abs(float(out[0, 0, 1]) - float(GAUSS[:5, :5].astype(np.float64).sum())) < 1e-6


def test_bloom_impulse_closed_form():
    spp, v = 4, 35.5  # excess 5.5 -> radius 5
    img = impulse(21, 21, 10, 10, v * spp)
    img[3, 3, :3] = 2.0 * spp  # dim pixel: passes through as 2
    out = orc.post_bloom(img, spp)
    e = np.float32(v) - np.float32(30.0)
    for dy in range(-6, 7):
        for dx in range(-6, 7):
            got = out[10 + dy, 10 + dx, 0]
            if dx == 0 and dy == 0:
                assert got == 30.0
            elif max(abs(dx), abs(dy)) <= 5:
                assert abs(got - math.exp(-math.hypot(dx, dy)) * e) <= 1e-6 * e
            else:
                assert got == 0.0
    assert out[3, 3, 0] == 2.0
    # radius capped at 7, border taps clamped (mass of a corner impulse)
    out = orc.post_bloom(impulse(9, 9, 0, 0, 100.0), 1)
    mass = 30.0 + sum(math.exp(-math.hypot(dx, dy)) * 70.0 for dy in range(-7, 8) for dx in range(-7, 8)
                      if (dx, dy) != (0, 0))
    assert abs(float(out[..., 2].sum()) - mass) < 1e-3


def test_tone_mapping_layout_and_curve():
    w, h = 5, 3
    rb = ca.bmp_row_bytes(w)
    assert rb == 3 * w + w % 4 and rb % 4 == 0
    img = np.zeros((h, w, 4), np.float32)
    img[0, 0, 0] = 1.0           # top-left pure red
    img[h - 1, w - 1, 2] = 50.0  # bottom-right blue beyond saturation
    img[1, 2, :3] = 1.0          # grey
    out = orc.post_tone(img, rb).reshape(h, rb)
    # top image row is the last BMP row; pixel bytes are B, G, R; padding 0
    top = out[h - 1]
    assert top[0] == 0 and top[1] == 0 and top[2] > 0
    assert (out[:, 3 * w:] == 0).all()
    assert out[0, 3 * (w - 1) + 0] > 0 and out[0, 3 * (w - 1) + 2] == 0
    sat = orc.post_tone(np.where(img > 30, 30, img).astype(np.float32), rb)
    np.testing.assert_array_equal(out.reshape(-1), sat)
    # grey 1.0: Y = 0.9999, byte = trunc(255 * ((1 - e^-Y) / Y)^(1/2.2))
    y = 0.2989 + 0.5866 + 0.1144
    g = int(255.0 * ((1.0 - math.exp(-y)) / y) ** (1 / 2.2))
    assert abs(int(out[1, 3 * 2]) - g) <= 1
    # monotone in the input
    ramp = np.zeros((1, 64, 4), np.float32)
    ramp[0, :, :3] = np.linspace(0, 40, 64, dtype=np.float32)[:, None]
    b = orc.post_tone(ramp, ca.bmp_row_bytes(64))[: 3 * 64 : 3].astype(int)
    assert (np.diff(b) >= 0).all() and b[0] == 0


def test_save_bmp_header(tmp_path):
    w, h = 7, 3
    rb = ca.bmp_row_bytes(w)
    px = np.arange(rb * h, dtype=np.uint32).astype(np.uint8)
    p = tmp_path / "x.bmp"
    ca.save_bmp(str(p), px, w, h)
    data = p.read_bytes()
    assert data[:2] == b"BM" and len(data) == 54 + rb * h
    size, off, info, ww, hh, planes, bits, comp, dsize, xppm, yppm = struct.unpack_from("<I4xIIiiHHIIii", data, 2)
    assert (size, off, info, ww, hh, planes, bits, comp, dsize, xppm, yppm) == (54 + rb * h, 54, 40, w, h, 1, 24, 0,
                                                                                 rb * h, 1, 1)
    assert data[54:] == px.tobytes()

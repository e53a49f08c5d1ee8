"""Pins the CPU render oracle (oracle/render_oracle.cpp) with closed forms.

The reference has no tests, so the oracle's shading pieces are fixed by
mathematics: Lambert reduces to R/pi, Fresnel matches its textbook values,
every BxDF's sampled (direction, pdf, value) agrees with its own evaluation,
pdfs integrate to at most 1 (Beckmann's to 1), BRDFs are reciprocal, the
concentric map is an area-preserving square->disk map, and a white furnace
returns the closed-form image (camera weight x albedo x environment).
"""
import math

import numpy as np
import pytest

import clear_amd as ca
import oracle as orc

PI = np.float32(np.pi)


def hemi(th, ph):
    return [math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)]


def test_lambert_is_r_over_pi():
    r = orc.bxdf_probe(0, 0.0, 0.0, [0.5, 0.25, 1.0], hemi(0.3, 0.1), hemi(1.1, 2.0), 0.3, 0.7)
    np.testing.assert_allclose(r["f"], np.float32([0.5, 0.25, 1.0]) / PI, rtol=1e-7)
    assert r["pdf"] == pytest.approx(math.cos(1.1) / math.pi, rel=1e-6)
    # opposite hemispheres -> 0 (reflection only)
    r2 = orc.bxdf_probe(0, 0.0, 0.0, [1, 1, 1], hemi(0.3, 0.1), [0.1, 0.2, -0.97], 0.3, 0.7)
    assert r2["f"] == [0.0, 0.0, 0.0]


def test_oren_nayar_reduces_and_peaks():
    # sigma -> 0 gives A = 1, B = 0 (reflection.cl:244-247)
    a = orc.bxdf_probe(0, 1e-4, 0.0, [1, 1, 1], hemi(0.5, 0.0), hemi(0.7, 0.0), 0.1, 0.1)
    assert a["f"][0] == pytest.approx(1 / math.pi, rel=1e-6)
    # sigma > 0 brightens retro-reflection relative to the forward direction
    retro = orc.bxdf_probe(0, 0.5, 0.0, [1, 1, 1], hemi(1.0, 0.0), hemi(1.0, 0.0), 0.1, 0.1)["f"][0]
    fwd = orc.bxdf_probe(0, 0.5, 0.0, [1, 1, 1], hemi(1.0, 0.0), hemi(1.0, math.pi), 0.1, 0.1)["f"][0]
    assert retro > fwd


def test_fresnel_closed_forms():
    f0 = orc.fresnel_probe(2, [0, 0, 0], [0, 0, 0], 1.0, 1.5, 1.0)
    assert f0[0] == pytest.approx(((1.5 - 1) / (1.5 + 1)) ** 2, rel=1e-6)
    # leaving the dense medium beyond the critical angle: total internal reflection
    tir = orc.fresnel_probe(2, [0, 0, 0], [0, 0, 0], 1.0, 1.5, -math.cos(1.2))
    assert tir == [1.0, 1.0, 1.0]
    eta, k = [0.2, 0.5, 1.4], [3.9, 2.4, 1.6]
    c = orc.fresnel_probe(1, eta, k, 0, 0, 1.0)
    for i in range(3):
        expect = ((eta[i] - 1) ** 2 + k[i] ** 2) / ((eta[i] + 1) ** 2 + k[i] ** 2)
        assert c[i] == pytest.approx(expect, rel=1e-5)
    assert orc.fresnel_probe(0, [0, 0, 0], [0, 0, 0], 0, 0, 0.3) == [1.0, 1.0, 1.0]


@pytest.mark.parametrize("kind,p0,p1", [(0, 0.4, 0), (4, 80.0, 20.0), (5, 0, 0), (6, 0.2, 1.5),
                                        (6, 0.45, 0.0)])
def test_sampling_consistent_with_evaluation(kind, p0, p1):
    rng = np.random.default_rng(kind)
    R = [0.7, 0.5, 0.3]
    for _ in range(40):
        wo = hemi(rng.uniform(0, 1.3), rng.uniform(0, 6.28))
        r = orc.bxdf_probe(kind, p0, p1, R, wo, [0, 0, 1], rng.random(), rng.random())
        if r["spdf"] == 0.0:
            continue
        e = orc.bxdf_probe(kind, p0, p1, R, wo, r["wi"], 0.5, 0.5)
        np.testing.assert_allclose(r["sf"], e["f"], rtol=1e-5, atol=1e-7)
        assert r["spdf"] == pytest.approx(e["pdf"], rel=1e-5)


@pytest.mark.parametrize("kind,p0,p1", [(0, 0.3, 0), (3, 0.2, 0.3), (4, 50.0, 50.0), (5, 0, 0),
                                        (6, 0.3, 1.5)])
def test_reciprocity(kind, p0, p1):
    rng = np.random.default_rng(10 + kind)
    for _ in range(30):
        a = hemi(rng.uniform(0, 1.4), rng.uniform(0, 6.28))
        b = hemi(rng.uniform(0, 1.4), rng.uniform(0, 6.28))
        fab = orc.bxdf_probe(kind, p0, p1, [1, 1, 1], a, b, 0.5, 0.5)["f"][0]
        fba = orc.bxdf_probe(kind, p0, p1, [1, 1, 1], b, a, 0.5, 0.5)["f"][0]
        assert fab == pytest.approx(fba, rel=2e-5, abs=1e-9)


def mc_pdf_integral(kind, p0, p1, wo, n=60000, seed=1):
    """Integral of pdf(wo, wi) over the sphere by uniform sphere sampling."""
    rng = np.random.default_rng(seed)
    z = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    rr = np.sqrt(1 - z * z)
    tot = 0.0
    for i in range(n):
        wi = [rr[i] * math.cos(ph[i]), rr[i] * math.sin(ph[i]), z[i]]
        tot += orc.bxdf_probe(kind, p0, p1, [1, 1, 1], wo, wi, 0.5, 0.5)["pdf"]
    return tot * 4 * np.pi / n


def test_pdfs_integrate_to_one():
    # Beckmann with wo at the normal: all of D's mass maps above the horizon
    # for small alpha, so the pdf integrates to 1 (normalisation of D).
    assert mc_pdf_integral(6, 0.2, 0.0, [0, 0, 1]) == pytest.approx(1.0, abs=0.03)
    # Cosine-weighted diffuse: the reference's pdf is |cos(wi)|/pi with no
    # same-hemisphere test (reflection.cl:587-588), so it is 1 over the upper
    # hemisphere and 2 over the whole sphere.
    assert mc_pdf_integral(0, 0.0, 0.0, hemi(0.6, 0.0), n=20000) == pytest.approx(2.0, abs=0.06)


def test_white_furnace_lambert_estimate():
    """E[f cos / pdf] = R for the Lambert sampler (no Monte Carlo error: exact per sample)."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        wo = hemi(rng.uniform(0, 1.4), rng.uniform(0, 6.28))
        r = orc.bxdf_probe(0, 0.0, 0.0, [0.8, 0.8, 0.8], wo, [0, 0, 1], rng.random(), rng.random())
        est = r["sf"][0] * abs(r["wi"][2]) / r["spdf"]
        assert est == pytest.approx(0.8, rel=1e-5)


def test_concentric_disk_is_area_preserving():
    assert orc.disk_probe(0.5, 0.5) == [0.0, 0.0]
    for u in [(0, 0), (1, 0), (0, 1), (1, 1)]:
        x, y = orc.disk_probe(*u)
        assert math.hypot(x, y) == pytest.approx(1.0, abs=1e-6)
    g = np.random.default_rng(4)
    pts = np.array([orc.disk_probe(a, b) for a, b in g.random((20000, 2))])
    r2 = (pts ** 2).sum(1)
    assert r2.max() <= 1.0 + 1e-6
    # uniform on the disk: r^2 is uniform on [0, 1]
    assert np.mean(r2) == pytest.approx(0.5, abs=0.01)
    assert np.mean(r2 < 0.25) == pytest.approx(0.25, abs=0.01)


def test_seed_states_follow_the_expansion():
    st = orc.seed_states(4, 2).reshape(-1, 4)
    for i in range(8):
        s = (i ^ 0x9E3779B9) & 0xFFFFFFFF
        for k in range(4):
            s = (1812433253 * (s ^ (s >> 30)) + k) & 0xFFFFFFFF
            assert st[i, k] == s


def camera_weight(width, height, fov_deg):
    """alpha of a pinhole camera ray through the pixel centre: |cos| / dirPDF
    = (-d.z)^4 * virtualPlaneArea with d the unit camera-space direction."""
    a = width / height
    t = math.tan(math.radians(fov_deg) / 2)
    xs = ((np.arange(width) + 0.5) / width * 2 - 1) * a * t
    ys = (1 - (np.arange(height) + 0.5) / height * 2) * t
    X, Y = np.meshgrid(xs, ys)
    dz = 1 / np.sqrt(1 + X * X + Y * Y)
    return dz ** 4 * a * t * t


@pytest.mark.parametrize("albedo", [0.5, 1.0])
def test_white_furnace_image(albedo):
    W = H = 48
    spp = 256
    s = ca.HostScene(ca.SCENE_FURNACE, W, H, detail=4, albedo=albedo, fov_y=40.0)
    o = orc.lbvh_build(s.vertices, s.faces, 10)
    st = orc.seed_states(W, H)
    px = np.zeros(W * H * 4, np.float32)
    orc.render(s.arrays, s.num_faces, o["inodes"], o["leaves"], W, H, spp, 64, st, px)
    img = px.reshape(H, W, 4)[:, :, 0] / spp
    ratio = img / camera_weight(W, H, 40.0)
    yy, xx = np.mgrid[0:H, 0:W]
    r = np.hypot(xx + 0.5 - W / 2, yy + 0.5 - H / 2) / (W / 2)
    inside = r < 0.45  # sphere of radius 0.5 at distance 2, fov 40: edge ~0.72
    outside = r > 0.85
    assert ratio[outside].mean() == pytest.approx(1.0, abs=0.01)
    assert ratio[inside].mean() == pytest.approx(albedo, abs=0.02 * albedo + 0.005)

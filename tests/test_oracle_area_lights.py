"""Pins the area-light half of the render oracle (oracle/render_oracle.cpp,
following light.cl:138-292, pathtracer.cl:106-128,154-163 and
reflection.cl:629-723) with checks that mathematics fixes:

* direct lighting of a Lambert floor under diffuse polygon emitters equals
  Lambert's polygon formula, E = L/2 * sum_i theta_i (n . g_i) with L = M / pi,
  for the combined (MIS) estimator and for light samples or BSDF samples
  alone, and is exactly 0 behind a one-sided emitter;
* the three strategies converge to the same Cornell-box image (depth 5);
* the BSDF mixture's sample_fs draws from the density fs_pdf reports: the
  estimate E[fs cos / pdf] over stratified samples equals a quadrature of
  fs cos, and fs_pdf integrates to (1 + the glossy lobe's mass) / 2.

Each group is shown to detect mistakes planted in the oracle (a dropped
1/area, a dropped |cos| at the light, swapped MIS terms, a dropped dist^2,
the light choice's probability dropped, emission without 1/pi or on both
sides, a wrong mixture average or pick).
"""
import numpy as np
import pytest

import clear_amd as ca
import oracle as orc

M_EMIT = np.array([4.0, 2.0, 1.0])  # CLR_SCENE_AREA_LIGHTS emission
RHO = 0.5                           # its default floor albedo
STRATEGIES = {"mis": 0, "light": ca.RENDER_LIGHT_SAMPLING_ONLY, "bsdf": ca.RENDER_BSDF_SAMPLING_ONLY}


def polygon_irradiance_factor(x, n, poly):
    """Lambert's formula: irradiance / radiance at points x [m, 3] with
    normal n from a uniform polygon poly [k, 3] (vertex order either way)."""
    tot = np.zeros(len(x))
    for i in range(len(poly)):
        a = poly[i][None, :] - x
        b = poly[(i + 1) % len(poly)][None, :] - x
        cos_ab = (a * b).sum(1) / (np.linalg.norm(a, axis=1) * np.linalg.norm(b, axis=1))
        theta = np.arccos(np.clip(cos_ab, -1.0, 1.0))
        g = np.cross(a, b)
        g /= np.linalg.norm(g, axis=1, keepdims=True)
        tot += theta * (g @ n)
    return 0.5 * np.abs(tot)


def test_polygon_formula_closed_forms():
    """The formula itself: a regular 4096-gon of radius r at height
    h over the point gives pi r^2 / (r^2 + h^2); a square seen from its
    corner's normal at height h gives the textbook form factor."""
    r, h = 0.7, 0.4
    ang = 2 * np.pi * np.arange(4096) / 4096
    poly = np.stack([r * np.cos(ang), np.full(4096, h), r * np.sin(ang)], -1)
    e = polygon_irradiance_factor(np.zeros((1, 3)), np.array([0.0, 1.0, 0.0]), poly)[0]
    assert e == pytest.approx(np.pi * r * r / (r * r + h * h), rel=1e-5)
    # Unit square [0,1]^2 at height 1 over the origin (corner below): pi * F with
    # F = 1/(2 pi) * (2 * atan(1/sqrt 2) / sqrt 2) for a 1x1 rectangle at unit height.
    sq = np.array([[0, 1, 0], [1, 1, 0], [1, 1, 1], [0, 1, 1]], float)
    e = polygon_irradiance_factor(np.zeros((1, 3)), np.array([0.0, 1.0, 0.0]), sq)[0]
    f = (1 / (2 * np.pi)) * 2 * (1 / np.sqrt(2)) * np.arctan(1 / np.sqrt(2))
    assert e == pytest.approx(np.pi * f, rel=1e-9)


def _direct_at(host, film):
    """For film points [n, 4] (px, py, lu0, lu1): alpha * rho / pi * E(x)
    per channel at the floor point x the camera ray reaches, whether the ray
    reaches the floor unhidden, and whether x lies in front of some emitter."""
    faces = host.faces
    V = host.vertices[:, :3].astype(np.float64)
    lights = [V[faces["p"][f].astype(np.int64)] for f in np.nonzero(faces["lightPtr"] != 0xFFFF)[0]]
    o, d, pdf, zaxis = orc.camera_rays(host.arrays, film)
    t = -o[:, 1] / d[:, 1]
    x = o + t[:, None] * d
    ok = (t > 0) & (np.abs(x[:, 0]) <= 2.0) & (np.abs(x[:, 2]) <= 2.0)
    lit = np.zeros(len(x), bool)
    irr = np.zeros(len(x))
    for tri in lights:
        e1, e2 = tri[1] - tri[0], tri[2] - tri[0]
        p = np.cross(d, e2)
        det = p @ e1
        s = o - tri[0]
        u = (s * p).sum(1) / det
        q = np.cross(s, e1)
        v = (d * q).sum(1) / det
        tt = (q @ e2) / det
        ok &= ~((u >= 0) & (v >= 0) & (u + v <= 1) & (tt > 0) & (tt < t))  # the emitter hides the floor
        nrm = np.cross(e1, e2)
        nrm /= np.linalg.norm(nrm)
        front = ((x - tri[0]) @ nrm) > 0  # emission leaves the front side only
        lit |= front
        irr += np.where(front, polygon_irradiance_factor(x, np.array([0.0, 1.0, 0.0]), tri), 0.0)
    alpha = np.abs(d @ zaxis) / pdf
    return (alpha * irr)[:, None] * (RHO / np.pi) * (M_EMIT / np.pi)[None, :], ok, lit


def _film_grid(W, H, j):
    gx, gy, jx, jy = np.meshgrid(np.arange(W), np.arange(H), j, j, indexing="ij")
    return np.stack([gx + jx, gy + jy, np.full(gx.shape, 0.5), np.full(gx.shape, 0.5)], -1).reshape(-1, 4)


def expected_direct(host, k=8):
    """Expected pixel values (H, W, 3) of the floor's direct lighting at
    maxDepth 1 (the mean over a k x k grid of film points of
    alpha * rho / pi * E(x); alpha = |cos| / dirPDF of the camera ray, whose
    model is pinned in test_oracle_shading_pins.py), the mask of pixels whose
    footprint sees only the floor, and the mask of pixels whose footprint
    lies wholly behind every emitter (masks from a (k+1)^2 grid reaching the
    pixel's edges)."""
    W, H = host.width, host.height
    val, _, _ = _direct_at(host, _film_grid(W, H, (np.arange(k) + 0.5) / k))
    val = val.reshape(W, H, k * k, 3).mean(2).transpose(1, 0, 2)
    _, ok, lit = _direct_at(host, _film_grid(W, H, np.arange(k + 1) / k))
    mask = ok.reshape(W, H, -1).all(2).T
    dark = ~lit.reshape(W, H, -1).any(2).T
    return val, mask, dark


def render_batches(host, flags, batches, spp, max_depth, ob=None):
    """`batches` oracle renders of `spp` samples continuing one random stream
    per pixel: float64 [batches, H, W, 3] of per-batch pixel means."""
    W, H = host.width, host.height
    ob = ob or orc.lbvh_build(host.vertices, host.faces, 10)
    st = orc.seed_states(W, H)
    out = []
    for _ in range(batches):
        px = np.zeros(W * H * 4, np.float32)
        orc.render(host.arrays, host.num_faces, ob["inodes"], ob["leaves"], W, H, spp, max_depth, st, px,
                   flags=flags)
        out.append(px.reshape(H, W, 4)[:, :, :3].astype(np.float64) / spp)
    return np.array(out)


_HOSTS = {}


def area_scene(detail):
    if detail not in _HOSTS:
        host = ca.HostScene(ca.SCENE_AREA_LIGHTS, detail=detail)
        _HOSTS[detail] = (host, expected_direct(host), orc.lbvh_build(host.vertices, host.faces, 10))
    return _HOSTS[detail]


def irradiance_check(detail, flags, batches=16, spp=128):
    """(passes, report) of the rendered direct lighting against the closed form."""
    host, (exp, mask, dark), ob = area_scene(detail)
    imgs = render_batches(host, flags, batches, spp, 1, ob)
    mean = imgs.mean(0)
    se = imgs.std(0, ddof=1) / np.sqrt(batches)
    tot = imgs[:, mask].sum((1, 2)) / exp[mask].sum()  # per-batch flux ratio
    ratio, ratio_se = tot.mean(), tot.std(ddof=1) / np.sqrt(batches)
    bad = np.abs(mean - exp) > 5.0 * se + 2e-3 * exp.max()
    dark_ok = bool((mean[mask & dark] == 0.0).all())
    ok = abs(ratio - 1.0) <= 4.0 * ratio_se + 1e-3 and bad[mask].mean() <= 0.01 and dark_ok
    return ok, f"flux ratio {ratio:.5f} +- {ratio_se:.5f}, {bad[mask].mean():.4f} of pixels off, dark ok {dark_ok}"


@pytest.mark.parametrize("strategy", list(STRATEGIES))
@pytest.mark.parametrize("detail", [0, 1, 2])
def test_direct_lighting_equals_polygon_formula(detail, strategy):
    ok, rep = irradiance_check(detail, STRATEGIES[strategy])
    assert ok, rep


def test_floor_behind_a_one_sided_emitter_is_dark():
    host, (exp, mask, dark), _ = area_scene(2)
    assert (mask & dark).sum() > 100  # the scene has such pixels to check
    assert (exp[mask & dark] == 0.0).all()


@pytest.mark.parametrize("plant", [orc.PLANT_LIGHT_NO_INV_AREA, orc.PLANT_AREA_PDF_NO_INV_AREA,
                                   orc.PLANT_NEE_NO_COS_L, orc.PLANT_MIS_NO_COS_L, orc.PLANT_NEE_MIS_SWAPPED,
                                   orc.PLANT_IMPLICIT_NO_DIST2, orc.PLANT_NO_LIGHT_CHOICE_PROB,
                                   orc.PLANT_LE_NO_INV_PI, orc.PLANT_LE_TWO_SIDED])
def test_polygon_pin_detects_planted_mistake(plant):
    with orc.planted(plant):
        for detail in (1, 2):
            for flags in STRATEGIES.values():
                if not irradiance_check(detail, flags)[0]:
                    return
    pytest.fail(f"planted mistake {plant} passed every polygon-irradiance check")


def strategy_agreement(host, ob, batches=16, spp=64, depth=5, blk=4):
    """Pairwise agreement of MIS, light-only and BSDF-only renders: per pair
    the fraction of 4x4 blocks whose paired per-batch difference is beyond
    4 standard errors, and whether the flux difference is within 4 standard
    errors + 1e-3."""
    imgs = {k: render_batches(host, f, batches, spp, depth, ob) for k, f in STRATEGIES.items()}
    B, H, W, C = imgs["mis"].shape
    res = {}
    for a, b in (("mis", "light"), ("mis", "bsdf"), ("light", "bsdf")):
        d = (imgs[a] - imgs[b]).reshape(B, H // blk, blk, W // blk, blk, C).mean((2, 4))
        z = d.mean(0) / np.maximum(d.std(0, ddof=1) / np.sqrt(B), 1e-12)
        g = (imgs[a] - imgs[b]).reshape(B, -1).sum(1) / imgs[b].reshape(B, -1).sum(1).mean()
        flux_ok = abs(g.mean()) <= 4.0 * g.std(ddof=1) / np.sqrt(B) + 1e-3
        res[(a, b)] = ((np.abs(z) > 4.0).mean(), flux_ok)
    return res


_CORNELL = {}


def cornell():
    if not _CORNELL:
        host = ca.HostScene(ca.SCENE_CORNELL, 32, 32, detail=3)
        _CORNELL["v"] = (host, orc.lbvh_build(host.vertices, host.faces, 10))
    return _CORNELL["v"]


def test_strategies_converge_to_one_image():
    host, ob = cornell()
    for pair, (frac, flux_ok) in strategy_agreement(host, ob).items():
        assert frac <= 0.03 and flux_ok, f"{pair}: {frac:.3f} of blocks differ, flux agrees: {flux_ok}"


@pytest.mark.parametrize("plant", [orc.PLANT_LIGHT_NO_INV_AREA, orc.PLANT_AREA_PDF_NO_INV_AREA,
                                   orc.PLANT_NEE_NO_COS_L, orc.PLANT_MIS_NO_COS_L, orc.PLANT_NEE_MIS_SWAPPED,
                                   orc.PLANT_IMPLICIT_NO_DIST2, orc.PLANT_NO_LIGHT_CHOICE_PROB])
def test_strategy_pin_detects_planted_mistake(plant):
    host, ob = cornell()
    with orc.planted(plant):
        res = strategy_agreement(host, ob)
    assert any(frac > 0.03 or not flux_ok for frac, flux_ok in res.values()), res


# ---- BSDF mixture (reflection.cl:629-723) -----------------------------------

_CLUSTERS = {}


def mixture_scene():
    """The batch scene's materials: faces 0-3 Lambert, 4-7 Lambert under a
    Beckmann coat (alpha in [0.05, 0.5], dielectric eta 1.5)."""
    if not _CLUSTERS:
        host = ca.HostScene(ca.SCENE_CLUSTERS, 32, 32, num_triangles=64)
        _CLUSTERS["v"] = (host, [int(m) for m in host.faces["matPtr"][:8]])
    return _CLUSTERS["v"]


def quadrature(host, mat, wo, nth=1500, nph=256):
    """(integral of fs(wo, wi) |cos_i| over the upper hemisphere, integral of
    fs_pdf(wo, wi)) by the midpoint rule over half vectors h (wi = 2 (wo.h) h
    - wo, dwi = 4 (wo.h) dwh), which resolves narrow lobes."""
    th = (np.arange(nth) + 0.5) * (np.pi / 2) / nth
    ph = (np.arange(nph) + 0.5) * 2 * np.pi / nph
    T, P = np.meshgrid(th, ph, indexing="ij")
    h = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    c = h @ wo
    wi = 2 * c[:, None] * h - wo[None, :]
    ok = (c > 0) & (wi[:, 2] > 0)
    r = orc.bsdf_probe(host.arrays, mat, np.broadcast_to(wo, wi.shape), wi, [0.5, 0.5, 0.5])
    w = 4 * c * np.sin(T.reshape(-1)) * ok * (np.pi / 2 / nth) * (2 * np.pi / nph)
    alb = (r["f"].astype(np.float64) * (np.abs(wi[:, 2]) * w)[:, None]).sum(0)
    return alb, (r["pdf"].astype(np.float64) * w).sum()


def sampled_estimate(host, mat, wo, n=512):
    """E[fs cos / pdf] of sample_fs over a stratified grid: the element
    choice uc at the centre of each of two strata, (u0, u1) at n x n cell
    centres; a failed sample (pdf 0) counts 0."""
    g = (np.arange(n) + 0.5) / n
    U0, U1 = np.meshgrid(g, g, indexing="ij")
    est = np.zeros(3)
    for uc in (0.25, 0.75):
        u = np.stack([np.full(U0.size, uc), U0.reshape(-1), U1.reshape(-1)], -1)
        r = orc.bsdf_probe(host.arrays, mat, np.broadcast_to(wo, u.shape), [0.0, 0.0, 1.0], u)
        ok = r["spdf"] > 0
        pdf = np.where(ok, r["spdf"], 1.0).astype(np.float64)
        v = np.where(ok[:, None], r["sf"].astype(np.float64) * np.abs(r["wi"][:, 2:3]) / pdf[:, None], 0.0)
        est += v.mean(0) / 2
    return est


def wo_at(theta):
    return np.array([np.sin(theta), 0.0, np.cos(theta)])


@pytest.mark.parametrize("theta", [0.0, 0.6, 1.2])
@pytest.mark.parametrize("face", [0, 4, 5, 6, 7])
def test_mixture_sampling_matches_quadrature(face, theta):
    host, mats = mixture_scene()
    q, pdf_int = quadrature(host, mats[face], wo_at(theta))
    s = sampled_estimate(host, mats[face], wo_at(theta))
    np.testing.assert_allclose(s, q, rtol=2e-3)
    # A Lambert element's pdf has mass 1 over the upper hemisphere and a
    # glossy lobe's at most 1, so the average of the two is in (1/2, 1].
    assert 0.5 < pdf_int <= 1.0 + 1e-3


@pytest.mark.parametrize("face", [0, 4, 5, 6, 7])
def test_mixture_pdf_mass_at_normal_incidence(face):
    """At wo = n the Beckmann sampler's tan^2(theta_h) is exponential with
    mean alpha^2, so its samples stay above the horizon with probability
    1 - exp(-1 / alpha^2) >= 0.98 for alpha <= 0.5: the mixture's pdf
    integrates to 1 within 0.01 (Lambert alone: exactly 1)."""
    host, mats = mixture_scene()
    _, pdf_int = quadrature(host, mats[face], wo_at(0.0))
    assert pdf_int == pytest.approx(1.0, abs=0.01)


@pytest.mark.parametrize("plant", [orc.PLANT_MIX_NO_AVERAGE, orc.PLANT_MIX_OWN_PDF_ONLY, orc.PLANT_MIX_PICK_FIRST])
def test_mixture_pin_detects_planted_mistake(plant):
    host, mats = mixture_scene()
    wo = wo_at(0.6)
    with orc.planted(plant):
        for face in (4, 5, 6, 7):
            q, pdf_int = quadrature(host, mats[face], wo)
            s = sampled_estimate(host, mats[face], wo)
            if not np.allclose(s, q, rtol=2e-3) or not 0.5 < pdf_int <= 1.0 + 1e-3:
                return
    pytest.fail(f"planted mistake {plant} passed the mixture checks")

"""Pins of the oracle's material models (Lambertian, NVisii, Glass) and its
half rounding against what the mathematics fixes: an independent library
routine (numpy float16), textbook optics (Snell's law, unpolarized Fresnel
reflectance, total internal reflection), moments of the cosine-weighted
hemisphere, closed forms of the NVisii lobes at normal incidence,
reciprocity, and the lobe-selection weights of NVisii.h:853-870.

The reference has no numeric fixtures for these functions, so each test is
chosen so that a dropped term, a wrong sign or a swapped operand in the
oracle fails it."""
import math

import numpy as np
import pytest

from oracle import refcpu as R

MATTE = dict(type="matte")
GLASS = dict(type="pbr", transmission=1.0, ior=1.5)


def nvisii(metallic=0.0, roughness=0.5):
    return dict(type="pbr", metallic=metallic, roughness=roughness, transmission=0.0, ior=1.45)


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def test_half_round_matches_numpy_float16():
    rng = np.random.default_rng(7)
    xs = np.concatenate([
        rng.uniform(0, 1, 4000), rng.uniform(-2, 2, 2000), rng.uniform(0, 1e-4, 2000),  # includes subnormals
        10.0 ** rng.uniform(-8, 4.8, 2000),
        np.array([0.8, 0.5, 0.04, 1.0 - 0.03 ** 2, 65504.0, 65519.0, 65520.0, 1e6, 2.0 ** -24, 2.0 ** -25,
                  3 * 2.0 ** -25, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11]),
    ]).astype(np.float32)
    with np.errstate(over="ignore"):
        want = xs.astype(np.float16).astype(np.float32)
    got = np.array([R.half_round(float(x)) for x in xs], np.float32)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_glass_exact_normal_incidence_reflects():
    """At exactly normal incidence refractionDirection (Glass.h:20-68)
    normalizes a zero vector; with fminf/fmaxf clamping the Fresnel
    coefficient becomes 1 and every sample reflects back along N (the
    head-on branch meant for this case can never be taken)."""
    N = np.array([0, 0, 1], np.float32)
    s = R.bsdf_sample(GLASS, (1, 1, 1), N, N, 1000, in_medium=False, seed=3)
    assert not s["medium"].any()
    assert np.allclose(s["dir"], N, atol=1e-7)


@pytest.mark.parametrize("theta_deg", [0.5, 30.0, 60.0, 80.0])
def test_glass_fresnel_and_snell_entering(theta_deg):
    """Outside -> glass (ior 1.5): reflected fraction = unpolarized Fresnel
    reflectance; transmitted rays obey n1 sin(t1) = n2 sin(t2) and bend into
    the surface; reflected rays are mirror images."""
    th = math.radians(theta_deg)
    N = np.array([0, 0, 1], np.float32)
    wo = np.array([math.sin(th), 0.0, math.cos(th)], np.float32)  # towards the viewer
    n = 100_000
    s = R.bsdf_sample(GLASS, (1, 1, 1), N, wo, n, in_medium=False, seed=3)
    n1, n2 = 1.0, 1.5
    ci = math.cos(th)
    st = n1 / n2 * math.sin(th)
    ct = math.sqrt(1 - st * st)
    rs = (n1 * ci - n2 * ct) / (n1 * ci + n2 * ct)
    rp = (n2 * ci - n1 * ct) / (n2 * ci + n1 * ct)
    Rf = 0.5 * (rs * rs + rp * rp)
    refl = ~s["medium"]
    frac = refl.mean()
    sigma = math.sqrt(Rf * (1 - Rf) / n)
    assert abs(frac - Rf) < 4 * sigma + 1e-4, (frac, Rf)
    assert np.isinf(s["pdf"]).all() and (s["type"] == 2).all()
    d = s["dir"].astype(np.float64)
    # reflection: the mirror image of the incoming direction -wo about N
    assert np.allclose(d[refl], [-wo[0], -wo[1], wo[2]], atol=2e-6)
    if (~refl).any():
        t = d[~refl]
        assert (t[:, 2] < 0).all()                        # into the surface
        sin_t = np.hypot(t[:, 0], t[:, 1])
        assert np.allclose(sin_t, st, atol=2e-6)
        assert (t[:, 0] * -wo[0] >= -1e-7).all()         # continues in the incidence plane
        assert (s["offset"][~refl] == -1.0).all()


def test_glass_total_internal_reflection_inside():
    """Inside glass beyond the critical angle (sin > 1/1.5) every sample
    reflects and stays inside; below it some transmit out."""
    N = np.array([0, 0, 1], np.float32)
    crit = math.asin(1 / 1.5)
    for th, expect_tir in ((crit + 0.05, True), (crit - 0.2, False)):
        wo = np.array([math.sin(th), 0.0, math.cos(th)], np.float32)
        s = R.bsdf_sample(GLASS, (1, 1, 1), N, wo, 20_000, in_medium=True, seed=5)
        if expect_tir:
            assert not s["medium"].any()
        else:
            assert s["medium"].mean() > 0.5
            t = s["dir"][s["medium"]].astype(np.float64)
            # leaving the dense medium: sin grows by 1.5
            assert np.allclose(np.hypot(t[:, 0], t[:, 1]), 1.5 * math.sin(th), atol=3e-6)


def test_glass_eval_and_pdf_are_zero():
    v, p, pp = R.bsdf_eval(GLASS, (1, 1, 1), (0, 0, 1), unit((0.3, 0, 1)), unit((0.1, 0.2, 1)))
    assert (v == 0).all() and p == 0 and pp == 0


def _cos_moments(d, N):
    c = d.astype(np.float64) @ np.asarray(N, np.float64)
    return c.mean(), (c * c).mean(), c


@pytest.mark.parametrize("mat", [MATTE, nvisii(0.0, 0.5)])
def test_diffuse_lobe_is_cosine_weighted(mat):
    """E[cos] = 2/3, E[cos^2] = 1/2 for the cosine-weighted hemisphere around
    the (front-facing) normal."""
    N = np.array([0, 0, 1], np.float32)
    wo = unit((0.3, -0.2, 1.0)).astype(np.float32)
    s = R.bsdf_sample(mat, (0.6, 0.5, 0.4), N, wo, 200_000, seed=11)
    sel = s["type"] == 0
    m1, m2, c = _cos_moments(s["dir"][sel], N)
    n = sel.sum()
    assert (c > -1e-6).all()
    assert abs(m1 - 2 / 3) < 5 * math.sqrt(1 / 18 / n)
    assert abs(m2 - 0.5) < 5 * math.sqrt(1 / 12 / n)


def test_lambert_eval_closed_form():
    col = np.array([0.6, 0.5, 0.4])
    N = np.array([0, 0, 1.0])
    wi = unit((0.4, 0.1, 0.8))
    v, p, pp = R.bsdf_eval(MATTE, col, N, unit((0, 0.3, 1)), wi)
    alb = 0.85 * col / math.pi
    np.testing.assert_allclose(v, alb * wi[2] / math.pi, rtol=2e-6)
    assert abs(p - wi[2] / math.pi) < 1e-7 and abs(pp - p) < 1e-7
    # below the surface: pdf() mirrors the normal, eval clamps to 0
    wb = wi * np.array([1, 1, -1.0])
    v, p, pp = R.bsdf_eval(MATTE, col, N, unit((0, 0.3, 1)), wb)
    assert (v == 0).all() and p == 0 and abs(pp - wi[2] / math.pi) < 1e-7


@pytest.mark.parametrize("metallic,roughness", [(0.0, 0.5), (0.0, 0.1), (1.0, 0.3), (0.5, 0.8)])
def test_nvisii_lobe_selection_weights(metallic, roughness):
    """P(diffuse) = dw / (dw + gw), dw = 1 - m, gw = 1 + m + specular(.5) + r."""
    s = R.bsdf_sample(nvisii(metallic, roughness), (0.5, 0.6, 0.7), (0, 0, 1), unit((0.2, 0.1, 1)), 100_000,
                      seed=13)
    m = R.half_round(metallic)
    r = R.half_round(min(max(roughness, 0.001), 0.9))
    dw, gw = 1 - m, 1 + m + 0.5 + r
    want = dw / (dw + gw)
    # glossy samples falling below the hemisphere come back with pdf 0 but
    # keep their lobe's type
    got = (s["type"] == 0).mean()
    assert abs(got - want) < 5 * math.sqrt(max(want * (1 - want), 1e-6) / 100_000) + 1e-5


def test_nvisii_glossy_smooth_is_mirror():
    """At roughness .001 (alpha = .002 floor) glossy samples are the mirror
    direction of wo about N."""
    wo = unit((0.5, -0.3, 0.8)).astype(np.float32)
    s = R.bsdf_sample(nvisii(1.0, 0.001), (0.9, 0.8, 0.7), (0, 0, 1), wo, 20_000, seed=17)
    g = (s["type"] == 1) & (s["pdf"] > 0)
    assert g.mean() > 0.95
    mirror = np.array([-wo[0], -wo[1], wo[2]])
    ang = np.arccos(np.clip(s["dir"][g].astype(np.float64) @ mirror, -1, 1))
    assert np.percentile(ang, 99) < 0.02


def test_nvisii_eval_normal_incidence_closed_form():
    """wi = wo = N: schlick(1) = 0 and G = 1/4, so
    metallic 1: value = base / (4 pi a^2);
    metallic 0: value = base + 0.04 / (4 pi a^2)  (spec = 0.08 * specular)."""
    N = np.array([0, 0, 1.0], np.float32)
    base = np.array([0.5, 0.25, 0.75])
    for m, r in ((1.0, 0.4), (0.0, 0.3), (0.0, 0.05)):
        v, p, pp = R.bsdf_eval(nvisii(m, r), base, N, N, N)
        rh = R.half_round(r)
        alpha = np.float32(max(0.002, np.float32(rh) * np.float32(rh)))
        # D(1) = 1 / (pi (1 + (a^2 - 1))): the fp32 cancellation in the
        # denominator is part of the formula at small roughness
        a2 = float(np.float32(1) + (alpha * alpha - np.float32(1)))
        want = base / (4 * math.pi * a2) if m == 1.0 else base + 0.04 / (4 * math.pi * a2)
        np.testing.assert_allclose(v, want, rtol=2e-5)
        # pdf: dw / pi + gw * D(1) * 1 / 4
        dw, gw = 1 - m, 1 + m + 0.5 + rh
        np.testing.assert_allclose(p, dw / math.pi + gw / (math.pi * a2) / 4, rtol=2e-5)
        assert p == pp


def test_nvisii_reciprocity():
    """With sheen and clear coat off every active lobe is symmetric in
    (wi, wo): brdf(wo, wi) == brdf(wi, wo)."""
    rng = np.random.default_rng(19)
    N = np.array([0, 0, 1.0], np.float32)
    for _ in range(200):
        a = unit(rng.normal(size=3) * [1, 1, 0.3] + [0, 0, 1])
        b = unit(rng.normal(size=3) * [1, 1, 0.3] + [0, 0, 1])
        if a[2] <= 0.05 or b[2] <= 0.05:
            continue
        mat = nvisii(float(rng.uniform()), float(rng.uniform(0.01, 0.9)))
        col = rng.uniform(0.1, 0.9, 3)
        v1, _, _ = R.bsdf_eval(mat, col, N, a, b)
        v2, _, _ = R.bsdf_eval(mat, col, N, b, a)
        np.testing.assert_allclose(v1, v2, rtol=3e-5)


@pytest.mark.parametrize("mat", [MATTE, nvisii(0.0, 0.4), nvisii(0.7, 0.2)])
def test_scatter_is_consistent_with_eval(mat):
    """The (f_r, pdf) a scatter returns equal eval()/pdf() at the returned
    direction (up to the final normalize of the direction)."""
    N = np.array([0, 0, 1], np.float32)
    wo = unit((-0.4, 0.2, 0.9)).astype(np.float32)
    col = (0.7, 0.4, 0.2)
    s = R.bsdf_sample(mat, col, N, wo, 300, seed=23)
    for k in range(300):
        if not s["pdf"][k] > 0:
            continue
        v, p, pp = R.bsdf_eval(mat, col, N, wo, s["dir"][k])
        if mat["type"] == "matte":
            # Lambert scatter returns the albedo; eval returns albedo cos / pi
            np.testing.assert_allclose(s["f"][k] * s["dir"][k][2] / math.pi, v, rtol=1e-5, atol=1e-8)
        else:
            np.testing.assert_allclose(s["f"][k], v, rtol=2e-3, atol=1e-6)
        np.testing.assert_allclose(s["pdf"][k], pp, rtol=2e-3, atol=1e-6)

"""Pins of the oracle's quad, directional and point lights (the samplers
sampleAreaLights / sampleDirLights / samplePointLights and their choice in
sampleLights, barney/kernels/shadeRays.cu:46-404; the lights' getDD in
barney/light/) against what the mathematics fixes. The scenes are a matte
plane of color c seen from above, direct light only (no diffuse bounce, or
an ambient of 0 so that the bounce adds nothing), where barney's shading of
one light sample is
    tp = f_r * L * (1/pi) mu,   f_r = (0.85 c / pi) (1/pi) mu
(AnariMatte albedo, Lambertian::eval, the shadow ray's cosine factor,
shadeRays.cu:680-690), so a light of radiance L at cosine mu gives
0.85 c L mu^2 / pi^3:
  * one directional light: every pixel equals that value;
  * two directional lights with different weights: the 2-candidate
    reservoir and the choice's 1 / share are unbiased (mean = sum);
  * one point light of intensity I at height h: I / d^2 with mu = h / d per
    pixel, and power P gives I = P / (4 pi);
  * shadow-ray length: an occluder beyond a point light casts no shadow, one
    between the light and the plane does; for a directional light any
    occluder does;
  * a square quad light: the 8-candidate reservoir estimates
    Le int mu^2 mu' / d^2 dA (quadrature) without bias, and a quad facing
    away lights nothing;
  * lights next to a uniform ambient environment: the mean is the sum of
    the two lights' means (sampleLights' choice by shares)."""
import math

import numpy as np
import pytest

from oracle import refcpu as R

C = 0.6
K = 0.85 * C / math.pi ** 3


def _plane(y=0.0, S=20.0, c=C):
    v = np.array([[-S, y, -S], [S, y, -S], [S, y, S], [-S, y, S]], np.float32)
    t = np.array([[0, 2, 1], [0, 3, 2]], np.int32)
    return dict(kind=0, vertices=v, indices=t, color=(c, c, c))


def _down_cam(h, fovy=20.0):
    return R.camera((0.0, h, 0.0), (0.0, -1.0, 0.0), up=(0.0, 0.0, 1.0), fovy=fovy)


def _hit_points(cam, p, W, H, y=0.0):
    """Where each pixel's primary ray meets the plane at height y."""
    P = np.empty((H, W, 3))
    for iy in range(H):
        for ix in range(W):
            o, d, _ = R.camera_ray(cam, p, ix, iy)
            t = (y - o[1]) / d[1]
            P[iy, ix] = o.astype(np.float64) + t * d.astype(np.float64)
    return P


def _render(geoms, cam, W, H, spp, lights, ambient=0.0, max_diffuse=0):
    s = R.Scene(geoms)
    p = R.render_params(W, H, spp=spp, ambient=ambient, max_diffuse=max_diffuse, lights=lights)
    acc, _, _ = s.render(cam, p)
    return acc[..., :3] / spp, p


def test_directional_light_closed_form():
    d = np.array([0.3, -1.0, 0.2])
    mu = 1.0 / np.linalg.norm(d)
    Lc, rad = np.array([1.0, 0.5, 0.25]), 2.0
    L = R.Lights(dirs=[[*d, *Lc, rad]])
    img, _ = _render([_plane()], _down_cam(5.0), 16, 12, 2, L)
    exp = K * rad * Lc * mu ** 2
    np.testing.assert_allclose(img.reshape(-1, 3), np.broadcast_to(exp, (16 * 12, 3)), rtol=2e-5)
    # a light from below the plane lights nothing
    L2 = R.Lights(dirs=[[0.0, 1.0, 0.0, 1.0, 1.0, 1.0, 1.0]])
    img2, _ = _render([_plane()], _down_cam(5.0), 8, 8, 2, L2)
    assert np.all(img2 == 0.0)


def test_two_directional_lights_unbiased():
    da, db = np.array([0.0, -1.0, 0.0]), np.array([1.0, -1.0, 0.0])
    ra, rb = 1.0, 3.0
    L = R.Lights(dirs=[[*da, 1, 1, 1, ra], [*db, 1, 1, 1, rb]])
    img, _ = _render([_plane()], _down_cam(5.0), 16, 16, 64, L)
    mua, mub = 1.0, 1.0 / math.sqrt(2.0)
    exp = K * (ra * mua ** 2 + rb * mub ** 2)
    s, H, W = 64, 16, 16
    acc, _ = _render([_plane()], _down_cam(5.0), W, H, 1, L)  # one path per pixel: per-sample spread
    se = acc[..., 0].std() / math.sqrt(H * W * s)
    assert abs(img[..., 0].mean() - exp) < 4 * se + 1e-6, (img[..., 0].mean(), exp, se)


@pytest.mark.parametrize("mode", ["intensity", "power"])
def test_point_light_inverse_square(mode):
    h, W, H = 2.0, 12, 10
    I = 5.0
    pt = [0.3, h, -0.2, 1.0, 0.8, 0.6, 1.0, I] if mode == "intensity" else \
         [0.3, h, -0.2, 1.0, 0.8, 0.6, 4 * math.pi * I, float("nan")]
    L = R.Lights(points=[pt])
    cam = _down_cam(6.0, fovy=40.0)
    img, p = _render([_plane()], cam, W, H, 1, L)
    P = _hit_points(cam, p, W, H)
    dv = np.array(pt[:3]) - P
    d2 = (dv ** 2).sum(-1)
    mu = dv[..., 1] / np.sqrt(d2)
    exp = K * I * mu ** 2 / d2
    np.testing.assert_allclose(img[..., 0], exp, rtol=1e-4)
    np.testing.assert_allclose(img[..., 2], 0.6 * exp, rtol=1e-4)


def test_shadow_ray_length():
    h = 2.0
    pt = R.Lights(points=[[0.0, h, 0.0, 1, 1, 1, 1.0, 4.0]])
    cam = _down_cam(0.5 * h, fovy=30.0)  # between the plane and the occluders
    base, _ = _render([_plane()], cam, 8, 8, 1, pt)
    assert base.min() > 0.0
    beyond, _ = _render([_plane(), _plane(y=2 * h)], cam, 8, 8, 1, pt)
    np.testing.assert_array_equal(beyond, base)  # the shadow ray stops at 0.9999 d
    between, _ = _render([_plane(), _plane(y=0.75 * h)], _down_cam(0.5 * h, fovy=30.0), 8, 8, 1, pt)
    assert np.all(between == 0.0)
    dl = R.Lights(dirs=[[0.0, -1.0, 0.0, 1, 1, 1, 1.0]])
    far, _ = _render([_plane(), _plane(y=1e3)], cam, 8, 8, 1, dl)
    assert np.all(far == 0.0)  # a directional light's shadow ray is infinite


def _square_integral(P, h, a, n=1500):
    """int over the a x a square at height h centred above the origin of
    mu^2 mu' / d^2 = h^3 / d^5 (parallel planes), midpoint rule."""
    g = (np.arange(n) + 0.5) / n * a - a / 2
    X, Z = np.meshgrid(g, g, indexing="ij")
    dx, dz = X - P[0], Z - P[2]
    d2 = dx * dx + h * h + dz * dz
    return (h ** 3 / d2 ** 2.5).sum() * (a / n) ** 2


def test_quad_light_unbiased():
    h, a, Le = 1.5, 1.2, np.array([2.0, 1.0, 0.5])
    # corner, edge0, edge1 with edge0 x edge1 pointing down (towards the plane)
    q = [[-a / 2, h, -a / 2, a, 0, 0, 0, 0, a, *Le]]
    assert np.cross([a, 0, 0], [0, 0, a])[1] < 0
    L = R.Lights(quads=q)
    W = H = 4
    cam = _down_cam(0.8, fovy=60.0)
    spp = 1024
    img, p = _render([_plane()], cam, W, H, spp, L)
    one, _ = _render([_plane()], cam, W, H, 1, L)
    P = _hit_points(cam, p, W, H)
    for iy in range(H):
        for ix in range(W):
            exp = K * Le[0] * _square_integral(P[iy, ix], h, a)
            got = img[iy, ix, 0]
            se = max(one[..., 0].std(), 1e-3 * exp) / math.sqrt(spp)
            assert abs(got - exp) < 5 * se + 2e-3 * exp, (ix, iy, got, exp, se)
    # the same quad facing up lights nothing below it
    qu = [[-a / 2, h, -a / 2, 0, 0, a, a, 0, 0, *Le]]
    dark, _ = _render([_plane()], cam, W, H, 8, R.Lights(quads=qu))
    assert np.all(dark == 0.0)


def test_lights_next_to_ambient_unbiased():
    """Ambient environment plus a directional and a point light: the mean is
    the sum of each alone (each alone is pinned above or by the ambient
    closed forms)."""
    dl = [0.2, -1.0, 0.1, 1, 1, 1, 1.5]
    pl = [0.5, 2.0, 0.0, 1, 1, 1, 3.0, 2.0]
    cam = _down_cam(4.0, fovy=30.0)
    W = H = 8
    spp = 512
    both, _ = _render([_plane()], cam, W, H, spp, R.Lights(dirs=[dl], points=[pl]), ambient=0.7)
    amb, _ = _render([_plane()], cam, W, H, spp, None, ambient=0.7)
    d_only, _ = _render([_plane()], cam, W, H, 1, R.Lights(dirs=[dl]))
    p_only, _ = _render([_plane()], cam, W, H, 1, R.Lights(points=[pl]))
    one, _ = _render([_plane()], cam, W, H, 1, R.Lights(dirs=[dl], points=[pl]), ambient=0.7)
    exp = amb[..., 0].mean() + d_only[..., 0].mean() + p_only[..., 0].mean()
    se = math.hypot(one[..., 0].std() / math.sqrt(W * H * spp), amb[..., 0].std() / math.sqrt(W * H))
    assert abs(both[..., 0].mean() - exp) < 5 * se, (both[..., 0].mean(), exp, se)

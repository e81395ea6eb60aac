"""Pins of the oracle's HDRI environment light (barney/light/EnvMap.cu,
EnvMap.h; radianceFromEnv and the env branch of sampleLights in
barney/kernels/shadeRays.cu) and of its 2D texture sampling (the embree
backend's rtcore/embree/Texture.cpp), against what the mathematics fixes:
  * the frame built from direction / up (closed form for the default
    parameters; toLocal inverts toWorld);
  * bilinear filtering reproduces a field linear in the texel index,
    wrap / clamp / mirror address arithmetic by hand, nearest filtering;
  * the CDFs of a constant map: rows (x+1)/W, the row CDF the running sum of
    sin(theta) (sum_j sin(pi (j + 1/2) / H) = 1 / sin(pi / 2H));
  * the light pdf integrates to 1 over the sphere; samples land in pixels
    with the probabilities the CDFs give (chi-square); E[radiance / pdf] is
    the integral of the map's radiance;
  * a Lambert plane under a constant map and under a map bright only above
    the horizon renders to its closed-form expectation (barney's MIS weights
    with the map's pdf), within 4 standard errors."""
import math

import numpy as np
import pytest

from oracle import refcpu as R


def _sphere_grid(n_theta=720, n_phi=1440):
    th = (np.arange(n_theta) + 0.5) * np.pi / n_theta
    ph = (np.arange(n_phi) + 0.5) * 2 * np.pi / n_phi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1)
    dw = np.sin(T) * (np.pi / n_theta) * (2 * np.pi / n_phi)
    return d.reshape(-1, 3), dw.ravel()


def _random_map(h=16, w=32, seed=0):
    rng = np.random.default_rng(seed)
    t = rng.uniform(0.0, 1.0, (h, w, 4)).astype(np.float32)
    t[..., :3] *= rng.uniform(0.2, 5.0, (h, 1, 1)).astype(np.float32)  # rows of very different power
    t[3:5, 7:9, :3] = 40.0  # a small bright "sun"
    return t


def test_frame_closed_form_and_inverse():
    e = R.EnvMap(np.ones((4, 8, 4), np.float32))  # direction (1,0,0), up (0,0,1)
    tw, tl = e.frame()
    # vz = -up, vy = -normalize(vz x dir), vx = normalize(vy x vz)
    assert np.array_equal(tw, np.array([[-1, 0, 0], [0, 1, 0], [0, 0, -1]], np.float32))
    for direction, up in (((0.3, -0.2, 0.9), (0.1, 1.0, 0.2)), ((1, 2, 3), (-3, 1, 0.5))):
        e = R.EnvMap(np.ones((4, 8, 4), np.float32), direction=direction, up=up)
        tw, tl = e.frame()
        W = tw.T.astype(np.float64)  # columns vx, vy, vz
        L = tl.T.astype(np.float64)
        assert np.allclose(L @ W, np.eye(3), atol=2e-6)
        assert np.allclose(W[:, 2], -np.array(up) / np.linalg.norm(up), atol=1e-6)
        assert abs(np.dot(W[:, 1], direction)) < 1e-5  # vy is normal to the view direction


def test_bilinear_wrap_clamp_mirror_and_nearest():
    h, w = 8, 16
    iy, ix = np.mgrid[0:h, 0:w]
    t = np.zeros((h, w, 4), np.float32)
    t[..., 0] = ix * 2.0 + iy * 0.5 + 1.0  # linear in the texel index
    t[..., 1] = iy
    t[..., 3] = 1.0
    e = R.EnvMap(t, address=("clamp", "clamp"))
    rng = np.random.default_rng(1)
    # interior: texel-centre convention x = u W - 1/2
    u = rng.uniform(1.0 / w, 1 - 1.0 / w, 2000).astype(np.float32)
    v = rng.uniform(1.0 / h, 1 - 1.0 / h, 2000).astype(np.float32)
    out = e.tex2d(np.stack([u, v], -1))
    x, y = u.astype(np.float64) * w - 0.5, v.astype(np.float64) * h - 0.5
    assert np.allclose(out[:, 0], 2 * x + 0.5 * y + 1, atol=1e-4)
    assert np.allclose(out[:, 1], y, atol=1e-5)
    # clamp: beyond the edge texel centres the edge value
    out = e.tex2d(np.array([[-0.3, 0.5], [1.7, 0.5], [0.2 / w, 0.2 / h]], np.float32))
    assert out[0, 0] == t[3, 0, 0] * 0.5 + t[4, 0, 0] * 0.5 and out[1, 0] == t[3, w - 1, 0] * 0.5 + t[4, w - 1, 0] * 0.5
    assert out[2, 0] == t[0, 0, 0]
    # wrap in x: u and u + 1 agree, and across the seam it blends the last
    # and the first column
    e = R.EnvMap(t, address=("wrap", "clamp"))
    a = e.tex2d(np.stack([u, v], -1))
    b = e.tex2d(np.stack([u + 1, v], -1))
    assert np.allclose(a, b, atol=1e-5)
    s = e.tex2d(np.array([[0.0, 0.5]], np.float32))
    assert np.isclose(s[0, 0], 0.5 * (t[3, 0, 0] + t[4, 0, 0]) * 0.5 + 0.5 * (t[3, w - 1, 0] + t[4, w - 1, 0]) * 0.5)
    # mirror as the embree sampler does it: texel indices reflected about
    # texel 0 (|i|, then about the last texel), so u = -0.25 (x = -4.5)
    # reads x = 4.5, and u = 1.25 (x = 19.5) reads x = 2 * 15 - 19.5 + ... =
    # indices 19, 20 -> 12, 11
    e = R.EnvMap(t, address=("mirror", "mirror"))
    m = e.tex2d(np.array([[-0.25, 0.5], [1.25, 0.5]], np.float32))
    assert m[0, 0] == 2 * 4.5 + 0.5 * 3.5 + 1
    assert m[1, 0] == 0.5 * (2 * 12 + 0.5 * 3.5 + 1) + 0.5 * (2 * 11 + 0.5 * 3.5 + 1)
    # nearest: texel floor(|u| W) mod W, floor(|v| H) mod H
    e = R.EnvMap(t, filter="point")
    uv = rng.uniform(-2, 2, (500, 2)).astype(np.float32)
    n = e.tex2d(uv)
    jx = (np.abs(uv[:, 0]) * w).astype(np.uint32) % w
    jy = (np.abs(uv[:, 1]) * h).astype(np.uint32) % h
    assert np.array_equal(n, t[jy, jx])


def test_constant_map_cdfs():
    h, w = 64, 128
    e = R.EnvMap(np.full((h, w, 4), 0.8, np.float32))
    cy, cx = e.cdfs()
    assert np.allclose(cx, (np.arange(w) + 1) / w, atol=2e-6) and (cx[:, -1] == 1).all()
    s = np.sin(np.pi * (np.arange(h) + 0.5) / h)
    assert np.allclose(cy, np.cumsum(s) * math.sin(math.pi / (2 * h)), atol=4e-6) and cy[-1] == 1


def test_pdf_integrates_to_one_and_samples_follow_it():
    t = _random_map()
    h, w = t.shape[:2]
    e = R.EnvMap(t, direction=(0.3, 0.4, 0.5), up=(0, 1, 0.2))
    d, dw = _sphere_grid()
    pdf, rad = e.eval(d)
    assert abs((pdf * dw).sum() - 1) < 2e-3
    # samples: which pixel, and E[rad / pdf] = integral of the radiance
    n = 400_000
    sd, srad, spdf = e.sample(n, seed=7)
    assert np.allclose(np.linalg.norm(sd, axis=1), 1, atol=1e-5)
    cy, cx = e.cdfs()
    py = np.diff(np.concatenate([[0.0], cy.astype(np.float64)]))
    px = np.diff(np.concatenate([np.zeros((h, 1)), cx.astype(np.float64)], 1), axis=1)
    prob = py[:, None] * px
    tw, tl = e.frame()
    loc = sd.astype(np.float64) @ tl.astype(np.float64)  # rows of tl are columns of toLocal
    th = np.arccos(np.clip(loc[:, 2], -1, 1))
    ph = np.mod(np.arctan2(loc[:, 1], loc[:, 0]), 2 * np.pi)
    jx = np.clip((ph / (2 * np.pi) * w).astype(int), 0, w - 1)
    jy = np.clip((th / np.pi * h).astype(int), 0, h - 1)
    cnt = np.bincount(jy * w + jx, minlength=h * w).reshape(h, w)
    expd = prob * n
    m = expd > 20
    chi2 = ((cnt[m] - expd[m]) ** 2 / expd[m]).sum()
    dof = int(m.sum()) - 1
    assert chi2 < dof + 6 * math.sqrt(2 * dof), (chi2, dof)
    assert cnt[~m].sum() < 0.01 * n
    for k in range(3):
        est = srad[:, k].astype(np.float64) / spdf
        integral = (rad[:, k].astype(np.float64) * dw).sum()
        assert abs(est.mean() - integral) < 4 * est.std() / math.sqrt(n) + 2e-3 * integral, (k, est.mean(), integral)


def _plane_scene(c):
    S = 1e4
    v = np.array([[-S, 0, -S], [S, 0, -S], [S, 0, S], [-S, 0, S]], np.float32)
    t = np.array([[0, 2, 1], [0, 3, 2]], np.int32)
    return R.Scene([dict(kind=0, vertices=v, indices=t, color=(c, c, c))])


@pytest.mark.parametrize("upper_only", [False, True])
def test_lambert_plane_under_env_map_closed_form(upper_only):
    """Plane y = 0 facing +y seen from above, one diffuse bounce, map of
    radiance L (everywhere, or only above the horizon with nearest
    filtering). With alb = 0.85 c / pi and pl the map's pdf in the lit
    directions (constant: H / (2 pi^2 sum_j sin theta_j) over the rows that
    carry weight), barney's estimator has the expectation
      NEE:     2 pi int_0^1 alb L mu^2 / pi^2 * pl / (pl + mu / pi) dmu
      bounce:  (1 - q) int_0^1 2 mu * pi alb L * (mu / pi) / (mu / pi + pl) dmu
               + q pi alb L
    where q is the chance that the light sample lands in a row that carries
    importance but no radiance (the dark row next to the lit half): that
    sample has no weight, sampleLights picks no light, and the bounced ray
    keeps MIS weight 1 (shadeRays.cu:866-877)."""
    c, L = 0.6, 0.8
    h, w = 64, 128
    t = np.full((h, w, 4), L, np.float32)
    if upper_only:
        t[: h // 2] = 0.0  # theta < pi/2 in the map's frame = below the horizon (vz = -up)
    e = R.EnvMap(t, direction=(1, 0, 0), up=(0, 1, 0), filter="point" if upper_only else "linear",
                 address=("wrap", "clamp"))
    s = _plane_scene(c)
    W, H, spp = 40, 30, 32
    acc, _, _ = s.render(R.camera((0, 10, 0), (0.01, -1, 0.02), fovy=30),
                         R.render_params(W, H, spp=spp, ambient=123.0, env=e))
    pix = acc[..., 0].ravel() / spp
    theta = np.pi * (np.arange(h) + 0.5) / h
    # rows with weight: the 3x3 half-texel stencil of row iy reaches row
    # iy + 1 (nearest filtering), so the row above the lit half counts too
    rows = np.arange(h) >= (h // 2 - 1 if upper_only else 0)
    pl = h / (2 * np.pi ** 2 * np.sin(theta[rows]).sum())
    q = np.sin(theta[h // 2 - 1]) / np.sin(theta[rows]).sum() if upper_only else 0.0
    alb = 0.85 * c / math.pi
    mu = np.linspace(0, 1, 200001)
    nee = 2 * np.pi * np.trapezoid(alb * L * mu ** 2 / np.pi ** 2 * pl / (pl + mu / np.pi), mu)
    bnc = (1 - q) * np.trapezoid(2 * mu * np.pi * alb * L * (mu / np.pi) / (mu / np.pi + pl), mu) + q * np.pi * alb * L
    exp = nee + bnc
    se = pix.std() / math.sqrt(len(pix))
    print(f"mean {pix.mean():.5f} expected {exp:.5f} se {se:.5f}")
    assert abs(pix.mean() - exp) < 4 * se + 1e-4, (pix.mean(), exp, se)


def test_primary_miss_shows_the_map():
    """A camera ray that hits nothing shows radianceFromEnv of its direction
    instead of the background color."""
    t = _random_map(8, 16, seed=3)
    e = R.EnvMap(t, direction=(0, 0, 1), up=(0, 1, 0), scale=2.0)
    s = _plane_scene(0.5)
    cam = R.camera((0, 10, 0), (0, 1, 0.3), fovy=60)  # looking up: every ray misses
    W, H = 16, 12
    acc, _, _ = s.render(cam, R.render_params(W, H, spp=1, env=e, bg=(0.1, 0.2, 0.3, 1.0)))
    dirs = np.array([R.camera_ray(cam, R.render_params(W, H), x, y)[1] for y in range(H) for x in range(W)])
    _, rad = e.eval(dirs)
    assert np.array_equal(acc[..., :3].reshape(-1, 3), rad)

"""Pins of the oracle's volume-scatter shading (the SCI_VIS bounce at a
delta-tracking collision) against its closed-form expectation.

A homogeneous medium (constant field, constant transfer function: colour c,
extinction sigma) fills a box that is exactly one macrocell, under uniform
ambient light `a`, with background colour b. A primary ray that crosses the
box over [0, L] collides at the first free-flight distance
dt = -fastLog(u) / sigma (Volume.h:174-223; u uniform in (0, 1]) if dt < L,
and otherwise shows b. At a collision P, bounce() (shadeRays.cu:476-878)
with the Phase BSDF (albedo 0.7 c, value and pdf 1/(4 pi), Phase.h:137-172)
draws a uniform sphere direction w for the environment light (radiance
4 pi a, pdf 1/(4 pi)), weights it by MIS pl / (pl + pb) = 1/2, and traces a
shadow ray that escapes with probability T(d) = P(-fastLog(u)/sigma >= d),
d the distance from P to the box along w; the scattered ray is not traced
(volumes end the path). So

  E[pixel] = P(dt >= L) b + E[ 1{dt < L} 0.5 * 0.7 c * 4 pi a * E_w T(d(P, w)) ]

computed here in float64 for every path's own jittered camera ray, from the
fastLog formula (not from the oracle). Albedo 1 instead of 0.7, a
hemisphere pdf 1/(2 pi) instead of 1/(4 pi), an unshadowed light (T = 1) or
exact-log free flights each move the image by many standard errors."""
import ctypes as C
import math

import numpy as np
from scipy.ndimage import map_coordinates

from oracle import refcpu

SIGMA, AMB = 0.15, 0.9
COLOR = np.array([0.8, 0.5, 0.3])
BG = (0.2, 0.3, 0.4, 0.5)
N = 17            # voxels per axis: 16 cells = one 16-cell macrocell
ORIGIN = np.array([1.0, 2.0, 3.0])
SPACING = 0.5
LO, HI = ORIGIN, ORIGIN + SPACING * (N - 1)


def fast_log(f):
    """Volume.h:174-184 in float64: 2 (z + z^3/3 + z^5/5), z = (f-1)/(f+1)."""
    z = (f - 1.0) / (f + 1.0)
    return 2.0 * (z + z ** 3 / 3.0 + z ** 5 / 5.0)


def _free_flights(m=1 << 20):
    u = (np.arange(m) + 0.5) / m
    return np.sort(-fast_log(u) / SIGMA)


def _exit_distance(P, w):
    """Distance from points P (inside the box) along unit directions w."""
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.where(w > 0, (HI - P) / w, np.where(w < 0, (LO - P) / w, np.inf))
    return t.min(-1)


def _light_visibility_grid(flights, k=2048, g=17):
    """E_w T(d(P, w)) on a g^3 lattice over the box (Fibonacci directions)."""
    i = np.arange(k) + 0.5
    z = 1 - 2 * i / k
    phi = math.pi * (1 + 5 ** 0.5) * i
    r = np.sqrt(1 - z * z)
    w = np.stack([r * np.cos(phi), r * np.sin(phi), z], -1)
    ax = np.linspace(0.0, 1.0, g)
    P = LO + (HI - LO) * np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 1, 3)
    d = _exit_distance(P, w[None])
    T = 1.0 - np.searchsorted(flights, d) / len(flights)
    return T.mean(-1).reshape(g, g, g)


def _expected_path(org, d, flights, vis, bins=96):
    """Expected RGB and alpha of the paths with camera rays org + t d."""
    with np.errstate(divide="ignore", invalid="ignore"):
        a = (LO - org) / d
        b = (HI - org) / d
    t0 = np.maximum(np.minimum(a, b).max(-1), 0.0)
    t1 = np.maximum(a, b).min(-1)
    L = np.where(t1 > t0, t1 - t0, 0.0)
    n = len(flights)
    p_coll = np.searchsorted(flights, L) / n
    # collision positions: bins of the free-flight distribution below L
    q = (np.arange(bins) + 0.5) / bins
    idx = np.minimum((q[None, :] * (p_coll[:, None] * n)).astype(np.int64), n - 1)
    t = flights[idx]                       # quantiles of dt given dt < L
    P = org[:, None, :] + (t0[:, None] + t)[..., None] * d[:, None, :]
    g = vis.shape[0]
    coords = ((P - LO) / (HI - LO) * (g - 1)).reshape(-1, 3).T
    v = map_coordinates(vis, coords, order=1, mode="nearest").reshape(P.shape[:2]).mean(1)
    light = 0.5 * 0.7 * 4 * math.pi * AMB * v * p_coll
    rgb = (1 - p_coll)[:, None] * np.array(BG[:3]) + light[:, None] * COLOR[None, :]
    alpha = p_coll + (1 - p_coll) * BG[3]
    return rgb, alpha


def test_homogeneous_medium_scatter_closed_form():
    field = np.full((N, N, N), 0.5, np.float32)
    vol = refcpu.Volume(field, origin=tuple(ORIGIN), spacing=(SPACING,) * 3, cells_per_mc=N - 1)
    assert vol.mc_dims() == (1, 1, 1)
    tf = np.array([[*COLOR, 0.25], [*COLOR, 0.25]], np.float32)
    vol.set_tf(tf, (0.0, 1.0), SIGMA / 0.25)
    W, H, spp, frames = 16, 12, 64, 12
    cam = refcpu.camera((5.0, 6.5, -9.0), (0.0, -0.2, 1.0), fovy=55)
    flights = _free_flights()
    vis = _light_visibility_grid(flights)
    Lib = refcpu.load()
    got = np.zeros((frames, H, W, 4))
    want = np.zeros((H, W, 4))
    org, dr = np.zeros(3, np.float32), np.zeros(3, np.float32)
    seed = C.c_uint64(0)
    for f in range(frames):
        p = refcpu.render_params(W, H, spp=spp, accum_id=f * spp, ambient=AMB, bg=BG)
        acc, _, _ = vol.render(cam, p)
        got[f] = acc / spp
        # the same paths' camera rays (generateRays: jitter from accumID)
        rays_o, rays_d = np.zeros((H * W * spp, 3)), np.zeros((H * W * spp, 3))
        k = 0
        for a in range(f * spp, (f + 1) * spp):
            pa = refcpu.render_params(W, H, accum_id=a)
            for iy in range(H):
                for ix in range(W):
                    Lib.ref_camera_ray(C.byref(cam), C.byref(pa), ix, iy, org.ctypes.data, dr.ctypes.data,
                                       C.byref(seed))
                    rays_o[k], rays_d[k] = org, dr
                    k += 1
        rgb, alpha = _expected_path(rays_o, rays_d, flights, vis)
        want[..., :3] += rgb.reshape(spp, H, W, 3).mean(0) / frames
        want[..., 3] += alpha.reshape(spp, H, W).mean(0) / frames
    mean = got.mean(0)
    # floor: float32 accumulation of spp background values in a pixel that
    # misses the box (zero variance there)
    se = np.sqrt((got.std(0, ddof=1) / math.sqrt(frames)) ** 2 + 1e-10)
    z = (mean - want) / se
    zall = (mean - want).sum((0, 1)) / np.sqrt((se ** 2).sum((0, 1)))
    print(f"image mean {mean[..., :3].mean():.4f} expected {want[..., :3].mean():.4f}; "
          f"whole-image z {np.round(zall, 2)}")
    # per pixel where collisions are not rare (a grazed pixel's sample
    # variance is unreliable); the whole-image z covers every pixel
    inside = (want[..., 3] - BG[3]) / (1 - BG[3]) >= 0.05
    assert inside.mean() > 0.3
    assert np.abs(z[inside]).max() <= 4.5
    assert np.all(np.abs(zall) <= 3.5)
    # the medium matters: pixels through the box differ from the background
    assert np.abs(want[..., :3] - np.array(BG[:3])).max() > 0.5

"""Pins of the oracle's triangle anyHit features and the renderer's cut plane
(barney/geometry/Triangles.dev.cu:19-136, Spheres.dev.cu:160-174,
Cylinders.dev.cu:257-259, render/OptixGlobals.h:33-42,
kernels/generateRays.cu:171-193), against what the mathematics fixes:
  * shading normals: the normal at a hit is the normalized barycentric
    interpolation of the vertex normals (float64 from the hit point's own
    barycentrics), the geometric normal without them;
  * cut plane: the closest hit equals a float64 brute force over analytic
    sphere roots in which a sphere whose nearest root lies on the invisible
    side is skipped (its program returns without trying the far root);
    primary rays start on the plane when the camera is on the invisible
    side;
  * stochastic opacity: the accepted hit does not depend on traversal order
    (BVH equals brute force bit for bit), alpha 1 always hits and alpha 0
    never, the hit rate of alpha a is a (binomial bound), shadow rays see
    the same draws, and the ID pass (opacity ignored) sees every candidate;
  * a plane of alpha a over a plane of another color renders to the
    a-weighted mix of the two single-plane renders (unbiased)."""
import math

import numpy as np
import pytest

from oracle import refcpu as R


def _quad(y, S=50.0):
    v = np.array([[-S, y, -S], [S, y, -S], [S, y, S], [-S, y, S]], np.float32)
    return v, np.array([[0, 2, 1], [0, 3, 2]], np.int32)


def _rays_down(n, seed=0, y=10.0, spread=20.0):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), np.float32)
    r[:, 0] = rng.uniform(-spread, spread, n)
    r[:, 1] = y
    r[:, 2] = rng.uniform(-spread, spread, n)
    d = np.stack([rng.uniform(-0.3, 0.3, n), -np.ones(n), rng.uniform(-0.3, 0.3, n)], 1)
    r[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r[:, 7] = 1e30
    return r


def test_vertex_normals_interpolate():
    rng = np.random.default_rng(1)
    v = rng.uniform(-5, 5, (60, 3)).astype(np.float32)
    v[:, 1] = rng.uniform(-0.5, 0.5, 60)
    t = rng.integers(0, 60, (40, 3)).astype(np.int32)
    t = t[(t[:, 0] != t[:, 1]) & (t[:, 1] != t[:, 2]) & (t[:, 0] != t[:, 2])]
    nrm = rng.normal(size=(60, 3)).astype(np.float32)
    nrm[:, 1] = np.abs(nrm[:, 1]) + 0.5
    s = R.Scene([dict(kind=0, vertices=v, indices=t, normals=nrm)])
    s0 = R.Scene([dict(kind=0, vertices=v, indices=t)])
    rays = _rays_down(4000, spread=5.0)
    h = s.trace(rays)
    h0 = s0.trace(rays)
    hit = h["primID"] >= 0
    assert hit.sum() > 500
    assert np.array_equal(h["t"], h0["t"]) and np.array_equal(h["primID"], h0["primID"])
    for k in np.nonzero(hit)[0]:
        tri = t[h["primID"][k]].astype(np.int64)
        a, b, c = v[tri].astype(np.float64)
        P = rays[k, :3].astype(np.float64) + float(h["t"][k]) * rays[k, 4:7].astype(np.float64)
        # barycentrics of P by least squares in the triangle's plane
        M = np.stack([b - a, c - a], 1)
        uv = np.linalg.lstsq(M, P - a, rcond=None)[0]
        n = (1 - uv.sum()) * nrm[tri[0]] + uv[0] * nrm[tri[1]] + uv[1] * nrm[tri[2]]
        n /= np.linalg.norm(n)
        assert np.allclose(h["N"][k], n, atol=2e-4), (k, h["N"][k], n)
        g = np.cross(b - a, c - a)
        g /= np.linalg.norm(g)
        assert np.allclose(h0["N"][k], g, atol=1e-5)


def _sphere_scene(seed=2, n=30):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-6, 6, (n, 3)).astype(np.float32)
    r = rng.uniform(0.5, 1.5, n).astype(np.float32)
    return c, r, R.Scene([dict(kind=1, vertices=c, radii=r)])


def _brute_spheres_cut(rays, c, r, plane):
    """float64: per ray the smallest 'nearest root' over spheres whose
    nearest root in (0, tmax) lies on the visible side."""
    n = np.array(plane[:3], np.float64)
    out = np.full(len(rays), np.inf)
    idx = np.full(len(rays), -1)
    for k, ray in enumerate(rays.astype(np.float64)):
        o, d = ray[:3], ray[4:7]
        best, bi = np.inf, -1
        for j in range(len(c)):
            oc = o - c[j]
            a, b, cc = d @ d, oc @ d, oc @ oc - float(r[j]) ** 2
            disc = b * b - a * cc
            if disc < 0:
                continue
            roots = [x for x in ((-b - math.sqrt(disc)) / a, (-b + math.sqrt(disc)) / a) if 0 < x < ray[7]]
            if not roots:
                continue
            t = min(roots)
            if plane[3] > -1e28 and n @ (o + t * d) + plane[3] < 0:
                continue
            if t < best:
                best, bi = t, j
        out[k], idx[k] = best, bi
    return out, idx


@pytest.mark.parametrize("plane", [(0.0, 0.0, 1.0, -0.5), (0.3, -1.0, 0.2, 1.0), (0.0, 0.0, 0.0, -1e30)])
def test_cut_plane_against_brute_force(plane):
    c, r, s = _sphere_scene()
    rng = np.random.default_rng(3)
    rays = np.zeros((1500, 8), np.float32)
    rays[:, :3] = rng.uniform(-12, 12, (1500, 3))
    d = rng.uniform(-6, 6, (1500, 3)) - rays[:, :3]
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 7] = 1e30
    p = R.render_params(4, 4, cut_plane=plane)
    h = s.trace_filtered(rays, p, np.zeros(len(rays), np.uint64))
    t, j = _brute_spheres_cut(rays, c, r, plane)
    assert np.array_equal(h["geomID"] >= 0, j >= 0) and (j >= 0).sum() > 200
    ok = j >= 0
    assert np.array_equal(h["primID"][ok], j[ok])
    assert np.allclose(h["t"][ok], t[ok], rtol=1e-5)


def test_cut_plane_moves_primary_rays_onto_the_plane():
    cam = R.camera((0, 10, 0), (0, -1, 0.0001), fovy=30)
    # camera above y = 4 looking down, visible side y < 4: rays start at y = 4
    p = R.render_params(8, 8, cut_plane=(0.0, -1.0, 0.0, 4.0))
    o, d, _ = R.camera_ray(cam, p, 3, 5)
    assert abs(o[1] - 4.0) < 1e-5 and d[1] < 0
    # visible side y > 4 and looking down: clipped where it leaves (t = 6)
    p = R.render_params(8, 8, cut_plane=(0.0, 1.0, 0.0, -4.0))
    s = R.Scene([dict(kind=0, vertices=_quad(0.0)[0], indices=_quad(0.0)[1])])
    acc, ids, _ = s.render(cam, p, with_ids=True)
    assert (ids[..., 0] == -1).all()  # the ground plane lies beyond the cut
    # looking away from the visible side from the invisible side: nothing
    p = R.render_params(8, 8, cut_plane=(0.0, 1.0, 0.0, -14.0), bg=(0.25, 0.5, 0.75, 1.0))
    acc, ids, _ = s.render(cam, p, with_ids=True)
    assert (ids[..., 0] == -1).all() and np.allclose(acc[..., :3], (0.25, 0.5, 0.75))


def _cutout_scene(alpha, mat="matte"):
    top, ti = _quad(2.0)
    bot, bi = _quad(0.0)
    m = dict(type="matte", color=(0.9, 0.2, 0.2), opacity=alpha) if mat == "matte" else \
        dict(type="pbr", color=(0.9, 0.2, 0.2), metallic=0.0, roughness=0.5, transmission=0.0, ior=1.0,
             opacity=alpha)
    return R.Scene([dict(kind=0, vertices=top, indices=ti, material=m),
                    dict(kind=0, vertices=bot, indices=bi, material=dict(type="matte", color=(0.2, 0.2, 0.9)))])


@pytest.mark.parametrize("alpha", [0.0, 0.3, 0.77, 1.0])
def test_stochastic_opacity_hit_rate_and_order_independence(alpha):
    s = _cutout_scene(alpha)
    n = 20000
    rays = _rays_down(n, seed=4)
    seeds = np.random.default_rng(5).integers(0, 2 ** 63, n, dtype=np.uint64)
    p = R.render_params(4, 4)
    h = s.trace_filtered(rays, p, seeds)
    hb = s.trace_filtered(rays, p, seeds, brute=True)
    assert np.array_equal(h["geomID"], hb["geomID"]) and np.array_equal(h["t"].view(np.uint32), hb["t"].view(np.uint32))
    assert (h["geomID"] >= 0).all()
    rate = (h["geomID"] == 0).mean()
    if alpha in (0.0, 1.0):
        assert rate == alpha
    else:
        assert abs(rate - alpha) < 5 * math.sqrt(alpha * (1 - alpha) / n), (rate, alpha)
    # occlusion draws the same way; the ID pass ignores opacity
    sh = s.trace_filtered(rays, p, seeds, shadow=True)
    assert np.array_equal(sh["geomID"] >= 0, h["geomID"] >= 0)
    hid = s.trace_filtered(rays, p, seeds, stochastic=False)
    assert (hid["geomID"] == 0).all()
    # other seeds, other decisions
    if 0 < alpha < 1:
        h2 = s.trace_filtered(rays, p, seeds + np.uint64(1))
        assert (h2["geomID"] != h["geomID"]).mean() > 0.2


def test_pbr_alpha_is_a_half_of_the_product():
    # NVisii keeps alpha as a half: 0.1 * (1 - 0.3) rounds to 0.06997680664
    s = _cutout_scene(0.1 * 1.0, mat="pbr")
    n = 40000
    rays = _rays_down(n, seed=6)
    seeds = np.random.default_rng(7).integers(0, 2 ** 63, n, dtype=np.uint64)
    h = s.trace_filtered(rays, R.render_params(4, 4), seeds)
    a = R.half_round(0.1)
    rate = (h["geomID"] == 0).mean()
    assert abs(rate - a) < 5 * math.sqrt(a * (1 - a) / n)


def test_cutout_render_is_the_alpha_mix():
    """Camera straight down at a plane of alpha a (red, y = 2) over an
    opaque blue plane (y = 0), light from the environment, NEE only
    (no diffuse bounce). The first hit is red with probability a; a blue
    hit's shadow ray passes the red plane with probability 1 - a (its own,
    independent draw). So E[pixel] = a E[red plane alone] +
    (1 - a)^2 E[blue plane alone]."""
    a = 0.35
    cam = R.camera((0, 10, 0), (0.01, -1, 0.02), fovy=20)
    W, H, spp = 24, 18, 64
    p = R.render_params(W, H, spp=spp, ambient=0.8, max_diffuse=0)

    def mean(scene):
        acc, _, _ = scene.render(cam, p)
        x = acc[..., :3].reshape(-1, 3) / spp
        return x.mean(0), x.std(0) / math.sqrt(len(x))

    got, se = mean(_cutout_scene(a))
    top, ti = _quad(2.0)
    red, _ = mean(R.Scene([dict(kind=0, vertices=top, indices=ti,
                                material=dict(type="matte", color=(0.9, 0.2, 0.2)))]))
    bot, bi = _quad(0.0)
    blue, _ = mean(R.Scene([dict(kind=0, vertices=bot, indices=bi,
                                 material=dict(type="matte", color=(0.2, 0.2, 0.9)))]))
    want = a * red + (1 - a) ** 2 * blue
    print(got, want, se)
    assert np.all(np.abs(got - want) < 5 * se + 1e-4), (got, want, se)

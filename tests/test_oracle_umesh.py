"""Pins of the oracle's unstructured-mesh functions against what the
mathematics fixes: the tet generator is conforming and fills the jittered
hull; point location agrees with float64 barycentric brute force;
barycentric interpolation reproduces linear fields; macrocell ranges
bound every sampled value; the iso-surface of a linear field is the exact
plane with the exact gradient normal; delta-tracked opacity through a
constant-density slab is 1 - exp(-sigma L)."""
import math

import numpy as np
import pytest

from barney_amd import scenes
from oracle import refcpu as R


@pytest.fixture(scope="module")
def mesh():
    V, S, T = scenes.tet_grid(6, seed=11)
    return V, S, T, R.UMesh(V, S, T)


def test_tet_grid_is_conforming_and_positive():
    V, S, T = scenes.tet_grid(5, seed=3)
    a, b, c, d = (V[T[:, q]].astype(np.float64) for q in range(4))
    det = np.einsum("ij,ij->i", np.cross(b - a, c - a), d - a)
    assert (det > 0).all()
    faces = np.sort(np.concatenate([T[:, [0, 1, 2]], T[:, [0, 1, 3]], T[:, [0, 2, 3]], T[:, [1, 2, 3]]]), 1)
    _, cnt = np.unique(faces, axis=0, return_counts=True)
    # interior faces shared by two tets, the 6 * 5^2 * 2 boundary triangles by one
    assert set(np.unique(cnt)) == {1, 2} and (cnt == 1).sum() == 6 * 25 * 2
    assert S.min() == 0.0 and S.max() == 1.0


def _brute(V, S, T, p):
    a, b, c, d = (V[T[:, q]].astype(np.float64) for q in range(4))
    M = np.stack([b - a, c - a, d - a], -1)
    lam = np.linalg.solve(M, (p - a)[..., None])[..., 0]
    L = np.concatenate([(1 - lam.sum(1))[:, None], lam], 1)
    ok = (L >= -1e-9).all(1)
    if not ok.any():
        return np.nan
    k = np.nonzero(ok)[0][0]
    return float((L[k] * S[T[k]]).sum())


def test_point_location_matches_brute_force(mesh):
    V, S, T, u = mesh
    rng = np.random.default_rng(1)
    P = rng.uniform(-0.6, 6.6, (1500, 3)).astype(np.float32)
    got = u.sample(P)
    for i in range(len(P)):
        want = _brute(V, S, T, P[i].astype(np.float64))
        if math.isnan(want):
            assert math.isnan(got[i]), i
        else:
            assert abs(got[i] - want) <= 2e-6, (i, got[i], want)


def test_linear_field_is_reproduced():
    V, _, T = scenes.tet_grid(4, seed=21)
    g = np.array([0.3, -0.2, 0.1])
    S = (V.astype(np.float64) @ g + 0.5).astype(np.float32)
    u = R.UMesh(V, S, T)
    rng = np.random.default_rng(2)
    P = rng.uniform(0.5, 3.5, (3000, 3)).astype(np.float32)
    got = u.sample(P)
    assert not np.isnan(got).any()
    np.testing.assert_allclose(got, P.astype(np.float64) @ g + 0.5, atol=2e-6)


def test_macrocell_ranges_bound_samples(mesh):
    V, S, T, u = mesh
    ranges, _ = u.macrocells()
    mx, my, mz = u.mc_dims()
    lo = V[T.reshape(-1)].min(0)
    hi = V[T.reshape(-1)].max(0)
    size = hi - lo
    sp = size * (1.0 / np.array([mx, my, mz], np.float32))
    rng = np.random.default_rng(3)
    P = rng.uniform(lo, hi, (20000, 3)).astype(np.float32)
    vals = u.sample(P)
    m = ~np.isnan(vals)
    c = np.clip(((P - lo) / sp).astype(np.int64), 0, np.array([mx, my, mz]) - 1)
    r = ranges[c[:, 2], c[:, 1], c[:, 0]]
    assert (r[m, 0] <= vals[m] + 1e-6).all() and (vals[m] <= r[m, 1] + 1e-6).all()
    # untouched cells stay empty (+inf, -inf) -> majorant 0
    u.set_tf(np.array([[1, 1, 1, 0.2], [1, 1, 1, 0.8]], np.float32), (0.0, 1.0), 2.0)
    _, maj = u.macrocells()
    empty = ranges[..., 0] > ranges[..., 1]
    assert (maj[empty] == 0).all() and (maj[~empty] > 0).all()


def test_iso_surface_of_linear_field_is_the_plane():
    """s = x / n: the iso-surface s = 0.5 is the plane x = n/2; the linear
    root inside a sub-step is exact for a linear field, so the hit is
    t = (n/2 - ox) / dx, and the central-difference gradient is +x."""
    n = 6
    V, _, T = scenes.tet_grid(n, seed=31, jitter=0.15)
    S = (V[:, 0].astype(np.float64) / n).astype(np.float32)
    u = R.UMesh(V, S, T)
    rng = np.random.default_rng(4)
    m = 400
    org = np.stack([np.full(m, -5.0), rng.uniform(1.5, 4.5, m), rng.uniform(1.5, 4.5, m)], 1)
    d = np.stack([np.ones(m), rng.uniform(-0.1, 0.1, m), rng.uniform(-0.1, 0.1, m)], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([org, np.zeros((m, 1)), d, np.full((m, 1), 1e30)], 1).astype(np.float32)
    out = u.iso_rays(rays, 0.5)
    want = (n / 2 - org[:, 0]) / rays[:, 4]
    assert np.isfinite(out[:, 0]).all()
    np.testing.assert_allclose(out[:, 0], want, rtol=2e-5)
    np.testing.assert_allclose(out[:, 1:], np.tile([1.0, 0.0, 0.0], (m, 1)), atol=2e-3)
    # iso values outside the field's range are never hit
    assert np.isinf(u.iso_rays(rays, 1.5)[:, 0]).all()


def test_delta_tracking_opacity_of_constant_slab():
    """Constant field, TF density w: the fraction of primary rays that
    scatter equals 1 - exp(-w * base * L) over the path length L through
    the (convex) box of the mesh."""
    n = 4
    V, _, T = scenes.tet_grid(n, seed=41, jitter=0.0)
    S = np.full(len(V), 0.5, np.float32)
    u = R.UMesh(V, S, T)
    w, base = 0.25, 1.0
    u.set_tf(np.array([[1, 1, 1, w], [1, 1, 1, w]], np.float32), (0.0, 1.0), base)
    # rays along +z through the 4 x 4 x 4 cube: L = 4
    W = H = 8
    spp = 256
    cam = R.camera((2.0, 2.0, -30.0), (0.0, 0.0, 1.0), fovy=2.0)
    p = R.render_params(W, H, spp=spp, ambient=1.0, bg=(0, 0, 0, 0), max_depth=1, max_diffuse=0)
    acc, _, _ = u.render(cam, p, volume=True)
    frac = acc[..., 3].mean() / spp
    want = 1 - math.exp(-w * base * n)
    sigma = math.sqrt(want * (1 - want) / (W * H * spp))
    assert abs(frac - want) < 4 * sigma, (frac, want)

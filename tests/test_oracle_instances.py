"""Pins of the oracle's instanced worlds (two-level tracing): identity
instancing reproduces the plain scene bit for bit; a translated instance
equals the translated geometry; rotated + scaled instances equal a float64
analytic intersection of the transformed primitives; duplicate instances
resolve ties to the smaller instance index; normals follow the rotation."""
import math

import numpy as np

from oracle import refcpu as R

ID = (1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0)


def _rays(n, seed, org_lo=-6, org_hi=6):
    rng = np.random.default_rng(seed)
    o = rng.uniform(org_lo, org_hi, (n, 3))
    tgt = rng.uniform(-1.5, 1.5, (n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, np.zeros((n, 1)), d, np.full((n, 1), 1e30)], 1).astype(np.float32)


def _tri_scene():
    rng = np.random.default_rng(3)
    v = rng.uniform(-1, 1, (60, 3)).astype(np.float32)
    idx = rng.integers(0, 60, (40, 3)).astype(np.int32)
    return v, idx


def test_identity_instance_is_the_scene():
    v, idx = _tri_scene()
    sc = R.Scene([dict(kind=0, vertices=v, indices=idx)])
    w = R.World([sc], [(0, 5, ID)])
    rays = _rays(4000, 1)
    a, b = sc.trace(rays), w.trace(rays)
    assert np.array_equal(a["primID"], b["primID"])
    assert np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32))
    assert np.array_equal(a["N"].view(np.uint32), b["N"].view(np.uint32))
    hit = b["primID"] >= 0
    assert hit.mean() > 0.2 and (b["instID"][hit] == 5).all() and (b["instIndex"][hit] == 0).all()


def test_translated_instance_equals_translated_geometry():
    v, idx = _tri_scene()
    T = np.array([0.75, -0.5, 0.25], np.float32)
    sc = R.Scene([dict(kind=0, vertices=v, indices=idx)])
    moved = R.Scene([dict(kind=0, vertices=v + T, indices=idx)])
    w = R.World([sc], [(0, 0, (1, 0, 0, 0, 1, 0, 0, 0, 1, *T))])
    rays = _rays(4000, 2)
    a, b = moved.trace(rays), w.trace(rays)
    same = a["primID"] == b["primID"]
    assert same.mean() > 0.999
    m = same & (a["primID"] >= 0)
    np.testing.assert_allclose(b["t"][m], a["t"][m], rtol=1e-5)
    np.testing.assert_allclose(b["N"][m], a["N"][m], atol=1e-5)


def _rot(axis, ang):
    axis = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K


def _xfm12(L, p):
    return tuple(np.concatenate([L[:, 0], L[:, 1], L[:, 2], p]).astype(np.float32))


def test_rotated_scaled_spheres_match_analytic():
    rng = np.random.default_rng(4)
    c = rng.uniform(-1, 1, (30, 3)).astype(np.float32)
    r = rng.uniform(0.05, 0.2, 30).astype(np.float32)
    sc = R.Scene([dict(kind=1, vertices=c, radii=r)])
    L = 1.7 * _rot((1, 2, 3), 0.7)  # similarity: spheres stay spheres
    p = np.array([0.3, -0.2, 0.5])
    w = R.World([sc], [(0, 0, _xfm12(L, p))])
    rays = _rays(3000, 5)
    g = w.trace(rays)
    Lf = np.array(_xfm12(L, p), np.float64)
    Lw = np.stack([Lf[0:3], Lf[3:6], Lf[6:9]], 1)
    cw = c.astype(np.float64) @ Lw.T + Lf[9:12]
    rw = r.astype(np.float64) * 1.7
    o, d = rays[:, 0:3].astype(np.float64), rays[:, 4:7].astype(np.float64)
    oc = o[:, None, :] - cw[None]
    b = (oc * d[:, None, :]).sum(-1)
    cc = (oc * oc).sum(-1) - rw[None] ** 2
    disc = b * b - cc
    sq = np.sqrt(np.maximum(disc, 0))
    t1 = np.where((disc >= 0) & (-b - sq > 0), -b - sq, np.inf)
    t2 = np.where((disc >= 0) & (-b + sq > 0), -b + sq, np.inf)  # origin inside
    t = np.minimum(t1, t2)
    want = t.min(1)
    hit = np.isfinite(want)
    assert hit.mean() > 0.1
    assert ((g["primID"] >= 0) == hit).mean() > 0.999
    m = hit & (g["primID"] >= 0)
    np.testing.assert_allclose(g["t"][m], want[m], rtol=2e-5)
    # world normal of a sphere = (P - c) direction, scaled by the uniform factor
    P = o[m] + g["t"][m, None] * d[m]
    n = P - cw[g["primID"][m]]
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    gn = g["N"][m] / np.linalg.norm(g["N"][m], axis=1, keepdims=True)
    assert (np.abs(gn - n).max(1) < 1e-3).mean() > 0.999


def test_duplicate_instances_tie_to_smaller_index():
    v, idx = _tri_scene()
    sc = R.Scene([dict(kind=0, vertices=v, indices=idx)])
    w = R.World([sc], [(0, 11, ID), (0, 22, ID)])
    g = w.trace(_rays(2000, 6))
    hit = g["primID"] >= 0
    assert hit.any() and (g["instIndex"][hit] == 0).all() and (g["instID"][hit] == 11).all()


def test_rotated_triangle_normal():
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)  # normal +z
    sc = R.Scene([dict(kind=0, vertices=v, indices=np.array([[0, 1, 2]], np.int32))])
    L = _rot((1, 0, 0), math.pi / 2)  # +z -> -y
    w = R.World([sc], [(0, 0, _xfm12(L, np.zeros(3)))])
    ray = np.array([[0.2, -5, 0.2, 0, 0, 1, 0, 1e30]], np.float32)
    g = w.trace(ray)
    assert g["primID"][0] == 0
    np.testing.assert_allclose(g["t"][0], 5.0, rtol=1e-6)
    np.testing.assert_allclose(g["N"][0], (0, -1, 0), atol=1e-6)

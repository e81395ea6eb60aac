"""GPU closest-hit parity: bnamd_trace (BVH8 on gfx950) against the oracle.

Bar: hit/miss, primID, geomID bit-exact; t bit-exact (both sides run the same
fp32 intersector with no contraction); normals within 1e-5.
"""
import numpy as np
import pytest

from helpers import small_sphere_scene, user_prim_scene, to_product_geoms, random_rays
from oracle import refcpu

pytestmark = pytest.mark.gpu


def _compare(gpu_hits, ref_hits, label):
    gh, rh = gpu_hits, ref_hits
    hit_g = gh["primID"] >= 0
    hit_r = rh["primID"] >= 0
    mism = np.nonzero(hit_g != hit_r)[0]
    assert len(mism) == 0, f"{label}: hit/miss differs on {len(mism)} rays, first {mism[:5]}"
    assert np.array_equal(gh["primID"], rh["primID"]), f"{label}: primID differs"
    assert np.array_equal(gh["geomID"], rh["geomID"]), f"{label}: geomID differs"
    h = hit_r
    assert np.array_equal(gh["t"][h].view(np.uint32), rh["t"][h].view(np.uint32)), f"{label}: t not bit-exact"
    np.testing.assert_allclose(gh["N"][h], rh["N"][h], rtol=0, atol=1e-5, err_msg=label)
    return int(h.sum())


def _world(gpu, geoms):
    pg, mats = to_product_geoms(geoms)
    w = gpu.World([pg], materials=mats)
    w.validate()  # structure: references in range, each primitive once, boxes contain their content
    return w


def test_triangle_spheres_random_rays(gpu):
    geoms = small_sphere_scene(num_spheres=64)
    w = _world(gpu, geoms)
    ref = refcpu.Scene(geoms)
    rays = random_rays(20000, (0, 0, 0), (30, 30, 30))
    nhit = _compare(w.trace(rays), ref.trace(rays), "tessellated spheres")
    assert nhit > 2000


def test_user_prims_random_rays(gpu):
    geoms = user_prim_scene()
    w = _world(gpu, geoms)
    ref = refcpu.Scene(geoms)
    rays = random_rays(20000, (0, 0, 0), (20, 20, 20), seed=11)
    nhit = _compare(w.trace(rays), ref.trace(rays), "spheres+cylinders")
    assert nhit > 2000


def test_mixed_and_camera_rays(gpu):
    geoms = small_sphere_scene(num_spheres=16, extent=20.0) + user_prim_scene(50, 50)
    w = _world(gpu, geoms)
    ref = refcpu.Scene(geoms)
    cam = refcpu.camera((-10, 10, -10), (20, 0, 20), fovy=60)
    from helpers import camera_rays
    rays = camera_rays(cam, 96, 64)
    _compare(w.trace(rays), ref.trace(rays), "mixed camera rays")


def test_ties_resolve_to_lowest_geom(gpu):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    t = np.array([[0, 1, 2]], np.int32)
    geoms = [dict(kind=0, vertices=v, indices=t), dict(kind=0, vertices=v.copy(), indices=t.copy())]
    w = _world(gpu, geoms)
    rays = np.array([[0.2, 0.2, -1, 0, 0, 0, 1, 1e30]], np.float32)
    h = w.trace(rays)
    assert h["primID"][0] == 0 and h["geomID"][0] == 0
    assert h["t"][0] == np.float32(1.0)


def test_empty_and_single(gpu):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    t = np.array([[0, 1, 2]], np.int32)
    w = _world(gpu, [dict(kind=0, vertices=v, indices=t)])
    rays = np.array([[0.2, 0.2, -1, 0, 0, 0, 1, 1e30], [2, 2, -1, 0, 0, 0, 1, 1e30],
                     [0.2, 0.2, -1, 0, 0, 0, 1, 0.5]], np.float32)
    h = w.trace(rays)
    assert list(h["primID"]) == [0, -1, -1]
    w2 = gpu.World([[]])
    h2 = w2.trace(rays)
    assert list(h2["primID"]) == [-1, -1, -1]
    assert h2["t"][2] == np.float32(0.5)


def test_large_random_triangles(gpu):
    rng = np.random.default_rng(3)
    n = 200000
    c = rng.uniform(0, 100, (n, 3)).astype(np.float32)
    v = (c[:, None, :] + rng.uniform(-0.6, 0.6, (n, 3, 3))).astype(np.float32).reshape(-1, 3)
    t = np.arange(3 * n, dtype=np.int32).reshape(-1, 3)
    geoms = [dict(kind=0, vertices=v, indices=t)]
    w = _world(gpu, geoms)
    info = w.info()
    assert info["numPrims"] == n and info["depth"] >= 3
    ref = refcpu.Scene(geoms)
    rays = random_rays(50000, (0, 0, 0), (100, 100, 100), seed=5)
    _compare(w.trace(rays), ref.trace(rays), "soup")

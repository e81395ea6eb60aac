"""Two-level worlds on the GPU against the oracle's instanced worlds.

Groups of triangles and of spheres + cylinders are placed by rotations,
translations, uniform and non-uniform scales (a duplicate instance checks
the tie rule). The GPU walks the instance BVH8 and each group's BVH8 in
object space; the oracle loops over instances with its binary BVHs. Both
transform the ray with the same float32 inverse and arithmetic, so hits are
expected bit-exact: IDs equal and t bit-equal on >= 99.9% of rays (all
within 1e-5), normals within 1e-5."""
import math

import numpy as np
import pytest

from barney_amd import scenes
from oracle import refcpu

pytestmark = pytest.mark.gpu


def _rot(axis, ang):
    axis = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(ang) * K + (1 - math.cos(ang)) * K @ K


def _xfm(L, p):
    return tuple(float(x) for x in np.concatenate([L[:, 0], L[:, 1], L[:, 2], p]).astype(np.float32))


def _groups():
    v, t = scenes.sphere_mesh((0.0, 0.0, 0.0), 1.0, 12, 12)
    rng = np.random.default_rng(8)
    c = rng.uniform(-1, 1, (40, 3)).astype(np.float32)
    r = rng.uniform(0.05, 0.2, 40).astype(np.float32)
    a = rng.uniform(-1, 1, (30, 3))
    b = a + rng.uniform(-0.5, 0.5, (30, 3))
    cv = np.concatenate([a, b]).astype(np.float32)
    ci = np.stack([np.arange(30), np.arange(30) + 30], 1).astype(np.int32)
    cr = rng.uniform(0.03, 0.1, 30).astype(np.float32)
    g0 = [dict(kind=0, vertices=v, indices=t)]
    g1 = [dict(kind=1, vertices=c, radii=r), dict(kind=2, vertices=cv, indices=ci, radii=cr)]
    return [g0, g1]


INSTANCES = [
    (0, 100, _xfm(np.eye(3), np.array([0.0, 0.0, 0.0]))),
    (0, 101, _xfm(_rot((0, 1, 1), 0.9), np.array([2.5, 0.3, -0.4]))),
    (1, 102, _xfm(1.5 * _rot((1, 0, 2), 0.4), np.array([-2.5, 0.0, 0.5]))),
    (1, 103, _xfm(np.diag([1.0, 2.0, 0.5]) @ _rot((1, 1, 0), 1.1), np.array([0.0, 2.6, 0.0]))),
    (0, 104, _xfm(np.eye(3), np.array([0.0, 0.0, 0.0]))),  # duplicate of instance 0
]


def _worlds(gpu):
    groups = _groups()
    gg = [[gpu.Geom(g["kind"], g["vertices"], g.get("indices"), g.get("radii"), material=0) for g in grp]
          for grp in groups]
    world = gpu.World(gg, materials=[gpu.matte((0.7, 0.6, 0.5))], instances=INSTANCES)
    world.validate()  # both levels
    ref = refcpu.World([refcpu.Scene(grp) for grp in groups], INSTANCES)
    return world, ref


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-7, 7, (n, 3))
    tgt = rng.uniform(-3, 3, (n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, np.zeros((n, 1)), d, np.full((n, 1), 1e30)], 1).astype(np.float32)


def test_instanced_trace_matches_oracle(gpu):
    world, ref = _worlds(gpu)
    info = world.info()
    assert info["numPrims"] == 2 * 288 + 2 * 70 + 288  # instanced primitive count
    rays = _rays(100_000, 9)
    g, r = world.trace(rays), ref.trace(rays)
    for k in ("primID", "geomID", "instID", "instIndex"):
        agree = (g[k] == r[k]).mean()
        assert agree >= 0.999, (k, agree)
    hit = (g["primID"] >= 0) & (r["primID"] >= 0) & (g["instIndex"] == r["instIndex"])
    print(f"hits {hit.mean():.3f} by instance {np.bincount(g['instIndex'][hit], minlength=5)}", flush=True)
    assert hit.mean() > 0.2
    # every instance is hit, and the duplicate never wins a tie
    counts = np.bincount(g["instIndex"][hit], minlength=5)
    assert (counts[:4] > 0).all() and counts[4] == 0
    exact = (g["t"][hit].view(np.uint32) == r["t"][hit].view(np.uint32)).mean()
    print(f"t bit-exact {exact:.5f}", flush=True)
    assert exact >= 0.999
    np.testing.assert_allclose(g["t"][hit], r["t"][hit], rtol=1e-5)
    assert (np.abs(g["N"][hit] - r["N"][hit]).max(1) <= 1e-5 * np.abs(r["N"][hit]).max(1).clip(1)).mean() >= 0.999


def test_instanced_render_paths_agree(gpu):
    world, _ = _worlds(gpu)
    cams = [gpu.camera((-6.0, 4.0, -7.0), (6.0, -4.0, 7.0), fovy=50.0),
            gpu.camera((5.0, 1.0, -6.0), (-5.0, 0.0, 6.0), fovy=40.0)]
    views = gpu.Views(cams, [96, 64], [72, 80])
    p = gpu.render_params(0, 0, spp=4, ambient=1.0, bg=(-1, -1, -1, -1), max_depth=4, max_diffuse=2)
    mega = gpu.render_batch(world, views, p, gpu.BATCH_MEGAKERNEL)["accum"].numpy()
    wave = gpu.render_batch(world, views, p, gpu.BATCH_WAVEFRONT)["accum"].numpy()
    assert np.isfinite(mega).all() and mega[:, 3].mean() > 0.2
    np.testing.assert_allclose(wave, mega, rtol=1e-4, atol=1e-5)
    # the single-view megakernel renders the same first view
    out = gpu.render_surface(world, cams[0], gpu.render_params(96, 72, spp=4, max_depth=4, max_diffuse=2,
                                                               bg=(-1, -1, -1, -1)), with_ids=True)
    np.testing.assert_allclose(out["accum"].numpy().reshape(-1, 4), mega[:96 * 72], rtol=1e-5, atol=1e-6)
    ids = out["ids"]
    seen = set(np.unique(ids[..., 1])) - {-1}
    assert {100, 101, 13} <= seen  # user instance IDs (spheres report 13)

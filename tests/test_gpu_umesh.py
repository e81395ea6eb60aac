"""Unstructured tetrahedral volume on the GPU against the oracle.

Point location picks the containing tet of smallest index on both sides (a
point on a face shared by two tets takes the same tet whatever the
traversal order), and tetScalar, rasterTet<5> macrocell ranges, majorants,
DDA, Woodcock with fastLog, the iso-surface sub-stepping and its
central-difference normal, the shading step (shared fmath.h sine and
cosine) and the per-pixel path-order sums round alike on both sides, so:
  * point samples equal on >= 99.9% of points, all within 1e-6;
  * macrocell ranges and majorants bit-exact;
  * iso-surface t and normal per ray within 1e-5;
  * volume-only, iso-only and volume + iso-surface renders with a diffuse
    bounce equal the oracle's bit for bit, with equal sample and ray
    counts."""
import numpy as np
import pytest

from barney_amd import scenes
from oracle import refcpu

pytestmark = pytest.mark.gpu

N = 8


@pytest.fixture(scope="module")
def meshes(gpu):
    V, S, T = scenes.tet_grid(N, seed=0x5EED0404)
    g = gpu.UMesh(V, S, T)
    r = refcpu.UMesh(V, S, T)
    tf = scenes.ramp_tf()
    g.set_tf(tf, (0.0, 1.0), 0.6)
    r.set_tf(tf, (0.0, 1.0), 0.6)
    return V, S, T, g, r


def _cam(mod):
    c = N / 2.0
    pos = (c - 1.2 * N, c + 0.8 * N, c - 1.9 * N)
    return mod.camera(pos, (c - pos[0], c - pos[1], c - pos[2]), fovy=40.0)


def test_point_samples(meshes):
    V, S, T, g, r = meshes
    rng = np.random.default_rng(5)
    P = rng.uniform(-0.5, N + 0.5, (200_000, 3)).astype(np.float32)
    a, b = g.sample(P), r.sample(P)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    m = ~np.isnan(a)
    exact = (a[m] == b[m]).mean()
    print(f"inside {m.mean():.3f} exact {exact:.6f}", flush=True)
    assert exact >= 0.999
    assert np.abs(a[m] - b[m]).max() <= 1e-6


def test_macrocells_bit_exact(meshes):
    V, S, T, g, r = meshes
    assert g.info()["mcDims"] == r.mc_dims()
    gr, gm = g.macrocells()
    rr, rm = r.macrocells()
    assert np.array_equal(gr, rr)
    assert np.array_equal(gm, rm)
    assert (gm > 0).mean() > 0.3


def test_iso_rays(meshes, gpu):
    V, S, T, g, r = meshes
    rng = np.random.default_rng(6)
    n = 20_000
    org = rng.uniform(-2, N + 2, (n, 3))
    tgt = rng.uniform(1, N - 1, (n, 3))
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([org, np.zeros((n, 1)), d, np.full((n, 1), 1e30)], 1).astype(np.float32)
    a, b = g.iso_rays(rays, 0.5), r.iso_rays(rays, 0.5)
    ha, hb = np.isfinite(a[:, 0]), np.isfinite(b[:, 0])
    agree = (ha == hb).mean()
    print(f"hit {hb.mean():.3f} hit agreement {agree:.5f}", flush=True)
    assert agree >= 0.999 and hb.mean() > 0.2
    m = ha & hb
    rel = np.abs(a[m, 0] - b[m, 0]) / np.abs(b[m, 0])
    assert (rel <= 1e-5).mean() >= 0.999
    assert (np.abs(a[m, 1:] - b[m, 1:]).max(1) <= 1e-5).mean() >= 0.999


def _compare(gpu, meshes, volume, iso, max_diffuse, W=96, H=72, spp=4):
    V, S, T, g, r = meshes
    pg = gpu.render_params(W, H, spp=spp, ambient=1.0, bg=(-1, -1, -1, -1), max_diffuse=max_diffuse)
    pr = refcpu.render_params(W, H, spp=spp, ambient=1.0, bg=(-1, -1, -1, -1), max_diffuse=max_diffuse)
    col = (0.9, 0.7, 0.4)
    out = g.render(_cam(gpu), pg, volume=volume, iso=iso, iso_color=col)
    ga = out["accum"].numpy()
    ra, nr, ns = r.render(_cam(refcpu), pr, volume=volume, iso=iso, iso_color=col)
    assert np.isfinite(ga).all()
    rel = np.abs(ga - ra) / np.maximum(np.abs(ra), 1e-3)
    close = (rel.max(-1) <= 1e-3).mean()
    return ga, ra, close, out["stats"], nr, ns


@pytest.mark.parametrize("spp", [4, 8])
def test_volume_only_per_pixel(gpu, meshes, spp):
    ga, ra, close, st, nr, ns = _compare(gpu, meshes, True, None, 0, spp=spp)
    print(f"volume close {close:.5f} samples gpu {st['samples']} oracle {ns}", flush=True)
    assert np.array_equal(ga.view(np.uint32), ra.view(np.uint32))
    assert abs(st["samples"] - ns) <= 1e-3 * ns and st["raysTraced"] == nr
    assert ra[..., 3].mean() > 0.2


@pytest.mark.parametrize("spp", [4, 8])
def test_iso_only_per_pixel(gpu, meshes, spp):
    ga, ra, close, st, nr, ns = _compare(gpu, meshes, False, 0.5, 0, spp=spp)
    print(f"iso close {close:.5f}", flush=True)
    assert np.array_equal(ga.view(np.uint32), ra.view(np.uint32))
    assert st["raysTraced"] == nr
    assert ra[..., 3].mean() > 0.2


def test_volume_and_iso_with_bounce(gpu, meshes):
    ga, ra, close, st, nr, ns = _compare(gpu, meshes, True, 0.5, 1, spp=8)
    d = (ga[..., :3] - ra[..., :3]).reshape(-1, 3).astype(np.float64)
    z = d.mean(0) / (d.std(0) / np.sqrt(len(d)) + 1e-12)
    exact = np.all(ga == ra, axis=-1).mean()
    print(f"both close {close:.5f} bit-exact {exact:.5f} z {np.round(z, 2)}", flush=True)
    assert exact == 1.0
    assert np.array_equal(ga[..., 3], ra[..., 3])


def test_workload5_reduced_renders(gpu):
    w5 = scenes.workload5(n=24)
    m = gpu.UMesh(w5["vertices"], w5["scalars"], w5["tets"])
    m.set_tf(w5["tf"], (0.0, 1.0), w5["base_density"])
    info = m.info()
    assert info["numTets"] == 5 * 24 ** 3 and info["bvhDepth"] >= 2
    p = gpu.render_params(64, 64, spp=2, ambient=1.0, bg=(-1, -1, -1, -1))
    out = m.render(gpu.camera(**w5["camera"]), p, volume=True, iso=w5["iso"], iso_color=w5["iso_color"])
    a = out["accum"].numpy()
    assert np.isfinite(a).all() and a[..., :3].mean() > 0.01
    assert out["stats"]["samples"] > 0


def test_counted_render_equals_plain(gpu, meshes):
    """The instrumented kernel (point location's node visits, tet tests and
    their bytes counted for the roofline) renders the same image with the
    same samples as the plain one; a tet test is 80 bytes."""
    V, S, T, g, r = meshes
    p = gpu.render_params(64, 48, spp=8, ambient=1.0, bg=(-1, -1, -1, -1), max_diffuse=1)
    a = g.render(_cam(gpu), p, volume=True, iso=0.5)
    b = g.render(_cam(gpu), p, volume=True, iso=0.5, count=True)
    assert np.array_equal(a["accum"].numpy().view(np.uint32), b["accum"].numpy().view(np.uint32))
    sa, sb = a["stats"], b["stats"]
    assert sa["samples"] == sb["samples"] and sa["raysTraced"] == sb["raysTraced"]
    assert sa["nodeVisits"] == sa["cellTests"] == sa["cellBytes"] == 0
    print(f"nodes {sb['nodeVisits']} tests {sb['cellTests']} samples {sb['samples']}", flush=True)
    assert sb["nodeVisits"] > 0 and sb["cellTests"] > 0
    assert sb["cellBytes"] == 80 * sb["cellTests"]

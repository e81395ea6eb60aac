"""General unstructured cells (pyramids, prisms, hexahedra and tets in one
mesh, UMeshField's layout) on the GPU against the oracle.

The element tests (tetScalar and the Newton inversions of
ElementIntersection.h), the cell boxes, the macrocell rasterization
(rasterTet<5> for tets, rasterBox for other cells) and the walks are
exactly rounded on both sides, and where inside tests overlap (near faces
two cells share) both sides take the containing cell with the smallest
index, whatever their BVHs' orders:
  * point samples on non-affine (jittered) single elements: bit-exact,
    NaN outside included;
  * point samples on the sheared mixed mesh, random and within 1e-6 of
    shared faces (hex / polyhedron / prism / pyramid / tet neighbours):
    bit-exact;
  * macrocell ranges and majorants bit-exact;
  * volume-only and iso-only renders (NEE only) equal per pixel on >= 99.9 %
    of pixels, with the same ray counts."""
import numpy as np
import pytest

from barney_amd import scenes
from oracle import refcpu

pytestmark = pytest.mark.gpu

N = 6


@pytest.fixture(scope="module")
def mixed(gpu):
    V, S, cells = scenes.mixed_cells(N, seed=0x5EED0505)
    g = gpu.UMesh(V, S, cells=cells)
    r = refcpu.UMesh(V, S, cells=cells)
    tf = scenes.ramp_tf()
    g.set_tf(tf, (0.0, 1.0), 0.8)
    r.set_tf(tf, (0.0, 1.0), 0.8)
    return V, S, cells, g, r


def test_deformed_elements_bit_exact(gpu):
    V, S, cells = scenes.deformed_elements(seed=0x5EED0506)
    g = gpu.UMesh(V, S, cells=cells)
    r = refcpu.UMesh(V, S, cells=cells)
    assert g.info()["numTets"] == 3
    rng = np.random.default_rng(12)
    ix, off, ty = cells
    P = []
    for e, n in enumerate((5, 6, 8)):  # points in and around each element's box
        X = V[ix[off[e]:off[e] + n]]
        P.append(rng.uniform(X.min(0) - 0.1, X.max(0) + 0.1, (40_000, 3)))
    P = np.concatenate(P).astype(np.float32)
    a, b = g.sample(P), r.sample(P)
    inside = ~np.isnan(b)
    print(f"inside {inside.mean():.3f}", flush=True)
    assert inside.mean() > 0.2
    assert np.array_equal(a, b, equal_nan=True)


def test_mixed_point_samples(mixed):
    V, S, cells, g, r = mixed
    rng = np.random.default_rng(13)
    P = rng.uniform(V.min(0) - 0.3, V.max(0) + 0.3, (200_000, 3)).astype(np.float32)
    a, b = g.sample(P), r.sample(P)
    m = ~np.isnan(a)
    print(f"inside {m.mean():.3f} exact {(a[m] == b[m]).mean():.6f}", flush=True)
    assert m.mean() > 0.3
    assert np.array_equal(a, b, equal_nan=True)


def _near_shared_faces(n, count, seed):
    """Points within 1e-6 of the lattice faces that neighbouring cells share
    (one grid coordinate an interior integer), mapped by the mesh's shear."""
    rng = np.random.default_rng(seed)
    g = rng.uniform(0.0, float(n), (count, 3))
    axis = rng.integers(0, 3, count)
    g[np.arange(count), axis] = rng.integers(1, n, count)
    A = np.array([[1.0, 0.2, 0.1], [0.0, 1.0, 0.15], [0.05, 0.0, 1.0]])
    P = g @ A.T + rng.uniform(-1e-6, 1e-6, (count, 3))
    return P.astype(np.float32)


@pytest.mark.parametrize("which", ["mixed", "poly"])
def test_shared_face_points_bit_exact(which, request):
    """Near faces shared by two cells (a polyhedron and a hex among them) the
    overlapping inside tests are resolved by the smallest cell index on both
    sides, so the value does not depend on either side's BVH order."""
    V, S, cells, g, r = request.getfixturevalue(which)
    P = _near_shared_faces(N, 100_000, 15)
    a, b = g.sample(P), r.sample(P)
    m = ~np.isnan(a)
    print(f"{which}: inside {m.mean():.3f} exact {(a[m] == b[m]).mean():.6f}", flush=True)
    assert m.mean() > 0.9
    assert np.array_equal(a, b, equal_nan=True)


def test_mixed_macrocells_bit_exact(mixed):
    V, S, cells, g, r = mixed
    assert g.info()["mcDims"] == r.mc_dims()
    gr, gm = g.macrocells()
    rr, rm = r.macrocells()
    assert np.array_equal(gr, rr)
    assert np.array_equal(gm, rm)


def _cam(mod):
    c = N / 2.0
    pos = (c - 1.2 * N, c + 0.9 * N, c - 1.9 * N)
    return mod.camera(pos, (c - pos[0], c - pos[1], c - pos[2]), fovy=40.0)


@pytest.mark.parametrize("volume,iso,spp", [(True, None, 4), (False, 0.5, 4), (True, 0.5, 1)])
def test_mixed_renders_per_pixel(mixed, gpu, volume, iso, spp):
    """4 spp: four lanes per pixel; 1 spp: one."""
    V, S, cells, g, r = mixed
    W, H = 80, 60
    pg = gpu.render_params(W, H, spp=spp, ambient=1.0, bg=(-1, -1, -1, -1), max_diffuse=0)
    pr = refcpu.render_params(W, H, spp=spp, ambient=1.0, bg=(-1, -1, -1, -1), max_diffuse=0)
    col = (0.9, 0.7, 0.4)
    out = g.render(_cam(gpu), pg, volume=volume, iso=iso, iso_color=col)
    ga = out["accum"].numpy()
    ra, nr, ns = r.render(_cam(refcpu), pr, volume=volume, iso=iso, iso_color=col)
    assert np.isfinite(ga).all()
    rel = np.abs(ga - ra) / np.maximum(np.abs(ra), 1e-3)
    close = (rel.max(-1) <= 1e-3).mean()
    print(f"volume {volume} iso {iso} close {close:.5f} samples {out['stats']['samples']} {ns}", flush=True)
    assert np.array_equal(ga.view(np.uint32), ra.view(np.uint32))
    assert out["stats"]["raysTraced"] == nr
    assert ra[..., 3].mean() > 0.2


@pytest.fixture(scope="module")
def poly(gpu):
    V, S, cells = scenes.mixed_cells(N, seed=0x5EED0505, polyhedra=True)
    assert (cells[2] == scenes.CELL_POLYHEDRON).sum() > 20
    g = gpu.UMesh(V, S, cells=cells)
    r = refcpu.UMesh(V, S, cells=cells)
    tf = scenes.ramp_tf()
    g.set_tf(tf, (0.0, 1.0), 0.8)
    r.set_tf(tf, (0.0, 1.0), 0.8)
    return V, S, cells, g, r


def test_polyhedra_point_samples(poly):
    """Polyhedra (face streams): crossing parity + inverse-distance weights,
    the same float operations on both sides."""
    V, S, cells, g, r = poly
    P = np.random.default_rng(14).uniform(V.min(0) - 0.3, V.max(0) + 0.3, (200_000, 3)).astype(np.float32)
    a, b = g.sample(P), r.sample(P)
    m = ~np.isnan(a)
    print(f"inside {m.mean():.3f} exact {(a[m] == b[m]).mean():.6f}", flush=True)
    assert m.mean() > 0.3
    assert np.array_equal(a, b, equal_nan=True)


def test_polyhedra_macrocells_bit_exact(poly):
    V, S, cells, g, r = poly
    assert g.info()["mcDims"] == r.mc_dims()
    gr, gm = g.macrocells()
    rr, rm = r.macrocells()
    assert np.array_equal(gr, rr) and np.array_equal(gm, rm)


def test_polyhedra_render_per_pixel(poly, gpu):
    V, S, cells, g, r = poly
    W, H = 80, 60
    pg = gpu.render_params(W, H, spp=4, ambient=1.0, bg=(-1, -1, -1, -1), max_diffuse=0)
    pr = refcpu.render_params(W, H, spp=4, ambient=1.0, bg=(-1, -1, -1, -1), max_diffuse=0)
    col = (0.9, 0.7, 0.4)
    out = g.render(_cam(gpu), pg, volume=True, iso=0.5, iso_color=col)
    ga = out["accum"].numpy()
    ra, nr, ns = r.render(_cam(refcpu), pr, volume=True, iso=0.5, iso_color=col)
    assert np.isfinite(ga).all()
    rel = np.abs(ga - ra) / np.maximum(np.abs(ra), 1e-3)
    close = (rel.max(-1) <= 1e-3).mean()
    print(f"close {close:.5f} samples {out['stats']['samples']} {ns}", flush=True)
    assert np.array_equal(ga.view(np.uint32), ra.view(np.uint32))
    assert out["stats"]["raysTraced"] == nr
    assert ra[..., 3].mean() > 0.2


def test_bad_face_stream_rejected(gpu):
    V, S, _ = scenes.mixed_cells(2, seed=3)
    for bad in ([0], [1, 2, 0, 1], [1, 3, 0, 1, len(V)], [2, 3, 0, 1, 2]):
        with pytest.raises(Exception):
            gpu.UMesh(V, S, cells=(np.array(bad, np.int32), np.array([0], np.int32), np.array([42], np.uint8)))


@pytest.mark.parametrize("which", ["mixed", "poly"])
def test_counted_render_equals_plain(which, request, gpu):
    """The instrumented kernel renders the same image with the same samples
    as the plain one; every test reads at least a tet's 80 bytes."""
    V, S, cells, g, r = request.getfixturevalue(which)
    p = gpu.render_params(64, 48, spp=4, ambient=1.0, bg=(-1, -1, -1, -1), max_diffuse=0)
    a = g.render(_cam(gpu), p, volume=True, iso=0.5)
    b = g.render(_cam(gpu), p, volume=True, iso=0.5, count=True)
    assert np.array_equal(a["accum"].numpy().view(np.uint32), b["accum"].numpy().view(np.uint32))
    sb = b["stats"]
    assert a["stats"]["samples"] == sb["samples"]
    print(f"nodes {sb['nodeVisits']} tests {sb['cellTests']} bytes {sb['cellBytes']}", flush=True)
    assert sb["nodeVisits"] > 0 and sb["cellTests"] > 0
    assert sb["cellBytes"] >= 80 * sb["cellTests"]

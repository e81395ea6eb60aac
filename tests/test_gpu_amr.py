"""Block-structured AMR volume on the GPU against the oracle. Both visit the
blocks containing a point in ascending order through the same lookup grid
and evaluate the tent basis with the same float operations, and both
rasterize every cell's basis support into the macrocell grid the same way,
so samples and macrocells are bit-exact; delta-tracked renders (NEE only)
equal per pixel on >= 99.9 % of pixels with the same ray counts."""
import numpy as np
import pytest

from barney_amd import scenes
from oracle import refcpu

pytestmark = pytest.mark.gpu

N, B = 16, 8


@pytest.fixture(scope="module")
def field(gpu):
    f = scenes.amr_field(N, B, seed=0x5EED0707)
    args = (f["origins"], f["dims"], f["levels"], f["offsets"], f["refinements"], f["scalars"])
    g, r = gpu.Amr(*args), refcpu.Amr(*args)
    tf = scenes.ramp_tf()
    g.set_tf(tf, (0.0, 1.0), 0.25)
    r.set_tf(tf, (0.0, 1.0), 0.25)
    return f, g, r


def test_samples_bit_exact(field):
    f, g, r = field
    assert g.info()["numBlocks"] == len(f["origins"])
    P = np.random.default_rng(21).uniform(-3, f["extent"] + 3, (200_000, 3)).astype(np.float32)
    a, b = g.sample(P), r.sample(P)
    print(f"inside {(~np.isnan(b)).mean():.3f}", flush=True)
    assert (~np.isnan(b)).mean() > 0.5
    assert np.array_equal(a, b, equal_nan=True)


def test_macrocells_bit_exact(field):
    f, g, r = field
    assert g.info()["mcDims"] == r.mc_dims()
    gr, gm = g.macrocells()
    rr, rm = r.macrocells()
    assert np.array_equal(gr, rr) and np.array_equal(gm, rm)
    assert (gm > 0).mean() > 0.1


@pytest.mark.parametrize("spp", [1, 4])
def test_render_per_pixel(field, gpu, spp):
    f, g, r = field
    W, H = 72, 56
    c = f["extent"] / 2.0
    pos = (c - 1.3 * f["extent"], c + 0.7 * f["extent"], c - 1.6 * f["extent"])
    d = (c - pos[0], c - pos[1], c - pos[2])
    pg = gpu.render_params(W, H, spp=spp, ambient=1.0, bg=(-1, -1, -1, -1))
    pr = refcpu.render_params(W, H, spp=spp, ambient=1.0, bg=(-1, -1, -1, -1))
    out = g.render(gpu.camera(pos, d, fovy=35.0), pg)
    ga = out["accum"].numpy()
    ra, nr, ns = r.render(refcpu.camera(pos, d, fovy=35.0), pr)
    assert np.isfinite(ga).all()
    rel = np.abs(ga - ra) / np.maximum(np.abs(ra), 1e-3)
    close = (rel.max(-1) <= 1e-3).mean()
    print(f"spp {spp} close {close:.5f} samples {out['stats']['samples']} {ns}", flush=True)
    assert np.array_equal(ga.view(np.uint32), ra.view(np.uint32))
    assert out["stats"]["raysTraced"] == nr
    assert out["stats"]["samples"] == ns
    assert ra[..., 3].mean() > 0.2 * spp


def test_render_with_lights_bit_exact(field, gpu):
    """Quad, directional and point lights next to the ambient light: the
    lit instantiation of the shared volume path tracer equals the oracle."""
    f, g, r = field
    e = float(f["extent"])
    c = e / 2.0
    quads = [[0.0, 1.5 * e, 0.0, e, 0.0, 0.0, 0.0, 0.0, e, 3.0, 3.0, 3.0]]
    dirs = [[0.3, -1.0, 0.2, 1.0, 0.9, 0.8, 1.5]]
    points = [[c, c, c, 1.0, 1.0, 1.0, 4.0 * e * e, float("nan")]]
    W, H, spp = 64, 48, 4
    pos = (c - 1.3 * e, c + 0.7 * e, c - 1.6 * e)
    d = (c - pos[0], c - pos[1], c - pos[2])
    out = g.render(gpu.camera(pos, d, fovy=35.0),
                   gpu.render_params(W, H, spp=spp, ambient=0.3, lights=gpu.Lights(quads, dirs, points)))
    ra, nr, ns = r.render(refcpu.camera(pos, d, fovy=35.0),
                          refcpu.render_params(W, H, spp=spp, ambient=0.3, lights=refcpu.Lights(quads, dirs, points)))
    assert np.array_equal(out["accum"].numpy().view(np.uint32), ra.view(np.uint32))
    assert out["stats"]["raysTraced"] == nr and out["stats"]["samples"] == ns

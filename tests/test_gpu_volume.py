"""GPU structured-volume parity against the oracle: macrocell ranges and
majorants (bit-exact), per-ray delta tracking with barney's fastLog
(bit-exact t and color: same RNG stream, same fp32 arithmetic), and a small
delta-tracking render (per-pixel, 1e-5 relative)."""
import numpy as np
import pytest

from barney_amd import scenes
from oracle import refcpu
from helpers import random_rays

pytestmark = pytest.mark.gpu


def _field(n=40, seed=0x5EED0101):
    return scenes.structured_field(n=n, blobs=6, sigma=(3.0, 10.0), octaves=2, seed=seed)


def _pair(gpu, field, cpm=8, base=0.4):
    tf = scenes.ramp_tf()
    gv = gpu.Volume(field, origin=(1.0, -2.0, 0.5), spacing=(0.5, 0.5, 0.75), cells_per_mc=cpm)
    gv.set_tf(tf, (0.0, 1.0), base)
    rv = refcpu.Volume(field, origin=(1.0, -2.0, 0.5), spacing=(0.5, 0.5, 0.75), cells_per_mc=cpm)
    rv.set_tf(tf, (0.0, 1.0), base)
    return gv, rv


@pytest.mark.parametrize("cpm", [8, 16, 5])
def test_macrocells_bit_exact(gpu, cpm):
    gv, rv = _pair(gpu, _field(37), cpm=cpm)
    assert gv.mc_dims() == rv.mc_dims()
    gr, gm = gv.macrocells()
    rr, rm = rv.macrocells()
    assert np.array_equal(gr, rr)
    assert np.array_equal(gm.view(np.uint32), rm.view(np.uint32))


def test_delta_tracking_rays_bit_exact(gpu):
    f = _field(48)
    gv, rv = _pair(gpu, f)
    lo = np.array([1.0, -2.0, 0.5])
    hi = lo + np.array([0.5, 0.5, 0.75]) * 47
    rays = random_rays(20000, lo, hi, seed=21)
    seeds = (np.arange(len(rays), dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)) ^ np.uint64(0x1234)
    g = gv.sample_rays(rays, seeds)
    r, _ = rv.sample_rays(rays, seeds)
    hit = np.isfinite(r[:, 0])
    assert np.array_equal(np.isfinite(g[:, 0]), hit)
    assert hit.sum() > 1000
    assert np.array_equal(g.view(np.uint32), r.view(np.uint32))


def test_delta_tracking_exact_log_close(gpu):
    f = _field(32)
    gv, rv = _pair(gpu, f)
    lo = np.array([1.0, -2.0, 0.5])
    hi = lo + np.array([0.5, 0.5, 0.75]) * 31
    rays = random_rays(5000, lo, hi, seed=22)
    seeds = np.arange(len(rays), dtype=np.uint64) + np.uint64(99)
    g = gv.sample_rays(rays, seeds, exact_log=True)
    r, _ = rv.sample_rays(rays, seeds, exact_log=True)
    # logf may differ by an ulp between device and host libm; almost all rays
    # still take identical decisions
    same = (np.isfinite(g[:, 0]) == np.isfinite(r[:, 0]))
    assert same.mean() > 0.99
    both = np.isfinite(g[:, 0]) & np.isfinite(r[:, 0])
    close = np.abs(g[both, 0] - r[both, 0]) <= 1e-4 * np.maximum(1, np.abs(r[both, 0]))
    assert close.mean() > 0.99


def test_volume_render_matches_oracle(gpu):
    f = _field(40)
    gv, rv = _pair(gpu, f, base=0.3)
    c = np.array([1.0, -2.0, 0.5]) + np.array([0.5, 0.5, 0.75]) * 19.5
    pos = c + np.array([-25.0, 12.0, -30.0])
    W, H, spp = 48, 32, 4
    gcam = gpu.camera(pos, c - pos, fovy=40)
    rcam = refcpu.camera(pos, c - pos, fovy=40)
    gout = gv.render(gcam, gpu.render_params(W, H, spp=spp, bg=(0.1, 0.2, 0.3, 1.0)))
    racc, nrays, nsamp = rv.render(rcam, refcpu.render_params(W, H, spp=spp, bg=(0.1, 0.2, 0.3, 1.0)))
    gacc = gout["accum"].numpy()
    np.testing.assert_allclose(gacc, racc, rtol=1e-5, atol=1e-6)
    assert gout["stats"]["samples"] == nsamp
    assert gout["stats"]["raysTraced"] == nrays


@pytest.mark.parametrize("W,H,spp", [(50, 30, 4), (13, 7, 1), (9, 70, 16)])
def test_volume_render_every_pixel_any_frame_size(gpu, W, H, spp):
    """Tile counts that are not a multiple of the 8 XCDs: every pixel is
    written (the output starts as NaN) and equals the oracle."""
    f = _field(24)
    gv, rv = _pair(gpu, f, base=0.3)
    c = np.array([1.0, -2.0, 0.5]) + np.array([0.5, 0.5, 0.75]) * 11.5
    pos = c + np.array([-15.0, 8.0, -18.0])
    acc = gpu.DeviceArray(np.full((H, W, 4), np.nan, np.float32))
    gout = gv.render(gpu.camera(pos, c - pos, fovy=40), gpu.render_params(W, H, spp=spp), accum=acc)
    racc, nrays, nsamp = rv.render(refcpu.camera(pos, c - pos, fovy=40), refcpu.render_params(W, H, spp=spp))
    gacc = gout["accum"].numpy()
    assert np.isfinite(gacc).all()
    np.testing.assert_allclose(gacc, racc, rtol=1e-5, atol=1e-6)
    assert gout["stats"]["samples"] == nsamp


@pytest.mark.parametrize("cpm", [8, 5])
def test_iso_rays_bit_exact(gpu, cpm):
    """Structured iso-surface walk: hit t and normal equal the oracle's
    bit for bit (same macrocell ranges, sub-steps and gradient arithmetic)."""
    f = _field(40)
    gv, rv = _pair(gpu, f, cpm=cpm)
    lo = np.array([1.0, -2.0, 0.5])
    hi = lo + np.array([0.5, 0.5, 0.75]) * 39
    rays = random_rays(20000, lo, hi, seed=23)
    for iso in (0.3, 0.55, 2.0):
        g = gv.iso_rays(rays, iso)
        r = rv.iso_rays(rays, iso)
        hit = np.isfinite(r[:, 0])
        assert np.array_equal(np.isfinite(g[:, 0]), hit)
        if iso < 1.0:
            assert hit.sum() > 1000
        else:
            assert hit.sum() == 0
        assert np.array_equal(g.view(np.uint32), r.view(np.uint32))


@pytest.mark.parametrize("W,H,spp", [(48, 32, 4), (37, 23, 16)])
def test_volume_render_bit_exact(gpu, W, H, spp):
    """Each pixel adds its paths in path order, as the oracle does, and the
    volume path uses correctly rounded operations only: the accumulation
    equals the oracle's bit for bit, as do the sample and ray counts."""
    f = _field(40)
    gv, rv = _pair(gpu, f, base=0.3)
    c = np.array([1.0, -2.0, 0.5]) + np.array([0.5, 0.5, 0.75]) * 19.5
    pos = c + np.array([-25.0, 12.0, -30.0])
    p = dict(spp=spp, bg=(0.1, 0.2, 0.3, 1.0))
    gout = gv.render(gpu.camera(pos, c - pos, fovy=40), gpu.render_params(W, H, **p))
    racc, nrays, nsamp = rv.render(refcpu.camera(pos, c - pos, fovy=40), refcpu.render_params(W, H, **p))
    gacc = gout["accum"].numpy()
    assert np.array_equal(gacc.view(np.uint32), racc.view(np.uint32))
    assert gout["stats"]["samples"] == nsamp and gout["stats"]["raysTraced"] == nrays

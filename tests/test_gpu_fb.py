"""Frame-buffer readback on the GPU against the oracle, bit for bit.

Tile ownership, the tiled -> linear copy of color and normals, the aux
channels, the resolve of the renderers' linear accumulation and the
clamp-and-quantise pass are all integer work or one float multiply, so the
GPU's outputs equal the oracle's exactly. The sRGB formats go through powf;
a byte may differ only where the oracle's value sits within 1e-4 of a
quantisation step (the two powf implementations differ by an ulp), and
never by more than 1."""
import numpy as np
import pytest

from oracle import refcpu as R

pytestmark = pytest.mark.gpu

T = 32


def _tiles(num, seed):
    rng = np.random.default_rng(seed)
    acc = rng.uniform(-0.3, 4.0, (num, T * T, 4)).astype(np.float32)
    acc[:, ::97] = np.nan  # NaN samples must quantise to 0
    acc[:, 5::211, 0] = 1e30
    nrm = rng.normal(size=(num, T * T, 3)).astype(np.float32)
    return np.concatenate([acc.reshape(num, -1), nrm.reshape(num, -1)], 1)


def _srgb_close(got, want, ref_lin):
    """Bytes of R, G, B equal except ±1 where the oracle's linear input maps
    within 1e-4 of a k/256 step after the sRGB curve."""
    if np.array_equal(got, want):
        return True
    d = got.astype(np.int64) - want.astype(np.int64)
    for ch in range(3):
        dc = ((got >> (8 * ch)) & 0xFF).astype(np.int64) - ((want >> (8 * ch)) & 0xFF).astype(np.int64)
        x = ref_lin[..., ch].astype(np.float64)
        s = np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.power(np.maximum(x, 0), 1 / 2.4) - 0.055) * 256
        near = np.abs(s - np.round(s)) < 1e-4
        if (np.abs(dc) > 1).any() or (dc[~near & np.isfinite(s)] != 0).any():
            return False
    return ((got >> 24) == (want >> 24)).all() and np.count_nonzero(d) <= max(3, d.size // 10000)


@pytest.mark.parametrize("w,h,rank,size", [(1024, 768, 0, 1), (1000, 700, 0, 1), (100, 70, 1, 3), (33, 1, 0, 1),
                                           (1, 1, 0, 1), (64, 64, 5, 8)])
def test_tile_descs_match_oracle(gpu, w, h, rank, size):
    d = gpu.fb_tile_descs(w, h, rank, size).numpy()
    assert np.array_equal(d, R.fb_tile_descs(w, h, rank, size))


@pytest.mark.parametrize("w,h,rank,size", [(1024, 768, 0, 1), (1000, 700, 0, 1), (100, 70, 2, 3), (1, 1, 0, 1)])
@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_linearize_matches_oracle(gpu, w, h, rank, size, fmt):
    print(f"linearize {w}x{h} rank {rank}/{size} format {fmt}", flush=True)
    lower = R.fb_tile_descs(w, h, rank, size)
    tiles = _tiles(len(lower), w + h + fmt)
    scale = np.float32(1) / np.float32(7)
    fill = 0.5 if fmt == 0 else 0xA5A5A5A5
    want, wn = R.fb_linearize(tiles, lower, w, h, scale, fmt, with_normal=True, fill=fill)
    d_tiles, d_desc = gpu.DeviceArray(tiles), gpu.fb_tile_descs(w, h, rank, size)
    out = gpu.DeviceArray(np.full(want.shape, fill, want.dtype))
    nrm = gpu.DeviceArray(np.full((h, w, 3), fill, np.float32))
    got = gpu.fb_linearize(d_tiles, d_desc, w, h, scale, fmt, normal=nrm, out=out).numpy()
    gn = nrm.numpy()
    mine = gn[..., 0].view(np.uint32) != np.float32(fill).view(np.uint32)
    assert np.array_equal(gn.view(np.uint32), wn.view(np.uint32))
    if fmt == 2:
        acc = tiles[:, :T * T * 4].reshape(-1, T * T, 4) * scale
        ref_lin = np.zeros((h, w, 4), np.float32)
        for t, (lx, ly) in enumerate(lower):
            ys, xs = np.divmod(np.arange(T * T), T)
            ok = (lx + xs < w) & (ly + ys < h)
            ref_lin[ly + ys[ok], lx + xs[ok]] = acc[t, ok]
        assert _srgb_close(got, want, ref_lin)
    else:
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    if size > 1:
        assert 0 < mine.mean() < 1  # only the owned tiles were written


def test_linearize_aux_matches_oracle(gpu):
    w, h = 1000, 700
    lower = R.fb_tile_descs(w, h)
    aux = np.random.default_rng(2).integers(0, 2**32, (len(lower), T * T), dtype=np.uint64).astype(np.uint32)
    want = R.fb_linearize_aux(aux, lower, w, h)
    got = gpu.fb_linearize_aux(gpu.DeviceArray(aux), gpu.fb_tile_descs(w, h), w, h).numpy()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("w,h", [(1024, 768), (1920, 1080), (7, 3)])
@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_resolve_matches_oracle(gpu, w, h, fmt):
    rng = np.random.default_rng(w + fmt)
    acc = rng.uniform(-0.5, 16.0, (h, w, 4)).astype(np.float32)
    acc.reshape(-1, 4)[::331] = np.nan
    scale = np.float32(1) / np.float32(16)
    want = R.fb_resolve(acc, scale, fmt)
    got = gpu.fb_resolve(gpu.DeviceArray(acc), w, h, scale, fmt).numpy()
    if fmt == 2:
        assert _srgb_close(got, want, acc * scale)
    else:
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("srgb", [False, True])
def test_to_fixed8_matches_oracle(gpu, srgb):
    w, h = 640, 480
    rng = np.random.default_rng(9)
    v = rng.uniform(-0.5, 1.5, (h, w, 4)).astype(np.float32)
    v.reshape(-1, 4)[::101] = np.nan
    want = R.fb_to_fixed8(v, srgb)
    got = gpu.fb_to_fixed8(gpu.DeviceArray(v), w, h, srgb).numpy()
    if srgb:
        sat = np.clip(np.nan_to_num(v, nan=1.0), 0, 1)
        assert _srgb_close(got, want, sat)
    else:
        assert np.array_equal(got, want)


def test_resolve_of_a_rendered_frame(gpu):
    """Readback of a real accumulation buffer: a structured-volume frame of
    4 samples per pixel, read with accumScale 1/4."""
    from barney_amd import scenes
    w, h, spp = 160, 120, 4
    f = scenes.structured_field(n=40, blobs=6, sigma=(3.0, 10.0), octaves=2, seed=0x5EED0101)
    v = gpu.Volume(f, cells_per_mc=8)
    v.set_tf(scenes.ramp_tf(), (0.0, 1.0), 0.3)
    c = np.array([19.5, 19.5, 19.5])
    pos = c + np.array([-50.0, 24.0, -60.0])
    acc = v.render(gpu.camera(pos, c - pos, fovy=40), gpu.render_params(w, h, spp=spp))["accum"]
    host = acc.numpy().reshape(h, w, 4)
    assert np.isfinite(host).all()
    assert host[..., :3].std() > 0  # a frame with content (alpha is spp everywhere)
    for fmt in (0, 1, 2):
        got = gpu.fb_resolve(acc, w, h, 1.0 / spp, fmt).numpy()
        want = R.fb_resolve(host, np.float32(1.0 / spp), fmt)
        if fmt == 2:
            assert _srgb_close(got, want, host * np.float32(1.0 / spp))
        else:
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32))

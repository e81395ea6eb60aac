"""Pins of the oracle's frame-buffer readback (tile ownership, tiled ->
linear copy, accumulation scaling, RGBA8 and sRGB quantisation).

The reference holds no golden frames for this path, so each function is
pinned by what the mathematics fixes: 8-bit quantisation against its closed
form (including NaN/inf/negative and every k/256 boundary), the sRGB
transfer function against IEC 61966-2-1 in float64, tile ownership as a
partition of the tile grid, and the tiled readback as the inverse of an
independent numpy tiling of a random frame."""
import numpy as np
import pytest

from oracle import refcpu as R

T = R.TILE


def _q8(f):
    """Closed form: min(255, max(0, trunc(f * 256))), NaN -> 0."""
    s = np.asarray(f, np.float32) * np.float32(256)
    out = np.zeros(s.shape, np.uint32)
    ok = s > 0
    out[ok] = np.floor(np.minimum(s[ok], 255.5)).astype(np.uint32)
    out[s >= 256] = 255
    return np.minimum(out, 255)


def _srgb64(x):
    x = np.asarray(x, np.float64)
    return np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.power(np.maximum(x, 0), 1 / 2.4) - 0.055)


def _pack(r, g, b, a):
    return _q8(r) | (_q8(g) << 8) | (_q8(b) << 16) | (_q8(a) << 24)


def test_make_8bit_closed_form():
    assert R.fb_make_8bit(0.0) == 0 and R.fb_make_8bit(1.0) == 255 and R.fb_make_8bit(0.5) == 128
    assert R.fb_make_8bit(-0.3) == 0 and R.fb_make_8bit(7.0) == 255
    assert R.fb_make_8bit(float("nan")) == 0
    assert R.fb_make_8bit(float("inf")) == 255 and R.fb_make_8bit(float("-inf")) == 0
    assert R.fb_make_8bit(1e30) == 255  # the GPU's conversion saturates
    for k in range(1, 256):
        f = np.float32(k / 256)
        assert R.fb_make_8bit(f) == k
        assert R.fb_make_8bit(np.nextafter(f, np.float32(0))) == k - 1
    rng = np.random.default_rng(1)
    xs = rng.uniform(-0.5, 1.5, 20000).astype(np.float32)
    got = np.array([R.fb_make_8bit(x) for x in xs[:2000]], np.uint32)
    assert np.array_equal(got, _q8(xs[:2000]))


def test_linear_to_srgb_matches_standard():
    # linear segment exact below the knee, the IEC curve above it
    for x in (0.0, 1e-4, 0.001, 0.0031308):
        assert R.fb_linear_to_srgb(x) == np.float32(np.float32(12.92) * np.float32(x))
    xs = np.concatenate([np.linspace(0.0031309, 1.0, 3000), np.geomspace(0.0032, 8.0, 500)]).astype(np.float32)
    got = np.array([R.fb_linear_to_srgb(x) for x in xs])
    np.testing.assert_allclose(got, _srgb64(xs), rtol=4e-7, atol=1e-7)
    # published anchors: mid grey 0.18 -> 0.4614, 0.5 -> 0.7354, white -> 1
    assert abs(R.fb_linear_to_srgb(0.18) - 0.46135612) < 1e-6
    assert abs(R.fb_linear_to_srgb(0.5) - 0.73535698) < 1e-6
    assert abs(R.fb_linear_to_srgb(1.0) - 1.0) < 2e-7
    # continuity at the knee (the two segments meet within 1e-7)
    assert abs(R.fb_linear_to_srgb(0.0031308) - R.fb_linear_to_srgb(np.nextafter(np.float32(0.0031308), 1))) < 1e-6


@pytest.mark.parametrize("w,h,size", [(1024, 768, 1), (1024, 768, 3), (100, 70, 4), (33, 1, 2), (1, 1, 5), (0, 9, 2)])
def test_tile_ownership_partitions_the_grid(w, h, size):
    tx, ty = -(-w // T), -(-h // T)
    allt = []
    for rank in range(size):
        d = R.fb_tile_descs(w, h, rank, size)
        assert len(d) == max(0, -(-(tx * ty - rank) // size))
        ids = (d[:, 1] // T) * tx + d[:, 0] // T
        assert np.array_equal(ids, np.arange(rank, tx * ty, size))  # tiles g, g+G, ...
        assert ((d % T) == 0).all()
        allt.append(ids)
    allt = np.concatenate(allt) if allt else np.zeros(0)
    assert np.array_equal(np.sort(allt), np.arange(tx * ty))
    if size == 1 and tx * ty:
        d = R.fb_tile_descs(w, h)
        ys, xs = np.meshgrid(np.arange(ty) * T, np.arange(tx) * T, indexing="ij")
        assert np.array_equal(d, np.stack([xs.ravel(), ys.ravel()], 1))


def _tiles_of(img, nrm):
    """Independent numpy tiling: [H, W, 4] + [H, W, 3] -> [tiles, 1024*7]."""
    h, w = img.shape[:2]
    ty, tx = -(-h // T), -(-w // T)
    pad = lambda a: np.pad(a, ((0, ty * T - h), (0, tx * T - w), (0, 0)), constant_values=np.nan)
    a = pad(img).reshape(ty, T, tx, T, 4).transpose(0, 2, 1, 3, 4).reshape(ty * tx, T * T * 4)
    n = pad(nrm).reshape(ty, T, tx, T, 3).transpose(0, 2, 1, 3, 4).reshape(ty * tx, T * T * 3)
    return np.concatenate([a, n], 1).astype(np.float32)


def _frame(w, h, seed):
    rng = np.random.default_rng(seed)
    img = rng.uniform(-0.2, 3.0, (h, w, 4)).astype(np.float32)
    nrm = rng.normal(size=(h, w, 3)).astype(np.float32)
    return img, nrm


@pytest.mark.parametrize("w,h", [(96, 64), (100, 70), (33, 31), (1, 1)])
def test_linearize_inverts_tiling(w, h):
    img, nrm = _frame(w, h, w * 7 + h)
    tiles = _tiles_of(img, nrm)
    lower = R.fb_tile_descs(w, h)
    scale = np.float32(1.0) / np.float32(3)  # accumScale = 1/accumID
    col, n = R.fb_linearize(tiles, lower, w, h, scale, R.FB_FLOAT4, with_normal=True)
    assert np.array_equal(col, img * scale) and np.array_equal(n, nrm)
    c8 = R.fb_linearize(tiles, lower, w, h, scale, R.FB_RGBA8)
    s = img * scale
    assert np.array_equal(c8, _pack(s[..., 0], s[..., 1], s[..., 2], s[..., 3]))
    # byte order: R in the low byte
    assert np.array_equal(c8 & 0xFF, _q8(s[..., 0])) and np.array_equal(c8 >> 24, _q8(s[..., 3]))


def test_linearize_srgb_against_float64_curve():
    w, h = 128, 96
    img, nrm = _frame(w, h, 5)
    img[..., :3] = np.random.default_rng(6).uniform(0, 1, (h, w, 3))
    tiles = _tiles_of(img, nrm)
    c = R.fb_linearize(tiles, R.fb_tile_descs(w, h), w, h, 1.0, R.FB_SRGB8)
    for ch in range(3):
        want = _srgb64(img[..., ch]) * 256
        got = (c >> (8 * ch)) & 0xFF
        exp = np.clip(np.floor(want), 0, 255)
        # a byte may differ only where the float64 value sits on a boundary
        diff = got.astype(np.int64) - exp.astype(np.int64)
        near = np.abs(want - np.round(want)) < 1e-3
        assert (diff[~near] == 0).all() and (np.abs(diff) <= 1).all()
    assert np.array_equal(c >> 24, _q8(img[..., 3]))  # alpha stays linear


def test_linearize_by_rank_writes_disjoint_tiles():
    w, h, size = 100, 70, 3
    img, nrm = _frame(w, h, 11)
    tiles = _tiles_of(img, nrm)
    full = np.full((h, w, 4), -7.0, np.float32)
    for rank in range(size):
        lower = R.fb_tile_descs(w, h, rank, size)
        ids = (lower[:, 1] // T) * (-(-w // T)) + lower[:, 0] // T
        part = R.fb_linearize(tiles[ids], lower, w, h, 1.0, R.FB_FLOAT4, fill=-7.0)
        mine = part != -7.0
        assert not (mine & (full != -7.0)).any()
        full[mine] = part[mine]
    assert np.array_equal(full, img)


def test_aux_linearize_and_resolve_consistency():
    w, h = 70, 45
    rng = np.random.default_rng(3)
    aux = rng.integers(0, 2**32, (h, w), dtype=np.uint64).astype(np.uint32)
    ty, tx = -(-h // T), -(-w // T)
    padded = np.pad(aux, ((0, ty * T - h), (0, tx * T - w)), constant_values=0xDEADBEEF)
    auxt = np.ascontiguousarray(padded.reshape(ty, T, tx, T).transpose(0, 2, 1, 3).reshape(ty * tx, T * T))
    got = R.fb_linearize_aux(auxt, R.fb_tile_descs(w, h), w, h)
    assert np.array_equal(got, aux)
    # the linear resolve equals the tiled readback
    img4, nrm = _frame(w, h, 4)
    for fmt in (R.FB_FLOAT4, R.FB_RGBA8, R.FB_SRGB8):
        a = R.fb_resolve(img4, 0.25, fmt)
        b = R.fb_linearize(_tiles_of(img4, nrm), R.fb_tile_descs(w, h), w, h, 0.25, fmt)
        assert np.array_equal(a, b)


def test_to_fixed8_saturates_first():
    v = np.array([[2.0, -1.0, 0.5, 1.5], [np.nan, 0.25, 1.0, np.nan], [0.0031308, 0.1, 0.9, -0.0]], np.float32)
    lin = R.fb_to_fixed8(v, srgb=False)
    sat = np.clip(np.nan_to_num(v, nan=1.0), 0, 1)  # fminf/fmaxf: NaN -> 1
    assert np.array_equal(lin, _pack(sat[:, 0], sat[:, 1], sat[:, 2], sat[:, 3]))
    assert lin[1] & 0xFF == 255  # NaN saturates to 1 here, unlike make_8bit(NaN) = 0
    s = R.fb_to_fixed8(v, srgb=True)
    for ch in range(3):
        want = np.clip(np.floor(_srgb64(sat[:, ch]) * 256), 0, 255).astype(np.uint32)
        assert np.array_equal((s >> (8 * ch)) & 0xFF, want)
    assert np.array_equal(s >> 24, lin >> 24)

"""Pins of the oracle's block-structured AMR field (BlockStructuredField.h:
100-248, BlockStructuredCuBQLSampler.h:66-80, BlockStructuredField.cu:
33-105) against what the mathematics fixes:
  * the lookup grid's ascending block lists give the same samples, bit for
    bit, as visiting every block in ascending order;
  * the normalized tent basis reproduces a constant anywhere inside the
    block domains and a linear field (sampled at the cell centers) between
    the outermost cell centers, on one level and where two levels blend;
  * NaN exactly outside every block domain;
  * every sample lies within the macrocell range of the cell holding it,
    up to the rounding of the normalization (each cell's basis support is
    rasterized into the grid);
  * invalid blocks are rejected."""
import numpy as np
import pytest

from barney_amd import scenes
from oracle import refcpu as R


def _single(values_fn, dims=(6, 5, 7), origin=(2, 1, 3), level=0, ref=2):
    d = np.array(dims)
    z, y, x = np.meshgrid(np.arange(d[2]), np.arange(d[1]), np.arange(d[0]), indexing="ij")
    cs = float(ref ** level)
    centers = (np.stack([x.ravel(), y.ravel(), z.ravel()], -1) + np.array(origin) + 0.5) * cs
    s = values_fn(centers).astype(np.float32)
    a = R.Amr([origin], [dims], [level], [0], [ref, ref], s)
    lo = (np.array(origin) - 0.5) * cs
    hi = (np.array(origin) + d + 0.5) * cs
    return a, lo, hi, cs


def test_lookup_equals_every_block():
    f = scenes.amr_field(16, 8)
    a = R.Amr(f["origins"], f["dims"], f["levels"], f["offsets"], f["refinements"], f["scalars"])
    P = np.random.default_rng(1).uniform(-3, f["extent"] + 3, (30000, 3)).astype(np.float32)
    assert np.array_equal(a.sample(P), a.sample(P, brute=True), equal_nan=True)


def test_constant_and_nan_outside():
    a, lo, hi, cs = _single(lambda c: np.full(len(c), 0.625), level=1)
    rng = np.random.default_rng(2)
    P = rng.uniform(lo - 2, hi + 2, (20000, 3)).astype(np.float32)
    v = a.sample(P)
    inside = np.all((P >= lo) & (P <= hi), 1)
    assert np.isnan(v[~inside]).all()
    assert np.abs(v[inside] - 0.625).max() <= 1e-6


def test_linear_field_between_cell_centers():
    coef = np.array([0.11, -0.07, 0.05])
    a, lo, hi, cs = _single(lambda c: c @ coef + 0.3, level=0)
    rng = np.random.default_rng(3)
    P = rng.uniform(lo + cs, hi - cs, (20000, 3)).astype(np.float32)  # between the outermost centers
    assert np.abs(a.sample(P) - (P.astype(np.float64) @ coef + 0.3)).max() <= 2e-6


def test_linear_field_where_levels_blend():
    n, B = 16, 8
    coef = np.array([0.013, 0.021, -0.017])
    f = scenes.amr_field(n, B)
    # the same linear field at every level's cell centers
    s = []
    for o, d, L in zip(f["origins"], f["dims"], f["levels"]):
        cs = 2.0 ** L
        z, y, x = np.meshgrid(np.arange(d[2]), np.arange(d[1]), np.arange(d[0]), indexing="ij")
        c = (np.stack([x.ravel(), y.ravel(), z.ravel()], -1) + o + 0.5) * cs
        s.append(c @ coef + 0.5)
    a = R.Amr(f["origins"], f["dims"], f["levels"], f["offsets"], f["refinements"], np.concatenate(s))
    rng = np.random.default_rng(4)
    P = rng.uniform(n / 2 + 1, 3 * n / 2 - 1, (20000, 3)).astype(np.float32)  # inside the fine level
    assert np.abs(a.sample(P) - (P.astype(np.float64) @ coef + 0.5)).max() <= 2e-6


def test_samples_within_macrocell_ranges():
    f = scenes.amr_field(16, 8)
    a = R.Amr(f["origins"], f["dims"], f["levels"], f["offsets"], f["refinements"], f["scalars"])
    ranges, _ = a.macrocells()
    mz, my, mx = ranges.shape[:3]
    # world box = union of block domains; macrocell spacing as buildMCs sets it
    lo = np.float32(-1.0)
    size = np.float32(f["extent"] + 2.0)
    dims = np.array([mx, my, mz], np.float32)
    sp = (np.full(3, size, np.float32) * (np.float32(1) / dims)).astype(np.float32)
    rcp = (np.float32(1) / sp).astype(np.float32)
    P = np.random.default_rng(5).uniform(-0.9, f["extent"] + 0.9, (20000, 3)).astype(np.float32)
    v = a.sample(P)
    idx = np.clip(((P - lo) * rcp).astype(np.int64), 0, np.array([mx, my, mz]) - 1)
    r = ranges[idx[:, 2], idx[:, 1], idx[:, 0]]
    ok = ~np.isnan(v)
    assert ok.mean() > 0.9
    tol = 1e-6 * np.maximum(np.abs(v[ok]), 1e-30)  # (w s) / w rounds to a neighbour of s
    assert (r[ok, 0] - tol <= v[ok]).all() and (v[ok] <= r[ok, 1] + tol).all()


def test_invalid_blocks_rejected():
    with pytest.raises(ValueError):
        R.Amr([[0, 0, 0]], [[2, 2, 2]], [1], [0], [2], np.zeros(8, np.float32))  # level 1 of 1
    with pytest.raises(ValueError):
        R.Amr([[0, 0, 0]], [[2, 2, 2]], [0], [1], [2], np.zeros(8, np.float32))  # scalars out of range

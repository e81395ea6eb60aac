"""Oracle pins for NanoVDB fields (barney/volume/NanoVDB.{h,cu}).

The golden vectors (tests/golden/nanovdb_*.npz, written by
tests/golden/make_nanovdb_vectors.py) come from the NanoVDB library that the
reference vendors: grids made by createNanoGrid (float, Fp4, Fp8, Fp16, FpN,
leaf and upper-node tiles, two root tiles, a non-unit map), samples by
barney's sampler statement over createSampler<1>, macrocell ranges by
barney's computeMCs statement over the library's accessor. The oracle must
equal them bit for bit; the package's dense-array writer is checked against
the oracle, against closed forms and, where oracle/_ref is built, against the
library reading the writer's buffers.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from barney_amd import nanovdb
from oracle import refcpu

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = ["float", "fp4", "fp8", "fp16", "fpn", "tiles"]
NVDB_REF = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "nvdb_ref")


def fixture(name):
    return np.load(os.path.join(HERE, "golden", f"nanovdb_{name}.npz"))


@pytest.mark.parametrize("case", CASES)
def test_oracle_samples_equal_library(case):
    f = fixture(case)
    o = refcpu.NanoVDB(f["grid"])
    s = o.sample(f["points"])
    assert np.array_equal(s.view(np.uint32), f["samples"].view(np.uint32))


@pytest.mark.parametrize("case", CASES)
def test_oracle_boxes_and_macrocells_equal_library(case):
    f = fixture(case)
    o = refcpu.NanoVDB(f["grid"])
    b = o.boxes()
    assert np.array_equal(b["world_lo"], f["world_lo"]) and np.array_equal(b["world_hi"], f["world_hi"])
    assert np.array_equal(b["index_lo"], f["index_lo"]) and np.array_equal(b["index_hi"], f["index_hi"])
    assert o.mc_dims() == tuple(int(x) for x in f["mc_dims"])
    assert np.array_equal(b["mc_origin"], f["world_lo"]) and np.array_equal(b["mc_spacing"], f["mc_spacing"])
    ranges, _ = o.macrocells()
    assert np.array_equal(ranges.reshape(-1, 2), f["ranges"])


def test_fixtures_cover_tiles_and_types():
    # the cases exercise what the reader must handle
    types = {c: int(fixture(c)["grid_type"]) for c in CASES}
    assert types == {"float": 1, "fp4": 13, "fp8": 14, "fp16": 15, "fpn": 16, "tiles": 1}
    f = fixture("tiles")
    assert int(f["root_tiles"]) == 2 and int(f["index_lo"][0]) == -256  # the 128^3 upper-node tile
    o = refcpu.NanoVDB(f["grid"])
    # inside the upper-node tile: its value; the 8^3 leaf tile: 0.625; far
    # outside every node: the background
    v = o.values([[-200, 10, 10], [-129, 127, 0], [12, 12, 12], [5000, 0, 0], [-5000, 7, 7]])
    assert v.tolist() == [0.25, 0.25, 0.625, 0.0, 0.0]


def _writer_cases():
    rng = np.random.default_rng(11)
    return rng.random((13, 10, 17), dtype=np.float32) * 2 - 0.5


@pytest.mark.parametrize("vt", ["float", "fp4", "fp8", "fp16", "fpn"])
def test_writer_voxel_centers(vt):
    v = _writer_cases()
    s, t, i0 = 0.5, np.array([1.0, -2.0, 3.0]), np.array([-5, 3, -12])
    g = nanovdb.from_dense(v, voxel_size=s, origin=t, index_origin=i0, value_type=vt, fpn_bits=8)
    o = refcpu.NanoVDB(g)
    nz, ny, nx = v.shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    idx = np.stack([i, j, k], -1).reshape(-1, 3) + i0
    P = (s * (idx + 0.5) + t).astype(np.float32)  # voxel centers in world space
    got = o.sample(P).reshape(v.shape)
    vals = o.values(idx).reshape(v.shape)
    assert np.array_equal(got, vals)  # a center samples its own voxel, weights 0
    if vt == "float":
        assert np.array_equal(vals, v)
    else:
        # code * quantum + min of the voxel's leaf, within half a quantum
        width = {"fp4": 4, "fp8": 8, "fp16": 16, "fpn": 8}[vt]
        span = float(v.max() - v.min())
        assert np.abs(vals - v).max() <= 0.5 * span / ((1 << width) - 1) * 1.0001 + 1e-7


def test_writer_trilinear_closed_form():
    # a linear field is reproduced by trilinear interpolation between voxel
    # centers (inside the box; outside, positions clamp to the boundary)
    nz, ny, nx = 9, 11, 14
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    v = (0.25 * i - 0.5 * j + 0.125 * k + 1.0).astype(np.float32)
    g = nanovdb.from_dense(v, voxel_size=2.0, origin=(0.0, 0.0, 0.0), index_origin=(4, -9, 0))
    o = refcpu.NanoVDB(g)
    rng = np.random.default_rng(5)
    idxp = rng.uniform([4.5, -8.5, 0.5], [4 + nx - 0.5, -9 + ny - 0.5, nz - 0.5], (4000, 3))
    P = (2.0 * idxp).astype(np.float32)
    c = P.astype(np.float64) / 2.0 - 0.5 - np.array([4, -9, 0])
    expect = 0.25 * c[:, 0] - 0.5 * c[:, 1] + 0.125 * c[:, 2] + 1.0
    assert np.abs(o.sample(P) - expect).max() < 2e-5
    # clamping: far outside the box equals the nearest boundary voxel center
    far = np.array([[-100.0, -100.0, -100.0], [100.0, 100.0, 100.0]], np.float32)
    assert o.sample(far).tolist() == [float(v[0, 0, 0]), float(v[-1, -1, -1])]


@pytest.mark.skipif(not os.path.exists(NVDB_REF), reason="oracle/_ref/nvdb_ref not built (no reference headers)")
@pytest.mark.parametrize("vt", ["float", "fp4", "fp8", "fp16", "fpn"])
def test_writer_read_by_library(vt):
    rng = np.random.default_rng(2)
    v = rng.random((19, 23, 21), dtype=np.float32)
    g = nanovdb.from_dense(v, voxel_size=0.75, origin=(0.5, 1.5, -2.0), index_origin=(-9, -3, 5), value_type=vt,
                           fpn_bits=2)
    P = rng.uniform(-10, 20, (4000, 3)).astype(np.float32)
    with tempfile.TemporaryDirectory() as d:
        g.tofile(os.path.join(d, "g"))
        P.tofile(os.path.join(d, "p"))
        subprocess.run([NVDB_REF, "--sample", os.path.join(d, "g"), os.path.join(d, "p"), os.path.join(d, "o")],
                       check=True)
        lib = np.fromfile(os.path.join(d, "o"), np.float32)
    assert np.array_equal(refcpu.NanoVDB(g).sample(P).view(np.uint32), lib.view(np.uint32))


def test_malformed_buffers_rejected():
    g = fixture("float")["grid"].copy()
    with pytest.raises(ValueError):
        refcpu.NanoVDB(g[:4000])  # truncated: nodes outside the buffer
    bad = g.copy()
    bad[0] ^= 0xFF  # magic
    with pytest.raises(ValueError):
        refcpu.NanoVDB(bad)
    bad = g.copy()
    bad[636:640] = np.frombuffer(np.uint32(2).tobytes(), np.uint8)  # double grid: unsupported
    with pytest.raises(ValueError):
        refcpu.NanoVDB(bad)
    bad = g.copy()
    root = 672 + 64
    bad[root + 64 + 8:root + 64 + 16] = np.frombuffer(np.int64(1 << 40).tobytes(), np.uint8)  # tile child
    with pytest.raises(ValueError):
        refcpu.NanoVDB(bad)

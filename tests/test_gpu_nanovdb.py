"""NanoVDB fields on the GPU against the oracle and the NanoVDB library's
golden vectors. The GPU walks the grid buffer as ReadAccessor::getValue does
and evaluates barney's sampler statement (double index position, fp32 map
rows, SampleFromVoxels<1> lerps) with the same operations as the oracle, so
samples, macrocell ranges, iso hits and delta-tracked renders are bit-exact;
the golden vectors come from the library itself (tests/golden)."""
import os

import numpy as np
import pytest

from barney_amd import nanovdb, scenes
from oracle import refcpu

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = ["float", "fp4", "fp8", "fp16", "fpn", "tiles"]


def fixture(name):
    return np.load(os.path.join(HERE, "golden", f"nanovdb_{name}.npz"))


@pytest.mark.parametrize("case", CASES)
def test_golden_samples_and_macrocells(gpu, case):
    f = fixture(case)
    g = gpu.NanoVDB(f["grid"])
    info = g.info()
    assert info["gridType"] == int(f["grid_type"])
    assert info["indexLower"] == f["index_lo"].tolist() and info["indexUpper"] == f["index_hi"].tolist()
    assert info["mcDims"] == tuple(int(x) for x in f["mc_dims"])
    assert np.array_equal(np.float32(info["lower"]), f["world_lo"])
    assert np.array_equal(np.float32(info["upper"]), f["world_hi"])
    assert np.array_equal(np.float32(info["mcSpacing"]), f["mc_spacing"])
    s = g.sample(f["points"])
    print(f"{case}: {len(s)} samples", flush=True)
    assert np.array_equal(s.view(np.uint32), f["samples"].view(np.uint32))
    ranges, _ = g.macrocells()
    assert np.array_equal(ranges.reshape(-1, 2), f["ranges"])


def test_device_buffer_equals_host_buffer(gpu):
    f = fixture("fpn")
    a = gpu.NanoVDB(f["grid"]).sample(f["points"])
    b = gpu.NanoVDB(f["grid"], on_device=True).sample(f["points"])
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("case", ["tiles", "fp4"])
def test_tree_walk_without_leaf_table(gpu, case):
    f = fixture(case)
    g = gpu.NanoVDB(f["grid"], leaf_table=False)
    assert np.array_equal(g.sample(f["points"]).view(np.uint32), f["samples"].view(np.uint32))
    ranges, _ = g.macrocells()
    assert np.array_equal(ranges.reshape(-1, 2), f["ranges"])


def test_malformed_buffers_rejected(gpu):
    from barney_amd.capi import BnamdError
    g = fixture("float")["grid"]
    for bad in (g[:4000], np.concatenate([g[:636], np.frombuffer(np.uint32(2).tobytes(), np.uint8), g[640:]])):
        with pytest.raises(BnamdError):
            gpu.NanoVDB(bad)


def _blob_grid(vt, n=48):
    f = scenes.structured_field(n, blobs=6, sigma=(4.0, 12.0), seed=0x5EED0909)
    return nanovdb.from_dense(f, voxel_size=0.5, origin=(-3.0, 1.0, 2.0), index_origin=(-7, 5, 0), value_type=vt,
                              fpn_bits=4), n


@pytest.fixture(scope="module", params=["float", "fp8", "fpn"])
def blob(gpu, request):
    grid, n = _blob_grid(request.param)
    g, r = gpu.NanoVDB(grid), refcpu.NanoVDB(grid)
    tf = scenes.ramp_tf()
    g.set_tf(tf, (0.0, 1.0), 2.0)
    r.set_tf(tf, (0.0, 1.0), 2.0)
    return request.param, n, g, r


def test_random_samples_and_majorants(blob):
    vt, n, g, r = blob
    b = r.boxes()
    P = np.random.default_rng(4).uniform(b["world_lo"] - 2, b["world_hi"] + 2, (200_000, 3)).astype(np.float32)
    assert np.array_equal(g.sample(P).view(np.uint32), r.sample(P).view(np.uint32))
    gr, gm = g.macrocells()
    rr, rm = r.macrocells()
    assert np.array_equal(gr, rr) and np.array_equal(gm, rm)
    assert (gm > 0).mean() > 0.2


def test_iso_rays(blob):
    vt, n, g, r = blob
    b = r.boxes()
    rng = np.random.default_rng(8)
    m = 20_000
    rays = np.zeros((m, 8), np.float32)
    c = (b["world_lo"] + b["world_hi"]) / 2
    org = c + rng.normal(size=(m, 3)) * 40.0
    tgt = c + rng.uniform(-6, 6, (m, 3))
    d = tgt - org
    rays[:, 0:3] = org
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 7] = 1e30
    a, e = g.iso_rays(rays, 0.4), r.iso_rays(rays, 0.4)
    print(f"{vt}: iso hits {np.isfinite(e[:, 0]).mean():.3f}", flush=True)
    assert np.isfinite(e[:, 0]).mean() > 0.2
    assert np.array_equal(a.view(np.uint32), e.view(np.uint32))


@pytest.mark.parametrize("spp", [1, 4])
def test_render_bit_exact(blob, gpu, spp):
    vt, n, g, r = blob
    b = r.boxes()
    c = (b["world_lo"] + b["world_hi"]) / 2
    ext = float((b["world_hi"] - b["world_lo"]).max())
    W, H = 64, 48
    pos = (c[0] - 1.2 * ext, c[1] + 0.6 * ext, c[2] - 1.5 * ext)
    d = (c[0] - pos[0], c[1] - pos[1], c[2] - pos[2])
    pg = gpu.render_params(W, H, spp=spp, ambient=1.0, bg=(-1, -1, -1, -1))
    pr = refcpu.render_params(W, H, spp=spp, ambient=1.0, bg=(-1, -1, -1, -1))
    out = g.render(gpu.camera(pos, d, fovy=40.0), pg)
    ga = out["accum"].numpy()
    ra, nr, ns = r.render(refcpu.camera(pos, d, fovy=40.0), pr)
    print(f"{vt} spp {spp}: samples {out['stats']['samples']} {ns}", flush=True)
    assert np.array_equal(ga.view(np.uint32), ra.view(np.uint32))
    assert out["stats"]["raysTraced"] == nr and out["stats"]["samples"] == ns
    assert ra[..., 3].mean() > 0.1 * spp and ns > 1000


def test_render_with_lights_bit_exact(blob, gpu):
    """Quad, directional and point lights next to the ambient light: the
    lit instantiation of the shared volume path tracer equals the oracle."""
    vt, n, g, r = blob
    b = r.boxes()
    c = (b["world_lo"] + b["world_hi"]) / 2
    ext = float((b["world_hi"] - b["world_lo"]).max())
    quads = [[c[0] - ext / 2, c[1] + ext, c[2] - ext / 2, ext, 0.0, 0.0, 0.0, 0.0, ext, 3.0, 3.0, 3.0]]
    dirs = [[0.3, -1.0, 0.2, 1.0, 0.9, 0.8, 1.5]]
    points = [[c[0], c[1], c[2], 1.0, 1.0, 1.0, 4.0 * ext * ext, float("nan")]]
    W, H, spp = 48, 36, 4
    pos = (c[0] - 1.2 * ext, c[1] + 0.6 * ext, c[2] - 1.5 * ext)
    d = (c[0] - pos[0], c[1] - pos[1], c[2] - pos[2])
    out = g.render(gpu.camera(pos, d, fovy=40.0),
                   gpu.render_params(W, H, spp=spp, ambient=0.3, lights=gpu.Lights(quads, dirs, points)))
    ra, nr, ns = r.render(refcpu.camera(pos, d, fovy=40.0),
                          refcpu.render_params(W, H, spp=spp, ambient=0.3, lights=refcpu.Lights(quads, dirs, points)))
    assert np.array_equal(out["accum"].numpy().view(np.uint32), ra.view(np.uint32))
    assert out["stats"]["raysTraced"] == nr and out["stats"]["samples"] == ns

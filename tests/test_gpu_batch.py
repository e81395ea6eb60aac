"""Batched multi-view rendering: the wavefront renderer (with and without
graph capture) computes the same per-pixel values as the megakernel — same
per-path RNG streams and arithmetic, only the order of the float
accumulation differs — on a reduced mixed-material workload-3 scene
(heightfield + spheres + cylinders; Lambert, NVisii and Glass; Russian
roulette; uneven view sizes)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scene(gpu, views=5):
    from barney_amd import scenes
    w3 = scenes.workload3(grid=200, spheres=3000, cylinders=1500, views=views, seed=0x5EED0202)
    geoms = [gpu.Geom(g["kind"], g["vertices"], g.get("indices"), g.get("radii"), material=g["material"])
             for g in w3["geoms"]]
    world = gpu.World([geoms], materials=scenes.product_materials(gpu, w3["materials"]))
    world.validate()
    cams = [gpu.camera(**c) for c in w3["cameras"]]
    ws = [max(16, w // 8) for w in w3["widths"]]
    hs = [max(16, h // 8) for h in w3["heights"]]
    return world, gpu.Views(cams, ws, hs), w3


def _params(gpu, w3, spp=4, accum_id=0):
    return gpu.render_params(0, 0, spp=spp, accum_id=accum_id, ambient=1.0, bg=(-1, -1, -1, -1),
                             max_depth=w3["max_depth"], max_diffuse=w3["max_diffuse"],
                             russian_roulette=w3["rr_depth"])


def test_wavefront_equals_megakernel(gpu):
    world, views, w3 = _scene(gpu)
    p = _params(gpu, w3)
    mega = gpu.render_batch(world, views, p, gpu.BATCH_MEGAKERNEL)
    wave = gpu.render_batch(world, views, p, gpu.BATCH_WAVEFRONT)
    a, b = mega["accum"].numpy(), wave["accum"].numpy()
    assert np.isfinite(a).all() and np.isfinite(b).all()
    np.testing.assert_allclose(b, a, rtol=1e-4, atol=1e-5)
    assert mega["stats"]["raysTraced"] == wave["stats"]["raysTraced"]
    assert mega["stats"]["shadowRays"] == wave["stats"]["shadowRays"]
    assert a[:, :3].mean() > 0.01
    assert wave["stats"]["graphReplays"] >= 1


def test_wavefront_without_graph_and_progressive(gpu):
    world, views, w3 = _scene(gpu, views=3)
    os.environ["BNAMD_NO_GRAPH"] = "1"
    try:
        nog = gpu.render_batch(world, views, _params(gpu, w3, spp=2))["accum"].numpy()
    finally:
        del os.environ["BNAMD_NO_GRAPH"]
    g = gpu.render_batch(world, views, _params(gpu, w3, spp=2))["accum"].numpy()
    np.testing.assert_allclose(nog, g, rtol=1e-4, atol=1e-5)
    # two progressive frames of one path each == one frame of two paths
    acc = gpu.render_batch(world, views, _params(gpu, w3, spp=1, accum_id=0))["accum"]
    acc = gpu.render_batch(world, views, _params(gpu, w3, spp=1, accum_id=1), accum=acc)["accum"]
    np.testing.assert_allclose(acc.numpy(), g, rtol=1e-4, atol=1e-5)


def test_autotune_writes_choice(gpu, tmp_path):
    """The tuner searches the wavefront grid and fetch chunk, the wavefront
    queue size (paths per pixel per batch) and the megakernel tile (lanes
    per pixel), and picks the faster renderer; every tuned renderer still
    equals the untuned megakernel's image (wavefront within 1e-4)."""
    world, views, w3 = _scene(gpu, views=2)
    path = str(tmp_path / "tune.json")
    ref = gpu.render_batch(world, views, _params(gpu, w3, spp=4), gpu.BATCH_MEGAKERNEL)["accum"].numpy()
    gpu.batch_autotune(world, views, _params(gpu, w3, spp=4), path)
    import json
    d = json.load(open(path))
    print(d)
    assert d["traceBlocksPerCU"] in (0, 2, 3, 4) and d["chunk"] in (0, 64, 256)  # flattened world
    assert d["sppBatch"] in (1, 2, 4) and d["megaLanes"] in (1, 2, 4) and d["mode"] in ("megakernel", "wavefront")
    assert d["wavefrontMs"] > 0 and d["megakernelMs"] > 0
    mega = gpu.render_batch(world, views, _params(gpu, w3, spp=4), gpu.BATCH_MEGAKERNEL)["accum"].numpy()
    assert np.array_equal(mega, ref)  # lanes change the schedule, not the sums (path order)
    for mode in (gpu.BATCH_WAVEFRONT, gpu.BATCH_AUTO):
        out = gpu.render_batch(world, views, _params(gpu, w3, spp=4), mode)["accum"].numpy()
        assert np.isfinite(out).all()
        np.testing.assert_allclose(out, ref, rtol=1e-4, atol=1e-5)


def test_graph_replay_follows_render_parameters(gpu):
    """The captured wavefront graph bakes its kernel arguments, so two frames
    of the same world and views that differ only in ambient radiance and
    background must not replay the first frame's values: each wavefront
    frame equals the megakernel frame with the same parameters, and the two
    frames differ."""
    world, views, w3 = _scene(gpu, views=3)
    frames = []
    for ambient, bg in ((1.0, (-1, -1, -1, -1)), (0.25, (0.1, 0.3, 0.7, 1.0))):
        p = gpu.render_params(0, 0, spp=2, ambient=ambient, bg=bg, max_depth=w3["max_depth"],
                              max_diffuse=w3["max_diffuse"], russian_roulette=w3["rr_depth"])
        wave = gpu.render_batch(world, views, p, gpu.BATCH_WAVEFRONT)
        mega = gpu.render_batch(world, views, p, gpu.BATCH_MEGAKERNEL)
        assert wave["stats"]["graphReplays"] >= 1
        a, b = mega["accum"].numpy(), wave["accum"].numpy()
        np.testing.assert_allclose(b, a, rtol=1e-4, atol=1e-5)
        frames.append(b)
    assert np.abs(frames[0] - frames[1]).max() > 0.05

"""Quad, directional and point lights on the GPU against the oracle
(sampleAreaLights / sampleDirLights / samplePointLights and their choice in
sampleLights, barney/kernels/shadeRays.cu:46-404; the lights' getDD in
barney/light/). Both sides run the same float operations in the same order
and draw the same random numbers, so:
  * surface renders lit by each kind of light, by all of them, and by all of
    them next to an HDRI map equal the oracle's bit for bit (accumulation,
    ID channels and ray counts, 1 to 8 bounces);
  * the batched megakernel equals the oracle per view, the wavefront
    renderer (float atomics per fragment) agrees within 1e-4;
  * structured-volume and tet-mesh renders (shadow rays through the medium
    ending at a point or quad light) equal the oracle's bit for bit;
  * bad light lists are rejected with an error."""
import numpy as np
import pytest

from barney_amd import scenes
from oracle import refcpu
from test_gpu_envmap import CAM, _bits, _hdr, _pair, _surface_scene

pytestmark = pytest.mark.gpu

QUADS = [[-3.0, 12.0, -3.0, 6.0, 0.0, 0.0, 0.0, 0.0, 6.0, 4.0, 3.5, 3.0],  # overhead, facing down
         [12.0, 1.0, -4.0, 0.0, 0.0, 8.0, 0.0, 6.0, 0.0, 1.0, 2.0, 6.0]]  # a wall facing -x
DIRS = [[0.4, -1.0, 0.3, 1.0, 0.9, 0.7, 2.5], [-0.8, -0.5, 0.1, 0.3, 0.4, 1.0, 1.2]]
POINTS = [[0.0, 6.0, 0.0, 1.0, 1.0, 1.0, 40.0, float("nan")], [-6.0, 3.0, 5.0, 1.0, 0.5, 0.2, 1.0, 6.0],
          [5.0, 2.0, -6.0, 0.2, 0.6, 1.0, 1.0, 3.0]]

SETS = {"quad": dict(quads=QUADS), "dir": dict(dirs=DIRS), "point": dict(points=POINTS),
        "all": dict(quads=QUADS, dirs=DIRS, points=POINTS)}


def _lights(gpu, kind):
    return gpu.Lights(**SETS[kind]), refcpu.Lights(**SETS[kind])


@pytest.mark.parametrize("kind,env", [("quad", False), ("dir", False), ("point", False), ("all", False),
                                      ("all", True)])
def test_surface_render_with_lights_bit_exact(gpu, kind, env):
    gl, rl = _lights(gpu, kind)
    ge, re = _pair(gpu, _hdr(), direction=(1, 0.2, 0), up=(0, 1, 0), scale=0.5) if env else (None, None)
    world, ref = _surface_scene(gpu)
    W, H = 64, 40
    for spp, depth, amb in ((4, 1, 0.0), (4, 8, 0.3)):
        gp = gpu.render_params(W, H, spp=spp, ambient=amb, bg=(0.1, 0.2, 0.3, 1.0), max_depth=depth, max_diffuse=3,
                               env=ge, lights=gl)
        rp = refcpu.render_params(W, H, spp=spp, ambient=amb, bg=(0.1, 0.2, 0.3, 1.0), max_depth=depth,
                                  max_diffuse=3, env=re, lights=rl)
        out = gpu.render_surface(world, gpu.camera(**CAM), gp, with_ids=True)
        racc, rids, nr = ref.render(refcpu.camera(**CAM), rp, with_ids=True)
        g = out["accum"].numpy()
        exact = np.all(_bits(g) == _bits(racc), axis=-1).mean()
        print(f"{kind} env {env} spp {spp} depth {depth}: bit-exact {exact:.5f}, mean {racc[..., :3].mean():.4f}",
              flush=True)
        assert np.array_equal(_bits(g), _bits(racc))
        assert np.array_equal(out["ids"], rids)
        assert out["stats"]["raysTraced"] == nr
        assert racc[..., :3].mean() > 0.02  # the lights light the scene


def test_batched_renderers_with_lights(gpu):
    gl, rl = _lights(gpu, "all")
    world, ref = _surface_scene(gpu)
    cams = [CAM, dict(position=(10.0, 5.0, -12.0), direction=(-10.0, -4.0, 12.0), fovy=40.0)]
    ws, hs = [40, 33], [30, 27]
    views = gpu.Views([gpu.camera(**c) for c in cams], ws, hs)
    p = gpu.render_params(0, 0, spp=4, ambient=0.2, bg=(-1, -1, -1, -1), max_depth=6, max_diffuse=2, lights=gl)
    mega = gpu.render_batch(world, views, p, gpu.BATCH_MEGAKERNEL)["accum"].numpy()
    wave = gpu.render_batch(world, views, p, gpu.BATCH_WAVEFRONT)["accum"].numpy()
    for v, c in enumerate(cams):
        racc, _, _ = ref.render(refcpu.camera(**c), refcpu.render_params(ws[v], hs[v], spp=4, ambient=0.2,
                                                                        bg=(-1, -1, -1, -1), max_depth=6,
                                                                        max_diffuse=2, lights=rl))
        assert np.array_equal(_bits(views.image(mega, v)), _bits(racc))
    rel = np.abs(wave - mega) / np.maximum(np.abs(mega), 1e-3)
    assert rel.max() <= 1e-4


def _volume_lights(gpu):
    quads = [[0.0, 40.0, 0.0, 32.0, 0.0, 0.0, 0.0, 0.0, 32.0, 3.0, 3.0, 3.0]]  # above the box, facing down
    dirs = [[0.3, -1.0, 0.2, 1.0, 0.9, 0.8, 1.5]]
    points = [[15.5, 15.5, 15.5, 1.0, 1.0, 1.0, 4000.0, float("nan")]]  # inside the medium
    return gpu.Lights(quads, dirs, points), refcpu.Lights(quads, dirs, points)


def test_volume_render_with_lights_bit_exact(gpu):
    f = scenes.structured_field(n=32, blobs=5, sigma=(3.0, 8.0), octaves=2, seed=0x5EED0111)
    tf = scenes.ramp_tf()
    gv = gpu.Volume(f, origin=(0, 0, 0), spacing=(1, 1, 1), cells_per_mc=8)
    gv.set_tf(tf, (0.0, 1.0), 0.4)
    rv = refcpu.Volume(f, origin=(0, 0, 0), spacing=(1, 1, 1), cells_per_mc=8)
    rv.set_tf(tf, (0.0, 1.0), 0.4)
    gl, rl = _volume_lights(gpu)
    c = np.array([15.5, 15.5, 15.5])
    pos = c + np.array([-30.0, 14.0, -36.0])
    W, H, spp = 48, 36, 8
    gout = gv.render(gpu.camera(pos, c - pos, fovy=45), gpu.render_params(W, H, spp=spp, ambient=0.3, lights=gl))
    racc, nrays, nsamp = rv.render(refcpu.camera(pos, c - pos, fovy=45),
                                   refcpu.render_params(W, H, spp=spp, ambient=0.3, lights=rl))
    assert np.array_equal(_bits(gout["accum"].numpy()), _bits(racc))
    assert gout["stats"]["samples"] == nsamp and gout["stats"]["raysTraced"] == nrays
    assert racc[..., :3].max() > 0.0


def test_tet_mesh_render_with_lights_bit_exact(gpu):
    V, S, T = scenes.tet_grid(8, seed=0x5EED0404)
    g = gpu.UMesh(V, S, T)
    r = refcpu.UMesh(V, S, T)
    tf = scenes.ramp_tf()
    g.set_tf(tf, (0.0, 1.0), 0.6)
    r.set_tf(tf, (0.0, 1.0), 0.6)
    dirs = [[0.2, -1.0, 0.4, 1.0, 1.0, 1.0, 2.0]]
    points = [[4.0, 12.0, 4.0, 1.0, 0.8, 0.6, 300.0, float("nan")]]
    gl, rl = gpu.Lights(dirs=dirs, points=points), refcpu.Lights(dirs=dirs, points=points)
    pos = (4 - 9.6, 4 + 6.4, 4 - 15.2)
    d = (4 - pos[0], 4 - pos[1], 4 - pos[2])
    W, H, spp = 40, 30, 4
    out = g.render(gpu.camera(pos, d, fovy=40.0), gpu.render_params(W, H, spp=spp, max_diffuse=1, lights=gl),
                   volume=True, iso=0.5, iso_color=(0.8, 0.6, 0.4))
    ra, nr, ns = r.render(refcpu.camera(pos, d, fovy=40.0),
                          refcpu.render_params(W, H, spp=spp, max_diffuse=1, lights=rl),
                          volume=True, iso=0.5, iso_color=(0.8, 0.6, 0.4))
    assert np.array_equal(_bits(out["accum"].numpy()), _bits(ra))
    assert out["stats"]["raysTraced"] == nr


@pytest.mark.parametrize("bad", [dict(quads=[[0, 0, 0, 1, 0, 0, 2, 0, 0, 1, 1, 1]]),  # zero area
                                 dict(dirs=[[0, 0, 0, 1, 1, 1, 1]])])  # no direction
def test_bad_lights_rejected(gpu, bad):
    from barney_amd.capi import BnamdError
    with pytest.raises(BnamdError):
        gpu.Lights(**bad)

"""Models holding geometry and a structured volume on the GPU against the
oracle: each ray's closest surface hit first, then delta tracking through
the volume up to it (MCVolumeAccel::isProg with the surface hit as tMax,
barney/volume/MCAccelerator.h:569-708); shadow rays that no surface blocks
are tracked through the volume. Same float operations and random streams on
both sides, so accumulation, ID channels, ray and sample counts are
bit-exact, with ambient light, with quad / directional / point lights and
with a cut plane; on a two-level (instanced) world a volume hidden behind
the geometry leaves the surface render unchanged bit for bit."""
import numpy as np
import pytest

from barney_amd import scenes
from oracle import refcpu
from test_gpu_envmap import CAM, _bits, _surface_scene
from test_gpu_lights import DIRS, POINTS, QUADS

pytestmark = pytest.mark.gpu


def _volumes(gpu, density=1.5):
    f = scenes.structured_field(n=24, blobs=4, sigma=(2.0, 6.0), octaves=2, seed=0x5EED0902)
    tf = scenes.ramp_tf()
    kw = dict(origin=(-6.0, 0.2, -6.0), spacing=(0.5, 0.25, 0.5), cells_per_mc=4)
    g = gpu.Volume(f, **kw)
    g.set_tf(tf, (0.0, 1.0), density)
    r = refcpu.Volume(f, **kw)
    r.set_tf(tf, (0.0, 1.0), density)
    return g, r


@pytest.mark.parametrize("case", ["ambient", "lights", "cut"])
def test_mixed_render_bit_exact(gpu, case):
    world, ref = _surface_scene(gpu)
    gv, rv = _volumes(gpu)
    kw = {}
    if case == "lights":
        kw = dict(quads=QUADS, dirs=DIRS, points=POINTS)
    gl = gpu.Lights(**kw) if kw else None
    rl = refcpu.Lights(**kw) if kw else None
    cut = (0.3, 0.0, 1.0, 2.0) if case == "cut" else None
    W, H = 64, 40
    for spp, depth in ((4, 1), (4, 6)):
        gp = gpu.render_params(W, H, spp=spp, ambient=0.8, bg=(0.1, 0.2, 0.3, 1.0), max_depth=depth,
                               max_diffuse=2, lights=gl, cut_plane=cut)
        rp = refcpu.render_params(W, H, spp=spp, ambient=0.8, bg=(0.1, 0.2, 0.3, 1.0), max_depth=depth,
                                  max_diffuse=2, lights=rl, cut_plane=cut)
        out = gpu.render_mixed(world, gv, gpu.camera(**CAM), gp, with_ids=True)
        racc, rids, nr, ns = ref.render_mixed(rv, refcpu.camera(**CAM), rp, with_ids=True)
        g = out["accum"].numpy()
        exact = np.all(_bits(g) == _bits(racc), axis=-1).mean()
        print(f"{case} spp {spp} depth {depth}: bit-exact {exact:.5f}, samples {ns}", flush=True)
        assert np.array_equal(_bits(g), _bits(racc))
        assert np.array_equal(out["ids"], rids)
        assert out["stats"]["raysTraced"] == nr and out["stats"]["samples"] == ns
        assert ns > 0


def test_mixed_render_two_level_world(gpu):
    """An instanced (two-level) wall: with the volume hidden behind it, the
    mixed render equals the surface render bit for bit (no ray reaches the
    box, so no seed advances); with the volume in front, it is sampled and
    changes the image. The volume step is the flat worlds' code, tested
    against the oracle above."""
    S = 30.0
    wv = np.array([[-S, -S, 0], [S, -S, 0], [S, S, 0], [-S, S, 0]], np.float32)
    wi = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    c, sn = np.cos(0.2), np.sin(0.2)
    xf = (c, 0.0, -sn, 0.0, 1.0, 0.0, sn, 0.0, c, 0.5, 0.0, 1.0)  # rotated about y, moved
    world = gpu.World([[gpu.Geom(0, wv, wi, material=0)]], materials=[gpu.matte((0.7, 0.6, 0.5))],
                      instances=[(0, 7, xf)])
    cam = gpu.camera((0.0, 1.0, -14.0), (0.0, -0.05, 1.0), fovy=50.0)
    p = gpu.render_params(48, 36, spp=4, ambient=1.0, max_diffuse=2)
    surf = gpu.render_surface(world, cam, p)["accum"].numpy()
    f = scenes.structured_field(n=16, blobs=3, sigma=(2.0, 5.0), octaves=1, seed=0x5EED0903)
    behind = gpu.Volume(f, origin=(-4.0, -4.0, 6.0), spacing=(0.5, 0.5, 0.5), cells_per_mc=4)
    behind.set_tf(scenes.ramp_tf(), (0.0, 1.0), 3.0)
    out = gpu.render_mixed(world, behind, cam, p)
    assert out["stats"]["samples"] == 0
    assert np.array_equal(_bits(out["accum"].numpy()), _bits(surf))
    front = gpu.Volume(f, origin=(-4.0, -3.0, -9.0), spacing=(0.5, 0.5, 0.5), cells_per_mc=4)
    front.set_tf(scenes.ramp_tf(), (0.0, 1.0), 3.0)
    out = gpu.render_mixed(world, front, cam, p)
    assert out["stats"]["samples"] > 0
    assert not np.array_equal(_bits(out["accum"].numpy()), _bits(surf))


def test_mixed_render_needs_a_volume_with_tf(gpu):
    from barney_amd.capi import BnamdError
    world, _ = _surface_scene(gpu)
    f = scenes.structured_field(n=8, blobs=2, sigma=(1.0, 2.0), octaves=1, seed=1)
    gv = gpu.Volume(f)
    with pytest.raises(BnamdError):
        gpu.render_mixed(world, gv, gpu.camera(**CAM), gpu.render_params(8, 8, spp=1))

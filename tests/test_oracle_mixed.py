"""Pins of the oracle's mixed scenes (a model holding geometry and a
structured volume: the surface path tracer, each ray delta-tracked through
the volume up to its closest surface hit, MCVolumeAccel::isProg with the
surface hit as tMax, barney/volume/MCAccelerator.h:569-708):
  * a volume hidden behind an opaque wall changes nothing: every camera ray
    stops at the wall before the volume's box, bounce and shadow rays leave
    on the camera's side, so the image equals the surface render bit for
    bit (a walk to the ray's tMax instead of the surface hit would advance
    the seed and scatter behind the wall);
  * with the surfaces out of every ray's reach, the image equals the
    structured-volume render bit for bit (the volume path is pinned by its
    own tests), so the composition adds nothing of its own;
  * a volume in front of a wall is sampled and changes the pixels behind
    it."""
import numpy as np

from barney_amd import scenes
from oracle import refcpu as R


def _volume(origin=(-4.0, 0.5, -4.0), density=2.0):
    f = scenes.structured_field(n=16, blobs=3, sigma=(2.0, 5.0), octaves=1, seed=0x5EED0901)
    v = R.Volume(f, origin=origin, spacing=(0.5, 0.5, 0.5), cells_per_mc=4)
    v.set_tf(scenes.ramp_tf(), (0.0, 1.0), density)
    return v


def _wall(z=0.0, S=50.0):
    v = np.array([[-S, -S, z], [S, -S, z], [S, S, z], [-S, S, z]], np.float32)
    return dict(kind=0, vertices=v, indices=np.array([[0, 1, 2], [0, 2, 3]], np.int32), color=(0.7, 0.6, 0.5))


def test_volume_behind_a_wall_changes_nothing():
    s = R.Scene([_wall(z=0.0)])
    v = _volume(origin=(-4.0, -4.0, 2.0))  # the box spans z in [2, 9.5], behind the wall
    cam = R.camera((1.0, 2.0, -12.0), (0.0, -0.1, 1.0), fovy=50)
    for spp, md in ((4, 1), (4, 3)):
        p = R.render_params(40, 30, spp=spp, max_diffuse=md)
        a0, i0, n0 = s.render(cam, p, with_ids=True)
        a1, i1, n1, ns = s.render_mixed(v, cam, p, with_ids=True)
        assert ns == 0
        assert np.array_equal(a0.view(np.uint32), a1.view(np.uint32))
        assert np.array_equal(i0, i1) and n0 == n1


def test_out_of_reach_surfaces_give_the_volume_render():
    # one tiny triangle far behind the camera: no ray of the frame meets it
    far = dict(kind=0, vertices=np.array([[0, 0, -1e4], [0.01, 0, -1e4], [0, 0.01, -1e4]], np.float32),
               indices=np.array([[0, 1, 2]], np.int32))
    s = R.Scene([far])
    v = _volume()
    c = np.array([0.0, 4.5, 0.0])
    pos = c + np.array([-9.0, 6.0, -11.0])
    cam = R.camera(pos, c - pos, fovy=45)
    p = R.render_params(36, 28, spp=4, bg=(0.1, 0.2, 0.3, 1.0))
    av, nv, sv = v.render(cam, p)
    am, _, nm, sm = s.render_mixed(v, cam, p)
    assert sv == sm and sm > 0
    assert np.array_equal(av.view(np.uint32), am.view(np.uint32))


def test_volume_in_front_of_a_wall_is_seen():
    """Camera rays through the box collide in the volume before the wall:
    the volume's samples are taken and the pixel at the box's centre
    changes."""
    s = R.Scene([_wall(z=0.0)])
    v = _volume(origin=(-4.0, -4.0, -9.0), density=4.0)  # z in [-9, -1.5], between camera and wall
    cam = R.camera((0.0, 0.0, -20.0), (0.0, 0.0, 1.0), fovy=20)
    p = R.render_params(32, 32, spp=4, bg=(0.0, 0.0, 0.0, 1.0), max_diffuse=0)
    wall, _, _ = s.render(cam, p)
    mixed, _, _, ns = s.render_mixed(v, cam, p)
    assert ns > 0
    assert not np.array_equal(mixed[16, 16], wall[16, 16])

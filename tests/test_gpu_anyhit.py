"""Triangle anyHit features and the renderer's cut plane on the GPU against
the oracle: interpolated vertex normals (Triangles.dev.cu:70-74), stochastic
opacity of triangles whose material alpha is below 1 (:121-133; the draw of
a candidate depends on the ray's seed at the start of the trace and on the
candidate's object-space point only, so it does not depend on traversal
order), and the cut plane (generateRays.cu:171-193 for primary rays,
OptixGlobals.h:33-42 for every candidate of every ray). Renders equal the
oracle's bit for bit, the ID channel holds the closest candidate whatever
its opacity, and the batched renderers agree."""
import numpy as np
import pytest

from barney_amd import scenes
from oracle import refcpu

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _scene(gpu, with_normals=True):
    rng = np.random.default_rng(11)
    S = 30.0
    geoms = [dict(kind=0, vertices=np.array([[-S, 0, -S], [S, 0, -S], [S, 0, S], [-S, 0, S]], np.float32),
                  indices=np.array([[0, 2, 1], [0, 3, 2]], np.int32))]
    mats = [dict(type="matte", color=(0.7, 0.7, 0.65))]
    for k in range(10):
        v, t = scenes.sphere_mesh(rng.uniform((-8, 1, -8), (8, 4, 8)), rng.uniform(0.8, 2.2), 12, 12)
        g = dict(kind=0, vertices=v, indices=t)
        if with_normals and k % 2 == 0:
            c = v.mean(0)
            n = v - c
            g["normals"] = (n / np.linalg.norm(n, axis=1, keepdims=True)).astype(np.float32)
        geoms.append(g)
        col = tuple(rng.uniform(0.2, 0.9, 3))
        mats.append([dict(type="matte", color=col, opacity=0.45),
                     dict(type="pbr", color=col, metallic=0.3, roughness=0.4, transmission=0.0, ior=1.0,
                          opacity=0.6),
                     dict(type="matte", color=col),
                     dict(type="pbr", color=col, metallic=0.7, roughness=0.2),
                     dict(type="glass", ior=1.5)][k % 5])
    geoms.append(dict(kind=1, vertices=rng.uniform((-10, 0.5, -10), (10, 3, 10), (30, 3)).astype(np.float32),
                      radii=rng.uniform(0.2, 0.7, 30).astype(np.float32)))
    mats.append(dict(type="matte", color=(0.9, 0.6, 0.3)))
    geoms.append(dict(kind=2, vertices=rng.uniform((-10, 0.2, -10), (10, 4, 10), (24, 3)).astype(np.float32),
                      indices=np.arange(24, dtype=np.int32).reshape(12, 2),
                      radii=rng.uniform(0.1, 0.3, 12).astype(np.float32)))
    mats.append(dict(type="matte", color=(0.3, 0.8, 0.4)))
    for k, g in enumerate(geoms):
        g["material"] = k
    gg = [gpu.Geom(g["kind"], g["vertices"], g.get("indices"), g.get("radii"), material=k, user_id=k,
                   normals=g.get("normals")) for k, g in enumerate(geoms)]
    world = gpu.World([gg], materials=scenes.product_materials(gpu, mats))

    def omat(m):
        if m["type"] == "matte":
            return dict(type="matte", color=m["color"], opacity=m.get("opacity", 1.0))
        if m["type"] == "pbr":
            return dict(type="pbr", color=m["color"], metallic=m["metallic"], roughness=m["roughness"],
                        transmission=m.get("transmission", 0.0), ior=m.get("ior", 1.45),
                        opacity=m.get("opacity", 1.0))
        return dict(type="pbr", color=(1, 1, 1), metallic=0.0, roughness=0.5, transmission=1.0, ior=m["ior"])

    ref = refcpu.Scene([dict(kind=g["kind"], vertices=g["vertices"], indices=g.get("indices"), radii=g.get("radii"),
                             normals=g.get("normals"), material=omat(mats[k]), user_id=k)
                        for k, g in enumerate(geoms)])
    return world, ref


CAMS = [dict(position=(-14.0, 9.0, -16.0), direction=(14.0, -7.0, 16.0), fovy=50.0),
        dict(position=(12.0, 4.0, 14.0), direction=(-12.0, -2.0, -14.0), fovy=40.0)]
PLANES = [None, (0.3, 0.0, 1.0, -1.0), (0.0, -1.0, 0.2, 2.5)]


@pytest.mark.parametrize("plane", PLANES)
def test_surface_render_bit_exact(gpu, plane):
    world, ref = _scene(gpu)
    W, H = 64, 48
    for spp, depth in ((4, 1), (2, 8)):
        gp = gpu.render_params(W, H, spp=spp, bg=(0.1, 0.2, 0.3, 1.0), max_depth=depth, max_diffuse=3,
                               cut_plane=plane)
        rp = refcpu.render_params(W, H, spp=spp, bg=(0.1, 0.2, 0.3, 1.0), max_depth=depth, max_diffuse=3,
                                  cut_plane=plane)
        out = gpu.render_surface(world, gpu.camera(**CAMS[0]), gp, with_ids=True)
        racc, rids, nr = ref.render(refcpu.camera(**CAMS[0]), rp, with_ids=True)
        g = out["accum"].numpy()
        exact = np.all(_bits(g) == _bits(racc), axis=-1).mean()
        print(f"plane {plane} spp {spp} depth {depth}: bit-exact {exact:.5f}", flush=True)
        assert np.array_equal(_bits(g), _bits(racc))
        assert np.array_equal(out["ids"], rids)
        assert out["stats"]["raysTraced"] == nr


def test_batched_renderers_bit_exact_and_agree(gpu):
    world, ref = _scene(gpu)
    ws, hs = [40, 35], [30, 26]
    views = gpu.Views([gpu.camera(**c) for c in CAMS], ws, hs)
    plane = PLANES[1]
    p = gpu.render_params(0, 0, spp=4, bg=(-1, -1, -1, -1), max_depth=6, max_diffuse=2, cut_plane=plane)
    mega = gpu.render_batch(world, views, p, gpu.BATCH_MEGAKERNEL, with_ids=True)
    wave = gpu.render_batch(world, views, p, gpu.BATCH_WAVEFRONT, with_ids=True)
    m, w = mega["accum"].numpy(), wave["accum"].numpy()
    for v, c in enumerate(CAMS):
        racc, rids, _ = ref.render(refcpu.camera(**c), refcpu.render_params(ws[v], hs[v], spp=4, bg=(-1, -1, -1, -1),
                                                                            max_depth=6, max_diffuse=2,
                                                                            cut_plane=plane), with_ids=True)
        assert np.array_equal(_bits(views.image(m, v)), _bits(racc))
        assert np.array_equal(views.image(mega["ids"], v), rids)
        assert np.array_equal(views.image(wave["ids"], v), rids)
    rel = np.abs(w - m) / np.maximum(np.abs(m), 1e-3)
    assert rel.max() <= 1e-4


def test_instanced_cutouts_equal_flattened(gpu):
    """An instanced world (one bottom level per group, identity transforms
    through the two-level path) draws the same opacity decisions as the
    flattened world: the draw depends on the object-space point only."""
    v, t = scenes.sphere_mesh((0.0, 2.0, 0.0), 2.0, 16, 16)
    mats = scenes.product_materials(gpu, [dict(type="matte", color=(0.8, 0.3, 0.2), opacity=0.5),
                                          dict(type="matte", color=(0.6, 0.6, 0.6))])
    S = 20.0
    ground = gpu.Geom(0, np.array([[-S, 0, -S], [S, 0, -S], [S, 0, S], [-S, 0, S]], np.float32),
                      np.array([[0, 2, 1], [0, 3, 2]], np.int32), material=1)
    ball = gpu.Geom(0, v, t, material=0)
    flat = gpu.World([[ground, ball]], materials=mats)
    # the same geometry through the two-level path: a small translation
    # forces it, applied to nothing that moves (both instances carry it)
    xf = (1, 0, 0, 0, 1, 0, 0, 0, 1, 0.0, 0.0, 1e-30)
    inst = gpu.World([[ground], [ball]], materials=mats, instances=[(0, 0, xf), (1, 1, xf)])
    p = gpu.render_params(48, 36, spp=2, max_depth=4)
    cam = gpu.camera((6.0, 5.0, -8.0), (-6.0, -3.0, 8.0), fovy=45.0)
    a = gpu.render_surface(flat, cam, p)["accum"].numpy()
    b = gpu.render_surface(inst, cam, p)["accum"].numpy()
    assert inst.info()["numPrims"] > 0
    rel = np.abs(a - b) / np.maximum(np.abs(a), 1e-3)
    print(f"instanced vs flattened: max rel {rel.max():.2e}, exact {np.all(a == b, -1).mean():.4f}")
    assert (rel.max(-1) <= 1e-4).mean() >= 0.99

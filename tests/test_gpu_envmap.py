"""HDRI environment light and 2D textures on the GPU against the oracle
(barney/light/EnvMap.cu, EnvMap.h; shadeRays.cu radianceFromEnv and the
env branch of sampleLights; the embree backend's tex2D): every value is
computed with the same float operations on both sides (shared acos, atan2,
sine and cosine of fmath.h, serial CDF sums in EnvMap.cu's order), so:
  * tex2D samples (bilinear wrap / clamp / mirror / border, nearest; float4,
    float, uchar4, uchar texels), the CDF tables and the frame are bit-exact;
  * the light's pdf, radiance and importance samples are bit-exact;
  * surface, volume and tet-mesh renders under a map (misses show it, NEE
    samples it, bounces weight it by MIS) equal the oracle's bit for bit,
    and the batched megakernel and wavefront renderers agree with them."""
import numpy as np
import pytest

from barney_amd import scenes
from oracle import refcpu

pytestmark = pytest.mark.gpu


def _hdr(h=32, w=64, seed=0):
    rng = np.random.default_rng(seed)
    t = rng.uniform(0.0, 1.0, (h, w, 4)).astype(np.float32)
    t[..., :3] *= rng.uniform(0.2, 4.0, (h, 1, 1)).astype(np.float32)
    t[h // 4: h // 4 + 2, w // 3: w // 3 + 3, :3] = 60.0  # a sun
    return t


def _pair(gpu, texels, filter="linear", address=("wrap", "clamp"), direction=(1, 0, 0), up=(0, 0, 1), scale=1.0):
    gt = gpu.Texture2D(texels, filter=filter, address=address)
    ge = gpu.EnvMap(gt, direction=direction, up=up, scale=scale)
    re = refcpu.EnvMap(texels, direction=direction, up=up, scale=scale, filter=filter, address=address)
    return ge, re


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


FORMATS = [
    ("float4", lambda t: t),
    ("float", lambda t: np.ascontiguousarray(t[..., 0])),
    ("uchar4", lambda t: np.clip(t * 40, 0, 255).astype(np.uint8)),
    ("uchar", lambda t: np.clip(t[..., 1] * 90, 0, 255).astype(np.uint8)),
]


@pytest.mark.parametrize("fmt", [f[0] for f in FORMATS])
@pytest.mark.parametrize("filter,address", [("linear", ("wrap", "clamp")), ("linear", ("mirror", "border")),
                                            ("linear", ("clamp", "wrap")), ("point", ("wrap", "wrap"))])
def test_tex2d_bit_exact(gpu, fmt, filter, address):
    texels = dict(FORMATS)[fmt](_hdr(13, 29))
    ge, re = _pair(gpu, texels, filter, address)
    rng = np.random.default_rng(2)
    uv = np.concatenate([rng.uniform(-2.5, 2.5, (4000, 2)), rng.uniform(0, 1, (4000, 2)),
                         np.array([[0, 0], [1, 1], [0.5, 0.5], [-0.0, 1.0], [1e-7, -1e-7]])]).astype(np.float32)
    g = ge.texture.sample(uv)
    r = re.tex2d(uv)
    assert np.array_equal(_bits(g), _bits(r))


@pytest.mark.parametrize("h,w", [(32, 64), (17, 5), (1, 1), (128, 256)])
def test_cdfs_and_frame_bit_exact(gpu, h, w):
    ge, re = _pair(gpu, _hdr(h, w, seed=h), direction=(0.3, -0.4, 0.8), up=(0.1, 1.0, 0.3))
    gy, gx = ge.cdfs()
    ry, rx = re.cdfs()
    assert np.array_equal(_bits(gy), _bits(ry)) and np.array_equal(_bits(gx), _bits(rx))
    assert gy[-1] == 1.0 and (gx[:, -1] == 1.0).all()
    for a, b in zip(ge.frame(), re.frame()):
        assert np.array_equal(_bits(a), _bits(b))


@pytest.mark.parametrize("filter", ["linear", "point"])
def test_pdf_radiance_and_samples_bit_exact(gpu, filter):
    ge, re = _pair(gpu, _hdr(), filter=filter, direction=(0.2, 0.9, -0.1), up=(0, 0.3, 1), scale=1.7)
    rng = np.random.default_rng(4)
    d = rng.normal(size=(20000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = np.concatenate([d, np.eye(3), -np.eye(3)]).astype(np.float32)
    gp, gr = ge.eval(d)
    rp, rr = re.eval(d)
    assert np.array_equal(_bits(gp), _bits(rp)) and np.array_equal(_bits(gr), _bits(rr))
    for seed in (0, 11):
        gs = ge.sample(30000, seed=seed)
        rs = re.sample(30000, seed=seed)
        for a, b in zip(gs, rs):
            assert np.array_equal(_bits(a), _bits(b))


def _surface_scene(gpu):
    """Matte, PBR and glass triangles and spheres over a ground plane."""
    rng = np.random.default_rng(7)
    S = 30.0
    ground = dict(kind=0, vertices=np.array([[-S, 0, -S], [S, 0, -S], [S, 0, S], [-S, 0, S]], np.float32),
                  indices=np.array([[0, 2, 1], [0, 3, 2]], np.int32))
    geoms, mats = [ground], [dict(type="matte", color=(0.7, 0.7, 0.65))]
    for k in range(12):
        v, t = scenes.sphere_mesh(rng.uniform((-8, 1, -8), (8, 4, 8)), rng.uniform(0.8, 2.0), 10, 10)
        geoms.append(dict(kind=0, vertices=v, indices=t))
        mats.append([dict(type="matte", color=tuple(rng.uniform(0.2, 0.9, 3))),
                     dict(type="pbr", color=tuple(rng.uniform(0.2, 0.9, 3)), metallic=0.8, roughness=0.3),
                     dict(type="glass", ior=1.5)][k % 3])
    geoms.append(dict(kind=1, vertices=rng.uniform((-10, 0.5, -10), (10, 3, 10), (40, 3)).astype(np.float32),
                      radii=rng.uniform(0.2, 0.6, 40).astype(np.float32)))
    mats.append(dict(type="pbr", color=(0.9, 0.6, 0.3), metallic=0.0, roughness=0.6))
    for k, g in enumerate(geoms):
        g["material"] = k
    gg = [gpu.Geom(g["kind"], g["vertices"], g.get("indices"), g.get("radii"), material=g["material"],
                   user_id=k) for k, g in enumerate(geoms)]
    world = gpu.World([gg], materials=scenes.product_materials(gpu, mats))

    def omat(m):
        if m["type"] == "matte":
            return dict(type="matte", color=m["color"])
        if m["type"] == "pbr":
            return dict(type="pbr", color=m["color"], metallic=m["metallic"], roughness=m["roughness"],
                        transmission=0.0, ior=1.45)
        return dict(type="pbr", color=(1, 1, 1), metallic=0.0, roughness=0.5, transmission=1.0, ior=m["ior"])

    ref = refcpu.Scene([dict(kind=g["kind"], vertices=g["vertices"], indices=g.get("indices"), radii=g.get("radii"),
                             material=omat(mats[g["material"]]), user_id=k) for k, g in enumerate(geoms)])
    return world, ref


CAM = dict(position=(-14.0, 9.0, -16.0), direction=(14.0, -7.0, 16.0), fovy=50.0)


@pytest.mark.parametrize("filter", ["linear", "point"])
def test_surface_render_under_map_bit_exact(gpu, filter):
    ge, re = _pair(gpu, _hdr(), filter=filter, direction=(1, 0.2, 0), up=(0, 1, 0), scale=0.8)
    world, ref = _surface_scene(gpu)
    W, H = 72, 48
    for spp, depth in ((4, 1), (4, 8)):
        gp = gpu.render_params(W, H, spp=spp, bg=(0.1, 0.2, 0.3, 1.0), max_depth=depth, max_diffuse=3, env=ge)
        rp = refcpu.render_params(W, H, spp=spp, bg=(0.1, 0.2, 0.3, 1.0), max_depth=depth, max_diffuse=3, env=re)
        out = gpu.render_surface(world, gpu.camera(**CAM), gp, with_ids=True)
        racc, rids, nr = ref.render(refcpu.camera(**CAM), rp, with_ids=True)
        g = out["accum"].numpy()
        exact = np.all(_bits(g) == _bits(racc), axis=-1).mean()
        print(f"{filter} spp {spp} depth {depth}: bit-exact {exact:.5f}, mean {racc[..., :3].mean():.4f}", flush=True)
        assert np.array_equal(_bits(g), _bits(racc))
        assert np.array_equal(out["ids"], rids)
        assert out["stats"]["raysTraced"] == nr
        # the map lights the scene and shows where rays miss
        assert racc[..., :3].mean() > 0.05


def test_batched_renderers_under_map(gpu):
    """The batched megakernel equals the oracle bit for bit per view; the
    wavefront renderer (float atomics per fragment) within 1e-4."""
    ge, re = _pair(gpu, _hdr(16, 48, seed=5), direction=(0.3, 0, 1), up=(0, 1, 0))
    world, ref = _surface_scene(gpu)
    cams = [CAM, dict(position=(10.0, 5.0, -12.0), direction=(-10.0, -4.0, 12.0), fovy=40.0)]
    ws, hs = [40, 33], [30, 27]
    views = gpu.Views([gpu.camera(**c) for c in cams], ws, hs)
    p = gpu.render_params(0, 0, spp=4, bg=(-1, -1, -1, -1), max_depth=6, max_diffuse=2, env=ge)
    mega = gpu.render_batch(world, views, p, gpu.BATCH_MEGAKERNEL)["accum"].numpy()
    wave = gpu.render_batch(world, views, p, gpu.BATCH_WAVEFRONT)["accum"].numpy()
    for v, c in enumerate(cams):
        racc, _, _ = ref.render(refcpu.camera(**c), refcpu.render_params(ws[v], hs[v], spp=4, bg=(-1, -1, -1, -1),
                                                                        max_depth=6, max_diffuse=2, env=re))
        assert np.array_equal(_bits(views.image(mega, v)), _bits(racc))
    rel = np.abs(wave - mega) / np.maximum(np.abs(mega), 1e-3)
    assert rel.max() <= 1e-4


def test_volume_render_under_map_bit_exact(gpu):
    f = scenes.structured_field(n=32, blobs=5, sigma=(3.0, 8.0), octaves=2, seed=0x5EED0111)
    tf = scenes.ramp_tf()
    gv = gpu.Volume(f, origin=(0, 0, 0), spacing=(1, 1, 1), cells_per_mc=8)
    gv.set_tf(tf, (0.0, 1.0), 0.4)
    rv = refcpu.Volume(f, origin=(0, 0, 0), spacing=(1, 1, 1), cells_per_mc=8)
    rv.set_tf(tf, (0.0, 1.0), 0.4)
    ge, re = _pair(gpu, _hdr(24, 40, seed=9), direction=(0, 0, 1), up=(0, 1, 0), scale=1.3)
    c = np.array([15.5, 15.5, 15.5])
    pos = c + np.array([-30.0, 14.0, -36.0])
    W, H, spp = 48, 36, 8
    gout = gv.render(gpu.camera(pos, c - pos, fovy=45), gpu.render_params(W, H, spp=spp, env=ge))
    racc, nrays, nsamp = rv.render(refcpu.camera(pos, c - pos, fovy=45), refcpu.render_params(W, H, spp=spp, env=re))
    assert np.array_equal(_bits(gout["accum"].numpy()), _bits(racc))
    assert gout["stats"]["samples"] == nsamp and gout["stats"]["raysTraced"] == nrays


def test_tet_mesh_render_under_map_bit_exact(gpu):
    V, S, T = scenes.tet_grid(8, seed=0x5EED0404)
    g = gpu.UMesh(V, S, T)
    r = refcpu.UMesh(V, S, T)
    tf = scenes.ramp_tf()
    g.set_tf(tf, (0.0, 1.0), 0.6)
    r.set_tf(tf, (0.0, 1.0), 0.6)
    ge, re = _pair(gpu, _hdr(16, 32, seed=10), direction=(1, 0, 0), up=(0, 1, 0))
    pos = (4 - 9.6, 4 + 6.4, 4 - 15.2)
    d = (4 - pos[0], 4 - pos[1], 4 - pos[2])
    W, H, spp = 40, 30, 4
    out = g.render(gpu.camera(pos, d, fovy=40.0), gpu.render_params(W, H, spp=spp, max_diffuse=1, env=ge),
                   volume=True, iso=0.5, iso_color=(0.8, 0.6, 0.4))
    ra, nr, ns = r.render(refcpu.camera(pos, d, fovy=40.0), refcpu.render_params(W, H, spp=spp, max_diffuse=1, env=re),
                          volume=True, iso=0.5, iso_color=(0.8, 0.6, 0.4))
    assert np.array_equal(_bits(out["accum"].numpy()), _bits(ra))
    assert out["stats"]["raysTraced"] == nr


def test_bad_texture_and_map_arguments_rejected(gpu):
    import ctypes as C
    from barney_amd import capi
    lib = capi.load()
    d = capi.bnamd_texture2d_desc()
    d.width, d.height, d.format = 4, 4, 99
    buf = np.zeros(64, np.float32)
    d.texels = buf.ctypes.data
    d.filterMode = 1
    h = C.c_void_p()
    assert lib.bnamd_texture2d_create(C.byref(d), None, C.byref(h)) != 0
    d.format, d.filterMode = 33, 7
    assert lib.bnamd_texture2d_create(C.byref(d), None, C.byref(h)) != 0
    d.filterMode, d.width = 1, 0
    assert lib.bnamd_texture2d_create(C.byref(d), None, C.byref(h)) != 0
    t = gpu.Texture2D(np.ones((4, 4, 4), np.float32))
    e = capi.bnamd_envmap_desc()
    e.direction[:] = [0, 0, 1]
    e.up[:] = [0, 0, 2]  # parallel to direction
    e.scale = 1.0
    assert lib.bnamd_envmap_create(t.handle, C.byref(e), None, C.byref(h)) != 0

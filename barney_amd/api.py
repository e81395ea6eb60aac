"""Python host layer over the C ABI: device arrays, worlds, volumes, renders.

This mirrors the object model barney's host core drives through rtcore
(rtc::Device buffers, Geom/Group/InstanceGroup builds, trace and compute
launches); every call goes to libbnamd.so.
"""
import ctypes as C

import numpy as np

from . import capi
from .capi import check

_initialized = False


def init(device=0):
    global _initialized
    lib = capi.load()
    if not _initialized:
        check(lib.bnamd_init(device))
        _initialized = True
    return lib


class DeviceArray:
    """A device allocation holding a copy of a numpy array."""

    def __init__(self, arr=None, nbytes=None, dtype=np.float32, shape=None):
        lib = init()
        if arr is not None:
            arr = np.ascontiguousarray(arr)
            self.dtype, self.shape, nbytes = arr.dtype, arr.shape, arr.nbytes
        else:
            self.dtype, self.shape = np.dtype(dtype), shape
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        check(lib.bnamd_device_malloc(C.byref(self.ptr), max(self.nbytes, 1)))
        if arr is not None and self.nbytes:
            check(lib.bnamd_memcpy_h2d(self.ptr, arr.ctypes.data, self.nbytes))

    @classmethod
    def empty(cls, shape, dtype=np.float32):
        dt = np.dtype(dtype)
        return cls(nbytes=int(np.prod(shape)) * dt.itemsize, dtype=dt, shape=tuple(shape))

    def zero(self):
        check(capi.load().bnamd_memset(self.ptr, 0, self.nbytes))
        return self

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            check(capi.load().bnamd_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes))
        return out

    @property
    def addr(self):
        return self.ptr.value

    def free(self):
        if self.ptr is not None and self.ptr.value:
            capi.load().bnamd_device_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


GEOM_TRIANGLES, GEOM_SPHERES, GEOM_CYLINDERS = 0, 1, 2
MAT_MATTE, MAT_PBR = 0, 1

IDENTITY = (1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0)


class Geom:
    """Host description of one geometry (triangles / spheres / cylinders)."""

    def __init__(self, kind, vertices, indices=None, radii=None, colors=None, default_radius=1.0,
                 material=0, user_id=0, normals=None):
        """normals: triangles only, one float3 per vertex (shading normals)."""
        self.kind = kind
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
        self.indices = None if indices is None else np.ascontiguousarray(indices, dtype=np.int32)
        self.radii = None if radii is None else np.ascontiguousarray(radii, dtype=np.float32)
        self.colors = None if colors is None else np.ascontiguousarray(colors, dtype=np.float32).reshape(-1, 3)
        self.normals = None if normals is None else np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
        self.default_radius = float(default_radius)
        self.material = int(material)
        self.user_id = int(user_id)

    @property
    def num_prims(self):
        if self.kind == GEOM_TRIANGLES:
            return len(self.indices.reshape(-1, 3))
        if self.kind == GEOM_CYLINDERS:
            return len(self.indices.reshape(-1, 2))
        return len(self.vertices)


def matte(color, opacity=1.0):
    m = capi.bnamd_material()
    m.type = MAT_MATTE
    m.color[:] = [float(c) for c in color]
    m.opacity = float(opacity)
    return m


class World:
    """Acceleration structure over one instance group (identity instances)."""

    def __init__(self, groups, materials=None, instances=None):
        lib = init()
        self._keep = []
        cgroups = (capi.bnamd_group * len(groups))()
        for gi, geoms in enumerate(groups):
            arr = (capi.bnamd_geom * max(len(geoms), 1))()
            for k, g in enumerate(geoms):
                dv = DeviceArray(g.vertices)
                di = DeviceArray(g.indices) if g.indices is not None else None
                dr = DeviceArray(g.radii) if g.radii is not None else None
                dc = DeviceArray(g.colors) if g.colors is not None else None
                dn = DeviceArray(g.normals) if getattr(g, "normals", None) is not None else None
                self._keep += [x for x in (dv, di, dr, dc, dn) if x is not None]
                e = arr[k]
                e.kind, e.numPrims, e.numVertices = g.kind, g.num_prims, len(g.vertices)
                e.materialID, e.userID = g.material, g.user_id
                e.vertices = dv.addr
                e.indices = di.addr if di else None
                e.radii = dr.addr if dr else None
                e.colors = dc.addr if dc else None
                e.normals = dn.addr if dn else None
                e.defaultRadius = g.default_radius
            cgroups[gi].geoms = C.cast(arr, C.POINTER(capi.bnamd_geom))
            cgroups[gi].numGeoms = len(geoms)
            self._keep.append(arr)
        if instances is None:
            instances = [(gi, 0, IDENTITY) for gi in range(len(groups))]
        cinst = (capi.bnamd_instance * max(len(instances), 1))()
        for k, (g, iid, xfm) in enumerate(instances):
            cinst[k].group, cinst[k].instID = g, iid
            cinst[k].xfm[:] = [float(x) for x in xfm]
        materials = materials or [matte((0.8, 0.8, 0.8))]
        cmats = (capi.bnamd_material * len(materials))(*materials)
        self.handle = C.c_void_p()
        check(lib.bnamd_world_build(cgroups, len(groups), cinst, len(instances), cmats, len(materials), None,
                                    C.byref(self.handle)))

    def validate(self):
        """Host-side structural check of the BVH8s (raises on a violation)."""
        check(capi.load().bnamd_world_validate(self.handle))

    def info(self):
        inf = capi.bnamd_world_info()
        check(capi.load().bnamd_world_info_get(self.handle, C.byref(inf)))
        return {"numPrims": inf.numPrims, "numNodes": inf.numNodes, "depth": inf.depth,
                "buildMs": inf.buildMs, "lower": list(inf.sceneLower), "upper": list(inf.sceneUpper)}

    def trace(self, rays):
        """rays: (n, 8) float32 [org, tmin, dir, tmax] -> structured hit array."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        n = len(rays)
        dr = DeviceArray(rays)
        dh = DeviceArray.empty((n, 10), np.float32)
        check(capi.load().bnamd_trace(self.handle, dr.addr, dh.addr, n, None))
        return hits_from_raw(dh.numpy())

    def destroy(self):
        if self.handle is not None and self.handle.value:
            capi.load().bnamd_world_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def hits_from_raw(raw):
    raw = np.ascontiguousarray(raw, dtype=np.float32)
    ints = raw.view(np.int32)
    return {"t": raw[:, 0].copy(), "u": raw[:, 1].copy(), "v": raw[:, 2].copy(),
            "primID": ints[:, 3].copy(), "geomID": ints[:, 4].copy(), "instID": ints[:, 5].copy(),
            "N": raw[:, 6:9].copy(), "instIndex": ints[:, 9].copy()}


def camera(position, direction, up=(0, 1, 0), fovy=60.0):
    c = capi.bnamd_camera()
    c.position[:] = [float(x) for x in position]
    c.direction[:] = [float(x) for x in direction]
    c.up[:] = [float(x) for x in up]
    c.fovy = float(fovy)
    return c


def render_params(width, height, spp=1, accum_id=0, ambient=1.0, bg=(0.0, 0.0, 0.0, 1.0), max_depth=10,
                  max_diffuse=1, russian_roulette=0, env=None, cut_plane=None, lights=None):
    """env: an EnvMap (HDRI environment light) or None (uniform ambient);
    cut_plane: (nx, ny, nz, d) or None (disabled); lights: Lights (quad,
    directional and point lights) or None."""
    p = capi.bnamd_render_params()
    p.width, p.height, p.accumID, p.pathsPerPixel = width, height, accum_id, spp
    p.ambientRadiance = float(ambient)
    p.bgColor[:] = [float(x) for x in bg]
    p.maxPathDepth, p.maxDiffuseBounces, p.russianRoulette = max_depth, max_diffuse, russian_roulette
    p.envMap = env.handle.value if env is not None else None
    p._env = env  # keeps the map alive as long as the parameters
    p.cutPlane[:] = [float(x) for x in (cut_plane if cut_plane is not None else (0.0, 0.0, 0.0, -1e30))]
    p.lights = lights.handle.value if lights is not None else None
    p._lights = lights  # keeps the lists alive as long as the parameters
    return p


class Lights:
    """Quad, directional and point lights on the device (World::DD's light
    lists). quads (n, 12) = corner, edge0, edge1, emission; dirs (n, 7) =
    direction, color, radiance; points (n, 8) = position, color, power,
    intensity (NaN: radiance = color * power / 4pi)."""

    def __init__(self, quads=(), dirs=(), points=()):
        lib = init()
        self.quads = np.ascontiguousarray(np.asarray(quads, np.float32).reshape(-1, 12))
        self.dirs = np.ascontiguousarray(np.asarray(dirs, np.float32).reshape(-1, 7))
        self.points = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 8))
        self.handle = C.c_void_p()
        check(lib.bnamd_lights_create(self.quads.ctypes.data, len(self.quads), self.dirs.ctypes.data,
                                      len(self.dirs), self.points.ctypes.data, len(self.points),
                                      C.byref(self.handle)))

    def __del__(self):
        if getattr(self, "handle", None) and self.handle.value:
            capi.load().bnamd_lights_destroy(self.handle)
            self.handle = C.c_void_p()


TEXEL_FORMATS = {"float4": 33, "float": 30, "uchar4": 1, "uchar": 0}
ADDRESS_MODES = {"wrap": 0, "clamp": 1, "border": 2, "mirror": 3}


class Texture2D:
    """A 2D texture (rtc::TextureData + rtc::Texture for images): texels
    (H, W, 4) / (H, W) float32 or uint8; filter 'linear' or 'point';
    address modes per axis ('wrap', 'clamp', 'border', 'mirror')."""

    def __init__(self, texels, filter="linear", address=("wrap", "clamp")):
        lib = init()
        t = np.ascontiguousarray(texels)
        if t.dtype != np.uint8:
            t = np.ascontiguousarray(t, np.float32)
        self.height, self.width = t.shape[:2]
        d = capi.bnamd_texture2d_desc()
        d.width, d.height = self.width, self.height
        d.format = TEXEL_FORMATS[("uchar" if t.dtype == np.uint8 else "float") + ("4" if t.ndim == 3 else "")]
        d.texels = t.ctypes.data
        d.filterMode = 1 if filter == "linear" else 0
        d.addressMode[:] = [ADDRESS_MODES[a] for a in address]
        self.handle = C.c_void_p()
        check(lib.bnamd_texture2d_create(C.byref(d), None, C.byref(self.handle)))

    def sample(self, uv):
        uv = DeviceArray(np.ascontiguousarray(uv, np.float32).reshape(-1, 2))
        out = DeviceArray.empty((uv.shape[0], 4), np.float32)
        check(capi.load().bnamd_texture2d_sample(self.handle, uv.addr, out.addr, uv.shape[0], None))
        return out.numpy()

    def __del__(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            capi.load().bnamd_texture2d_destroy(self.handle)
            self.handle = None


class EnvMap:
    """The HDRI environment light (EnvMapLight) over a Texture2D."""

    def __init__(self, texture, direction=(1, 0, 0), up=(0, 0, 1), scale=1.0):
        lib = init()
        self.texture = texture  # the map reads the texture's texels
        d = capi.bnamd_envmap_desc()
        d.direction[:] = [float(x) for x in direction]
        d.up[:] = [float(x) for x in up]
        d.scale = float(scale)
        self.handle = C.c_void_p()
        check(lib.bnamd_envmap_create(texture.handle, C.byref(d), None, C.byref(self.handle)))

    def cdfs(self):
        h, w = self.texture.height, self.texture.width
        y = np.empty(h, np.float32)
        x = np.empty((h, w), np.float32)
        check(capi.load().bnamd_envmap_get_cdfs(self.handle, y.ctypes.data, x.ctypes.data))
        return y, x

    def frame(self):
        w = np.empty(9, np.float32)
        l = np.empty(9, np.float32)
        check(capi.load().bnamd_envmap_get_frame(self.handle, w.ctypes.data, l.ctypes.data))
        return w.reshape(3, 3), l.reshape(3, 3)

    def eval(self, dirs):
        d = DeviceArray(np.ascontiguousarray(dirs, np.float32).reshape(-1, 3))
        n = d.shape[0]
        pdf = DeviceArray.empty((n,), np.float32)
        rad = DeviceArray.empty((n, 3), np.float32)
        check(capi.load().bnamd_envmap_eval(self.handle, d.addr, pdf.addr, rad.addr, n, None))
        return pdf.numpy(), rad.numpy()

    def sample(self, n, seed=0):
        out = DeviceArray.empty((n, 7), np.float32)
        check(capi.load().bnamd_envmap_sample(self.handle, seed, out.addr, n, None))
        o = out.numpy()
        return o[:, :3], o[:, 3:6], o[:, 6]

    def __del__(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            capi.load().bnamd_envmap_destroy(self.handle)
            self.handle = None


def render_surface(world, cam, params, with_ids=False, accum=None):
    n = params.width * params.height
    acc = accum if accum is not None else DeviceArray.empty((params.height, params.width, 4), np.float32)
    ids = DeviceArray.empty((params.height, params.width, 4), np.int32) if with_ids else None
    st = capi.bnamd_render_stats()
    check(capi.load().bnamd_render_surface(world.handle, C.byref(cam), C.byref(params), acc.addr,
                                           ids.addr if ids else None, None, C.byref(st)))
    out = {"accum": acc, "stats": {"raysTraced": st.raysTraced, "primaryRays": st.primaryRays, "ms": st.ms,
                                   "nodeVisits": st.nodeVisits, "primBytes": st.primBytes},
           "pixels": n}
    if ids is not None:
        out["ids"] = ids.numpy()
    return out


def render_mixed(world, volume, cam, params, with_ids=False, accum=None):
    """A scene of the world's surfaces and a structured Volume: each ray's
    closest surface hit, then delta tracking through the volume up to it."""
    n = params.width * params.height
    acc = accum if accum is not None else DeviceArray.empty((params.height, params.width, 4), np.float32)
    ids = DeviceArray.empty((params.height, params.width, 4), np.int32) if with_ids else None
    st = capi.bnamd_render_stats()
    ns = C.c_longlong(0)
    check(capi.load().bnamd_render_mixed(world.handle, volume.handle, C.byref(cam), C.byref(params), acc.addr,
                                         ids.addr if ids else None, None, C.byref(st), C.byref(ns)))
    out = {"accum": acc, "stats": {"raysTraced": st.raysTraced, "primaryRays": st.primaryRays, "ms": st.ms,
                                   "nodeVisits": st.nodeVisits, "primBytes": st.primBytes, "samples": ns.value},
           "pixels": n}
    if ids is not None:
        out["ids"] = ids.numpy()
    return out


BATCH_MEGAKERNEL, BATCH_WAVEFRONT, BATCH_AUTO = 0, 1, 2


class Views:
    """Several cameras of (possibly) different resolutions, rendered in one
    batch; pixels of view v are accum[offsets[v]:offsets[v+1]] (row-major)."""

    def __init__(self, cams, widths, heights):
        n = len(cams)
        self.cams = (capi.bnamd_camera * n)(*cams)
        self.widths = (C.c_int * n)(*[int(w) for w in widths])
        self.heights = (C.c_int * n)(*[int(h) for h in heights])
        self.c = capi.bnamd_views()
        self.c.numViews = n
        self.c.cameras = C.cast(self.cams, C.POINTER(capi.bnamd_camera))
        self.c.widths = C.cast(self.widths, C.POINTER(C.c_int))
        self.c.heights = C.cast(self.heights, C.POINTER(C.c_int))
        sizes = [int(w) * int(h) for w, h in zip(widths, heights)]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        self.num_pixels = int(self.offsets[-1])

    def image(self, accum_np, v):
        w, h = self.widths[v], self.heights[v]
        return accum_np[self.offsets[v]:self.offsets[v + 1]].reshape(h, w, 4)


def render_batch(world, views, params, mode=BATCH_WAVEFRONT, accum=None, with_ids=False, ids=None):
    """ids: a DeviceArray of int32 (num_pixels, 4) to write the ID channels
    into (with_ids=True allocates one); out["ids"] is its host copy."""
    acc = accum if accum is not None else DeviceArray.empty((views.num_pixels, 4), np.float32)
    if ids is None and with_ids:
        ids = DeviceArray.empty((views.num_pixels, 4), np.int32)
    st = capi.bnamd_batch_stats()
    check(capi.load().bnamd_render_batch(world.handle, C.byref(views.c), C.byref(params), acc.addr,
                                         ids.addr if ids is not None else None, int(mode), None, C.byref(st)))
    out = {"accum": acc, "stats": {"raysTraced": st.raysTraced, "primaryRays": st.primaryRays,
                                   "shadowRays": st.shadowRays, "ms": st.ms, "graphReplays": st.graphReplays,
                                   "nodeVisits": st.nodeVisits, "primBytes": st.primBytes}}
    if ids is not None:
        out["ids"] = ids.numpy()
    return out


def batch_autotune(world, views, params, path=None):
    check(capi.load().bnamd_batch_autotune(world.handle, C.byref(views.c), C.byref(params),
                                           None if path is None else path.encode()))


def pbr(color, metallic=0.0, roughness=0.5, transmission=0.0, ior=1.45, opacity=1.0):
    m = capi.bnamd_material()
    m.type = MAT_PBR
    m.color[:] = [float(c) for c in color]
    m.opacity = float(opacity)
    m.metallic, m.roughness, m.transmission, m.ior, m.specular = metallic, roughness, transmission, ior, 0.5
    return m


class Volume:
    """Structured scalar field + macrocell grid + transfer function."""

    def __init__(self, scalars, origin=(0, 0, 0), spacing=(1, 1, 1), cells_per_mc=8):
        lib = init()
        scalars = np.ascontiguousarray(scalars, dtype=np.float32)
        nz, ny, nx = scalars.shape
        self.dims = (nx, ny, nz)
        self.field = DeviceArray(scalars)
        d = capi.bnamd_structured_desc()
        d.scalars = self.field.addr
        d.dims[:] = [nx, ny, nz]
        d.gridOrigin[:] = [float(x) for x in origin]
        d.gridSpacing[:] = [float(x) for x in spacing]
        d.cellsPerMC = int(cells_per_mc)
        self.handle = C.c_void_p()
        check(lib.bnamd_volume_create(C.byref(d), C.byref(self.handle)))

    def set_tf(self, values, domain=(0.0, 1.0), base_density=1.0):
        values = np.ascontiguousarray(values, dtype=np.float32).reshape(-1, 4)
        check(capi.load().bnamd_volume_set_tf(self.handle, values.ctypes.data, len(values), float(domain[0]),
                                              float(domain[1]), float(base_density)))

    def mc_dims(self):
        d = (C.c_int * 3)()
        check(capi.load().bnamd_volume_mc_dims(self.handle, d))
        return tuple(d)

    def macrocells(self):
        mx, my, mz = self.mc_dims()
        ranges = np.empty((mz, my, mx, 2), np.float32)
        maj = np.empty((mz, my, mx), np.float32)
        check(capi.load().bnamd_volume_get_mcs(self.handle, ranges.ctypes.data, maj.ctypes.data))
        return ranges, maj

    def sample_rays(self, rays, seeds, exact_log=False):
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        n = len(rays)
        dr, ds = DeviceArray(rays), DeviceArray(seeds)
        do = DeviceArray.empty((n, 4), np.float32)
        check(capi.load().bnamd_volume_sample_rays(self.handle, dr.addr, ds.addr, do.addr, n, int(exact_log), None))
        return do.numpy()

    def iso_rays(self, rays, iso):
        """rays: [n, 8] (org, tmin, dir, tmax) -> [n, 4] (t or inf, normal)."""
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        dr = DeviceArray(r)
        do = DeviceArray.empty((len(r), 4), np.float32)
        check(capi.load().bnamd_volume_iso_rays(self.handle, dr.addr, float(iso), do.addr, len(r), None))
        return do.numpy()

    def render(self, cam, params, accum=None):
        acc = accum if accum is not None else DeviceArray.empty((params.height, params.width, 4), np.float32)
        st = capi.bnamd_volume_stats()
        check(capi.load().bnamd_volume_render(self.handle, C.byref(cam), C.byref(params), acc.addr, None,
                                              C.byref(st)))
        return {"accum": acc, "stats": {"samples": st.samples, "ddaSteps": st.ddaSteps,
                                        "raysTraced": st.raysTraced, "ms": st.ms,
                                        "zeroMajorantSteps": st.zeroMajorantSteps}}

    def destroy(self):
        if self.handle is not None and self.handle.value:
            capi.load().bnamd_volume_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class NanoVDB:
    """NanoVDB field (barney's "NanoVDB" scalar field): a NanoVDB grid buffer
    (bytes or a uint8 array, e.g. from barney_amd.nanovdb.from_dense) of
    float, Fp4, Fp8, Fp16 or FpN values; macrocell grid, transfer function,
    point samples, iso-surface rays and the delta-tracked render."""

    def __init__(self, grid, on_device=False, leaf_table=True):
        lib = init()
        g = np.frombuffer(bytes(grid), np.uint8) if not isinstance(grid, np.ndarray) else \
            np.ascontiguousarray(grid, np.uint8).reshape(-1)
        d = capi.bnamd_nanovdb_desc()
        d.noLeafTable = 0 if leaf_table else 1
        if on_device:
            dg = DeviceArray(g)
            d.data, d.bytes, d.dataOnDevice = dg.addr, g.size, 1
        else:
            d.data, d.bytes, d.dataOnDevice = g.ctypes.data, g.size, 0
        self.handle = C.c_void_p()
        check(lib.bnamd_nanovdb_create(C.byref(d), None, C.byref(self.handle)))

    def info(self):
        i = capi.bnamd_nanovdb_info()
        check(capi.load().bnamd_nanovdb_info_get(self.handle, C.byref(i)))
        return {"gridType": i.gridType, "indexLower": list(i.indexLower), "indexUpper": list(i.indexUpper),
                "mcDims": tuple(i.mcDims), "lower": list(i.worldLower), "upper": list(i.worldUpper),
                "mcOrigin": list(i.mcOrigin), "mcSpacing": list(i.mcSpacing), "voxelSize": list(i.voxelSize),
                "gridBytes": i.gridBytes, "buildMs": i.buildMs}

    def set_tf(self, values, domain=(0.0, 1.0), base_density=1.0):
        values = np.ascontiguousarray(values, dtype=np.float32).reshape(-1, 4)
        check(capi.load().bnamd_nanovdb_set_tf(self.handle, values.ctypes.data, len(values), float(domain[0]),
                                               float(domain[1]), float(base_density)))

    def macrocells(self):
        mx, my, mz = self.info()["mcDims"]
        ranges = np.empty((mz, my, mx, 2), np.float32)
        maj = np.empty((mz, my, mx), np.float32)
        check(capi.load().bnamd_nanovdb_get_mcs(self.handle, ranges.ctypes.data, maj.ctypes.data))
        return ranges, maj

    def sample(self, points):
        p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        dp = DeviceArray(p)
        do = DeviceArray.empty((len(p),), np.float32)
        check(capi.load().bnamd_nanovdb_sample_points(self.handle, dp.addr, do.addr, len(p), None))
        return do.numpy()

    def iso_rays(self, rays, iso):
        """rays: [n, 8] (org, tmin, dir, tmax) -> [n, 4] (t or inf, normal)."""
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        dr = DeviceArray(r)
        do = DeviceArray.empty((len(r), 4), np.float32)
        check(capi.load().bnamd_nanovdb_iso_rays(self.handle, dr.addr, float(iso), do.addr, len(r), None))
        return do.numpy()

    def render(self, cam, params, accum=None):
        acc = accum if accum is not None else DeviceArray.empty((params.height, params.width, 4), np.float32)
        st = capi.bnamd_volume_stats()
        check(capi.load().bnamd_nanovdb_render(self.handle, C.byref(cam), C.byref(params), acc.addr, None,
                                               C.byref(st)))
        return {"accum": acc, "stats": {"samples": st.samples, "ddaSteps": st.ddaSteps,
                                        "raysTraced": st.raysTraced, "ms": st.ms,
                                        "zeroMajorantSteps": st.zeroMajorantSteps}}

    def __del__(self):
        try:
            if self.handle is not None and self.handle.value:
                capi.load().bnamd_nanovdb_destroy(self.handle)
            self.handle = None
        except Exception:
            pass


class Amr:
    """Block-structured AMR volume (BlockStructuredField's layout): blocks
    (origins, dims int3; levels; offsets of their first scalar), per-level
    refinements, scalars; macrocell grid, transfer function, point samples
    and the delta-tracked render."""

    def __init__(self, origins, dims, levels, offsets, refinements, scalars):
        lib = init()
        self._o = np.ascontiguousarray(origins, dtype=np.int32).reshape(-1, 3)
        self._d = np.ascontiguousarray(dims, dtype=np.int32).reshape(-1, 3)
        self._l = np.ascontiguousarray(levels, dtype=np.int32).reshape(-1)
        self._off = np.ascontiguousarray(offsets, dtype=np.uint64).reshape(-1)
        self._r = np.ascontiguousarray(refinements, dtype=np.int32).reshape(-1)
        sc = np.ascontiguousarray(scalars, dtype=np.float32).reshape(-1)
        self.scalars = DeviceArray(sc)  # the field keeps the caller's device scalars
        d = capi.bnamd_amr_desc()
        d.numBlocks = len(self._o)
        d.origins, d.dims, d.levels = self._o.ctypes.data, self._d.ctypes.data, self._l.ctypes.data
        d.offsets, d.numLevels, d.refinements = self._off.ctypes.data, len(self._r), self._r.ctypes.data
        d.scalars, d.numScalars = self.scalars.addr, len(sc)
        self.handle = C.c_void_p()
        check(lib.bnamd_amr_create(C.byref(d), None, C.byref(self.handle)))

    def info(self):
        i = capi.bnamd_amr_info()
        check(capi.load().bnamd_amr_info_get(self.handle, C.byref(i)))
        return {"numBlocks": i.numBlocks, "mcDims": tuple(i.mcDims), "lookupDims": tuple(i.lookupDims),
                "lower": list(i.worldLower), "upper": list(i.worldUpper), "buildMs": i.buildMs}

    def set_tf(self, values, domain=(0.0, 1.0), base_density=1.0):
        values = np.ascontiguousarray(values, dtype=np.float32).reshape(-1, 4)
        check(capi.load().bnamd_amr_set_tf(self.handle, values.ctypes.data, len(values), float(domain[0]),
                                           float(domain[1]), float(base_density)))

    def macrocells(self):
        mx, my, mz = self.info()["mcDims"]
        ranges = np.empty((mz, my, mx, 2), np.float32)
        maj = np.empty((mz, my, mx), np.float32)
        check(capi.load().bnamd_amr_get_mcs(self.handle, ranges.ctypes.data, maj.ctypes.data))
        return ranges, maj

    def sample(self, points):
        p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        dp = DeviceArray(p)
        do = DeviceArray.empty((len(p),), np.float32)
        check(capi.load().bnamd_amr_sample_points(self.handle, dp.addr, do.addr, len(p), None))
        return do.numpy()

    def render(self, cam, params, accum=None):
        acc = accum if accum is not None else DeviceArray.empty((params.height, params.width, 4), np.float32)
        st = capi.bnamd_volume_stats()
        check(capi.load().bnamd_amr_render(self.handle, C.byref(cam), C.byref(params), acc.addr, None, C.byref(st)))
        return {"accum": acc, "stats": {"samples": st.samples, "ddaSteps": st.ddaSteps,
                                        "raysTraced": st.raysTraced, "ms": st.ms,
                                        "zeroMajorantSteps": st.zeroMajorantSteps}}

    def __del__(self):
        try:
            if self.handle is not None and self.handle.value:
                capi.load().bnamd_amr_destroy(self.handle)
            self.handle = None
        except Exception:
            pass


class UMesh:
    """Unstructured volume (per-vertex scalars): point-location BVH,
    macrocell grid, transfer function; volume and iso-surface render. Either
    tets (int4, positively oriented) or general cells = (indices, offsets,
    types) in UMeshField's layout (types: BNAMD_CELL_* / ANARI codes)."""

    def __init__(self, vertices, scalars, tets=None, cells=None):
        lib = init()
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
        sc = np.ascontiguousarray(scalars, dtype=np.float32).reshape(-1)
        if len(sc) != len(v):
            raise ValueError("one scalar per vertex expected")
        dv, ds = DeviceArray(v), DeviceArray(sc)
        d = capi.bnamd_umesh_desc()
        d.vertices, d.scalars, d.numVertices = dv.addr, ds.addr, len(v)
        bufs = [dv, ds]
        if cells is None:
            t = np.ascontiguousarray(tets, dtype=np.int32).reshape(-1, 4)
            dt = DeviceArray(t)
            bufs.append(dt)
            d.tets, d.numTets = dt.addr, len(t)
        else:
            ix = np.ascontiguousarray(cells[0], dtype=np.int32).reshape(-1)
            off = np.ascontiguousarray(cells[1], dtype=np.int32).reshape(-1)
            ty = np.ascontiguousarray(cells[2], dtype=np.uint8).reshape(-1)
            if len(off) != len(ty):
                raise ValueError("one offset and one type per cell expected")
            di, do, dty = DeviceArray(ix), DeviceArray(off), DeviceArray(ty)
            bufs += [di, do, dty]
            d.indices, d.numIndices = di.addr, len(ix)
            d.cellOffsets, d.cellTypes, d.numCells = do.addr, dty.addr, len(off)
        self.handle = C.c_void_p()
        try:
            check(lib.bnamd_umesh_create(C.byref(d), None, C.byref(self.handle)))
        finally:
            for b in bufs:
                b.free()

    def info(self):
        i = capi.bnamd_umesh_info()
        check(capi.load().bnamd_umesh_info_get(self.handle, C.byref(i)))
        return {"numTets": i.numTets, "numNodes": i.numNodes, "bvhDepth": i.bvhDepth, "mcDims": tuple(i.mcDims),
                "lower": list(i.worldLower), "upper": list(i.worldUpper), "buildMs": i.buildMs, "mcMs": i.mcMs}

    def set_tf(self, values, domain=(0.0, 1.0), base_density=1.0):
        values = np.ascontiguousarray(values, dtype=np.float32).reshape(-1, 4)
        check(capi.load().bnamd_umesh_set_tf(self.handle, values.ctypes.data, len(values), float(domain[0]),
                                             float(domain[1]), float(base_density)))

    def macrocells(self):
        mx, my, mz = self.info()["mcDims"]
        ranges = np.empty((mz, my, mx, 2), np.float32)
        maj = np.empty((mz, my, mx), np.float32)
        check(capi.load().bnamd_umesh_get_mcs(self.handle, ranges.ctypes.data, maj.ctypes.data))
        return ranges, maj

    def sample(self, points):
        p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        dp = DeviceArray(p)
        do = DeviceArray.empty((len(p),), np.float32)
        check(capi.load().bnamd_umesh_sample_points(self.handle, dp.addr, do.addr, len(p), None))
        return do.numpy()

    def iso_rays(self, rays, iso):
        """rays: [n, 8] (org, tmin, dir, tmax) -> [n, 4] (t or inf, normal)."""
        r = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 8)
        dr = DeviceArray(r)
        do = DeviceArray.empty((len(r), 4), np.float32)
        check(capi.load().bnamd_umesh_iso_rays(self.handle, dr.addr, float(iso), do.addr, len(r), None))
        return do.numpy()

    def render(self, cam, params, volume=True, iso=None, iso_color=(0.8, 0.8, 0.8), accum=None, count=False):
        """count: the instrumented kernel, which also reports point
        location's node visits, element tests and their bytes."""
        acc = accum if accum is not None else DeviceArray.empty((params.height, params.width, 4), np.float32)
        o = capi.bnamd_umesh_render_opts()
        o.volume = int(bool(volume))
        o.iso = int(iso is not None)
        o.isoValue = float(iso if iso is not None else 0.0)
        o.isoColor[:] = [float(c) for c in iso_color]
        o.countTraversal = int(bool(count))
        st = capi.bnamd_volume_stats()
        check(capi.load().bnamd_umesh_render(self.handle, C.byref(cam), C.byref(params), C.byref(o), acc.addr, None,
                                             C.byref(st)))
        return {"accum": acc, "stats": {"samples": st.samples, "ddaSteps": st.ddaSteps,
                                        "raysTraced": st.raysTraced, "ms": st.ms,
                                        "zeroMajorantSteps": st.zeroMajorantSteps, "nodeVisits": st.nodeVisits,
                                        "cellTests": st.cellTests, "cellBytes": st.cellBytes}}

    def destroy(self):
        if self.handle is not None and self.handle.value:
            capi.load().bnamd_umesh_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# ------------------------------------------------------------ frame-buffer readback
FB_FLOAT4, FB_UFIXED8_RGBA, FB_UFIXED8_RGBA_SRGB = 0, 1, 2
TILE_SIZE = 32


def fb_tile_descs(width, height, rank=0, size=1):
    """Device int32 [numActive, 2]: lower corners of the 32x32 tiles that
    device `rank` of `size` owns (global tiles rank, rank+size, ...)."""
    lib = init()
    n = C.c_int()
    check(lib.bnamd_fb_tile_descs(width, height, rank, size, None, C.byref(n), None))
    d = DeviceArray.empty((n.value, 2), np.int32)
    check(lib.bnamd_fb_tile_descs(width, height, rank, size, d.ptr, C.byref(n), None))
    return d


def _color_out(width, height, fmt, out):
    if out is not None:
        return out
    if fmt == FB_FLOAT4:
        return DeviceArray.empty((height, width, 4), np.float32)
    return DeviceArray.empty((height, width), np.uint32)


def fb_linearize(tiles, descs, width, height, accum_scale, fmt, normal=None, out=None):
    """tiles: DeviceArray of accumulation tiles (float accum[1024][4] then
    float normal[1024][3] per tile), descs from fb_tile_descs. Returns the
    linear color DeviceArray ([H, W, 4] float32 or [H, W] uint32); normal,
    if given, is a DeviceArray [H, W, 3] that receives the normals."""
    o = _color_out(width, height, fmt, out)
    check(init().bnamd_fb_linearize(tiles.ptr, descs.ptr, descs.shape[0], width, height, float(accum_scale), fmt,
                                    o.ptr, None if normal is None else normal.ptr, None))
    return o


def fb_linearize_aux(aux, descs, width, height, out=None):
    o = out if out is not None else DeviceArray.empty((height, width), np.uint32)
    check(init().bnamd_fb_linearize_aux(aux.ptr, descs.ptr, descs.shape[0], width, height, o.ptr, None))
    return o


def fb_resolve(accum, width, height, accum_scale, fmt, out=None):
    """Linear float4 accumulation (as the renderers write it) -> color."""
    o = _color_out(width, height, fmt, out)
    check(init().bnamd_fb_resolve(accum.ptr, width, height, float(accum_scale), fmt, o.ptr, None))
    return o


def fb_to_fixed8(rgba, width, height, srgb, out=None):
    o = out if out is not None else DeviceArray.empty((height, width), np.uint32)
    check(init().bnamd_fb_to_fixed8(rgba.ptr, width, height, int(bool(srgb)), o.ptr, None))
    return o

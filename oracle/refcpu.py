"""ctypes binding of the CPU oracle (oracle/librefcpu.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker. The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "librefcpu.so")


# the NanoVDB headers the reference vendors; present in the build container
# only (the golden-vector generator built on them is not needed elsewhere)
NANOVDB_INC = "/root/reference/3rdParty/nanovdb/include"


def build():
    r = subprocess.run(["make", "-C", HERE, "-s"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)
    if os.path.isdir(NANOVDB_INC):
        r = subprocess.run(["make", "-C", HERE, "-s", "_ref/nvdb_ref", f"NANOVDB_INC={NANOVDB_INC}"],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("oracle/_ref build failed:\n" + r.stdout + r.stderr)
    return LIB


class ref_geom(C.Structure):
    _fields_ = [("kind", C.c_int), ("numPrims", C.c_int), ("vertices", C.c_void_p), ("indices", C.c_void_p),
                ("radii", C.c_void_p), ("defaultRadius", C.c_float), ("colors", C.c_void_p),
                ("color", C.c_float * 3), ("userID", C.c_int), ("matType", C.c_int), ("metallic", C.c_float),
                ("roughness", C.c_float), ("transmission", C.c_float), ("ior", C.c_float),
                ("opacity", C.c_float), ("normals", C.c_void_p)]


class ref_hit(C.Structure):
    _fields_ = [("t", C.c_float), ("u", C.c_float), ("v", C.c_float), ("primID", C.c_int), ("geomID", C.c_int),
                ("instID", C.c_int), ("N", C.c_float * 3), ("instIndex", C.c_int)]


class ref_camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("direction", C.c_float * 3), ("up", C.c_float * 3),
                ("fovy", C.c_float)]


class ref_render_params(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("accumID", C.c_int), ("pathsPerPixel", C.c_int),
                ("ambientRadiance", C.c_float), ("bgColor", C.c_float * 4), ("maxPathDepth", C.c_int),
                ("maxDiffuseBounces", C.c_int), ("russianRoulette", C.c_int), ("envMap", C.c_void_p),
                ("cutPlane", C.c_float * 4), ("lights", C.c_void_p)]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    L = C.CDLL(LIB)
    P = C.c_void_p
    sig = {
        "ref_rng_seed": (C.c_uint64, [C.c_uint64, C.c_uint32, C.c_uint32]),
        "ref_rng_sequence": (None, [P, P, C.c_int]),
        "ref_random2_floats": (None, [C.c_uint32, C.c_uint32, P, C.c_int]),
        "ref_hash3": (C.c_uint64, [C.c_uint32, C.c_uint32, C.c_uint32]),
        "ref_fast_log": (C.c_float, [C.c_float]),
        "ref_half_round": (C.c_float, [C.c_float]),
        "ref_bsdf_sample": (None, [P, P, P, P, C.c_int, C.c_uint32, C.c_int, P]),
        "ref_bsdf_eval": (None, [P, P, P, P, P, P]),
        "ref_scene_create": (P, [P, C.c_int]),
        "ref_scene_destroy": (None, [P]),
        "ref_trace": (None, [P, P, P, C.c_int, C.c_int, C.c_int, P]),
        "ref_triangle_t_f64": (C.c_double, [P, P, P, P, P]),
        "ref_camera_ray": (None, [P, P, C.c_int, C.c_int, P, P, P]),
        "ref_render_surface": (C.c_longlong, [P, P, P, P, P, C.c_int, C.c_int, C.c_int]),
        "ref_render_mixed": (C.c_longlong, [P, P, P, P, P, P, C.c_int, C.c_int, C.c_int, P]),
        "ref_volume_create": (P, [P, P, P, P, C.c_int]),
        "ref_volume_destroy": (None, [P]),
        "ref_volume_set_tf": (None, [P, P, C.c_int, C.c_float, C.c_float, C.c_float]),
        "ref_volume_mc_dims": (None, [P, P]),
        "ref_volume_get_mcs": (None, [P, P, P]),
        "ref_volume_sample": (C.c_float, [P, P]),
        "ref_tf_map": (None, [P, C.c_float, P]),
        "ref_tf_majorant": (C.c_float, [P, C.c_float, C.c_float]),
        "ref_volume_sample_rays": (None, [P, P, P, P, C.c_int, C.c_int, C.c_int, P]),
        "ref_volume_iso_rays": (None, [P, P, C.c_float, P, C.c_int, C.c_int]),
        "ref_dda_cells": (C.c_int, [P, P, C.c_float, P, P, P, C.c_int]),
        "ref_render_volume": (C.c_longlong, [P, P, P, P, C.c_int, C.c_int, C.c_int, P]),
        "ref_world_create": (P, [P, C.c_int, P, P, P, C.c_int]),
        "ref_world_destroy": (None, [P]),
        "ref_world_trace": (None, [P, P, P, C.c_int, C.c_int]),
        "ref_umesh_create": (P, [P, P, C.c_int, P, C.c_int]),
        "ref_amr_create": (P, [P, P, P, P, C.c_int, P, C.c_int, P, C.c_ulonglong]),
        "ref_amr_destroy": (None, [P]),
        "ref_amr_set_tf": (None, [P, P, C.c_int, C.c_float, C.c_float, C.c_float]),
        "ref_amr_mc_dims": (None, [P, P]),
        "ref_amr_get_mcs": (None, [P, P, P]),
        "ref_amr_sample": (None, [P, P, P, C.c_int, C.c_int, C.c_int]),
        "ref_amr_render": (C.c_longlong, [P, P, P, P, C.c_int, C.c_int, C.c_int, P]),
        "ref_nvdb_create": (P, [P, C.c_ulonglong]),
        "ref_nvdb_destroy": (None, [P]),
        "ref_nvdb_set_tf": (None, [P, P, C.c_int, C.c_float, C.c_float, C.c_float]),
        "ref_nvdb_mc_dims": (None, [P, P]),
        "ref_nvdb_get_mcs": (None, [P, P, P]),
        "ref_nvdb_boxes": (None, [P, P, P, P]),
        "ref_nvdb_sample": (None, [P, P, P, C.c_int, C.c_int]),
        "ref_nvdb_values": (None, [P, P, P, C.c_int]),
        "ref_nvdb_render": (C.c_longlong, [P, P, P, P, C.c_int, C.c_int, C.c_int, P]),
        "ref_nvdb_iso_rays": (None, [P, P, C.c_float, P, C.c_int, C.c_int]),
        "ref_umesh_create_cells": (P, [P, P, C.c_int, P, C.c_int, P, P, C.c_int]),
        "ref_umesh_destroy": (None, [P]),
        "ref_umesh_set_tf": (None, [P, P, C.c_int, C.c_float, C.c_float, C.c_float]),
        "ref_umesh_mc_dims": (None, [P, P]),
        "ref_umesh_get_mcs": (None, [P, P, P]),
        "ref_umesh_sample": (None, [P, P, P, C.c_int, C.c_int]),
        "ref_fmath_eval": (None, [P, P, P, P, C.c_int]),
        "ref_fmath_eval_inv": (None, [P, P, P, P, C.c_int]),
        "ref_trace_filtered": (None, [P, P, P, P, P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
        "ref_envmap_create": (P, [P, C.c_int, C.c_int, C.c_int, C.c_int, P, P, P, C.c_float]),
        "ref_envmap_destroy": (None, [P]),
        "ref_lights_create": (P, [P, C.c_int, P, C.c_int, P, C.c_int]),
        "ref_lights_destroy": (None, [P]),
        "ref_envmap_get_cdfs": (None, [P, P, P]),
        "ref_envmap_frame": (None, [P, P, P]),
        "ref_envmap_tex2d": (None, [P, P, P, C.c_int]),
        "ref_envmap_eval": (None, [P, P, P, P, C.c_int]),
        "ref_envmap_sample": (None, [P, C.c_uint32, P, C.c_int]),
        "ref_umesh_sample_brute": (None, [P, P, P, C.c_int, C.c_int]),
        "ref_umesh_iso_rays": (None, [P, P, C.c_float, P, C.c_int, C.c_int]),
        "ref_umesh_render": (C.c_longlong, [P, P, P, C.c_int, C.c_int, C.c_float, P, P, C.c_int, C.c_int, C.c_int,
                                            P]),
        "ref_fb_make_8bit": (C.c_uint32, [C.c_float]),
        "ref_fb_linear_to_srgb": (C.c_float, [C.c_float]),
        "ref_fb_tile_descs": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, P]),
        "ref_fb_linearize": (None, [P, P, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, P, P]),
        "ref_fb_linearize_aux": (None, [P, P, C.c_int, C.c_int, C.c_int, P]),
        "ref_fb_resolve": (None, [P, C.c_size_t, C.c_float, C.c_int, P]),
        "ref_fb_to_fixed8": (None, [P, C.c_size_t, C.c_int, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data


class Scene:
    """Oracle scene over host geometry (one identity instance)."""

    def __init__(self, geoms):
        """geoms: list of dicts kind, vertices, indices, radii, colors,
        normals (triangles: per-vertex), default_radius, color, user_id and
        optionally material = dict(type "matte" | "pbr", metallic,
        roughness, transmission, ior, opacity)."""
        L = load()
        self._keep = []
        arr = (ref_geom * max(len(geoms), 1))()
        for k, g in enumerate(geoms):
            v = np.ascontiguousarray(g["vertices"], np.float32)
            i = None if g.get("indices") is None else np.ascontiguousarray(g["indices"], np.int32)
            r = None if g.get("radii") is None else np.ascontiguousarray(g["radii"], np.float32)
            c = None if g.get("colors") is None else np.ascontiguousarray(g["colors"], np.float32)
            nr = None if g.get("normals") is None else np.ascontiguousarray(g["normals"], np.float32)
            self._keep += [v, i, r, c, nr]
            kind = g["kind"]
            n = len(v.reshape(-1, 3)) if kind == 1 else len(i.reshape(-1, 3 if kind == 0 else 2))
            e = arr[k]
            e.kind, e.numPrims = kind, n
            e.vertices, e.indices, e.radii, e.colors = _ptr(v), _ptr(i), _ptr(r), _ptr(c)
            e.defaultRadius = float(g.get("default_radius", 1.0))
            e.color[:] = [float(x) for x in g.get("color", (0.8, 0.8, 0.8))]
            e.userID = int(g.get("user_id", 0))
            m = g.get("material") or {}
            e.matType = 1 if m.get("type", "matte") == "pbr" else 0
            e.metallic = float(m.get("metallic", 0.0))
            e.roughness = float(m.get("roughness", 0.5))
            e.transmission = float(m.get("transmission", 0.0))
            e.ior = float(m.get("ior", 1.45))
            e.opacity = float(m.get("opacity", 1.0))
            e.normals = _ptr(nr)
            if "color" in m:
                e.color[:] = [float(x) for x in m["color"]]
        self._arr = arr
        self.handle = L.ref_scene_create(arr, len(geoms))

    def trace_filtered(self, rays, params, seeds, brute=False, shadow=False, stochastic=True, threads=0):
        """Closest hit (or occlusion) with the renderer's candidate filter:
        params' cut plane, stochastic opacity drawn from seeds (uint64 per
        ray)."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        seeds = np.ascontiguousarray(seeds, np.uint64)
        n = len(rays)
        hits = (ref_hit * max(n, 1))()
        load().ref_trace_filtered(self.handle, C.byref(params), rays.ctypes.data, seeds.ctypes.data, hits, n,
                                  int(brute), int(shadow), int(stochastic), threads)
        raw = np.frombuffer(hits, dtype=np.float32, count=10 * n).reshape(n, 10)
        ints = raw.view(np.int32)
        return {"t": raw[:, 0].copy(), "primID": ints[:, 3].copy(), "geomID": ints[:, 4].copy(),
                "N": raw[:, 6:9].copy()}

    def trace(self, rays, brute=False, threads=0):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = len(rays)
        hits = (ref_hit * max(n, 1))()
        tests = C.c_longlong(0)
        load().ref_trace(self.handle, rays.ctypes.data, hits, n, int(brute), threads, C.byref(tests))
        raw = np.frombuffer(hits, dtype=np.float32, count=10 * n).reshape(n, 10)
        ints = raw.view(np.int32)
        return {"t": raw[:, 0].copy(), "u": raw[:, 1].copy(), "v": raw[:, 2].copy(), "primID": ints[:, 3].copy(),
                "geomID": ints[:, 4].copy(), "instID": ints[:, 5].copy(), "N": raw[:, 6:9].copy(),
                "instIndex": ints[:, 9].copy(), "primTests": tests.value}

    def render(self, cam, params, y0=0, y1=None, threads=0, with_ids=False, accum=None):
        y1 = params.height if y1 is None else y1
        acc = accum if accum is not None else np.zeros((params.height, params.width, 4), np.float32)
        ids = np.zeros((params.height, params.width, 4), np.int32) if with_ids else None
        n = load().ref_render_surface(self.handle, C.byref(cam), C.byref(params), acc.ctypes.data, _ptr(ids), y0,
                                      y1, threads)
        return acc, ids, n

    def render_mixed(self, volume, cam, params, y0=0, y1=None, threads=0, with_ids=False, accum=None):
        """The scene together with a structured Volume: each ray's closest
        surface hit first, then delta tracking through the volume up to it.
        Returns (accum, ids, rays traced, volume samples)."""
        y1 = params.height if y1 is None else y1
        acc = accum if accum is not None else np.zeros((params.height, params.width, 4), np.float32)
        ids = np.zeros((params.height, params.width, 4), np.int32) if with_ids else None
        ns = C.c_longlong(0)
        n = load().ref_render_mixed(self.handle, volume.handle, C.byref(cam), C.byref(params), acc.ctypes.data,
                                    _ptr(ids), y0, y1, threads, C.byref(ns))
        return acc, ids, n, ns.value

    def __del__(self):
        try:
            load().ref_scene_destroy(self.handle)
        except Exception:
            pass


class World:
    """Instanced oracle world: scenes placed by affine transforms (12 floats,
    columns vx, vy, vz, p)."""

    def __init__(self, scenes, instances):
        """scenes: list of Scene; instances: list of (scene index, instID,
        xfm[12])."""
        self._scenes = scenes
        g = (C.c_void_p * max(len(scenes), 1))(*[sc.handle for sc in scenes])
        ig = np.array([i[0] for i in instances], np.int32)
        ids = np.array([i[1] for i in instances], np.int32)
        x = np.ascontiguousarray(np.array([i[2] for i in instances], np.float32).reshape(-1, 12))
        self.handle = load().ref_world_create(g, len(scenes), ig.ctypes.data, ids.ctypes.data, x.ctypes.data,
                                              len(instances))

    def trace(self, rays, threads=0):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = len(rays)
        hits = (ref_hit * max(n, 1))()
        load().ref_world_trace(self.handle, rays.ctypes.data, hits, n, threads)
        raw = np.frombuffer(hits, dtype=np.float32, count=10 * n).reshape(n, 10)
        ints = raw.view(np.int32)
        return {"t": raw[:, 0].copy(), "u": raw[:, 1].copy(), "v": raw[:, 2].copy(), "primID": ints[:, 3].copy(),
                "geomID": ints[:, 4].copy(), "instID": ints[:, 5].copy(), "N": raw[:, 6:9].copy(),
                "instIndex": ints[:, 9].copy()}

    def __del__(self):
        try:
            load().ref_world_destroy(self.handle)
        except Exception:
            pass


class Volume:
    def __init__(self, scalars, origin=(0, 0, 0), spacing=(1, 1, 1), cells_per_mc=8):
        scalars = np.ascontiguousarray(scalars, np.float32)
        nz, ny, nx = scalars.shape
        dims = np.array([nx, ny, nz], np.int32)
        o = np.array(origin, np.float32)
        s = np.array(spacing, np.float32)
        self.handle = load().ref_volume_create(scalars.ctypes.data, dims.ctypes.data, o.ctypes.data,
                                               s.ctypes.data, int(cells_per_mc))

    def set_tf(self, values, domain=(0.0, 1.0), base_density=1.0):
        values = np.ascontiguousarray(values, np.float32).reshape(-1, 4)
        load().ref_volume_set_tf(self.handle, values.ctypes.data, len(values), float(domain[0]), float(domain[1]),
                                 float(base_density))

    def mc_dims(self):
        d = np.zeros(3, np.int32)
        load().ref_volume_mc_dims(self.handle, d.ctypes.data)
        return tuple(int(x) for x in d)

    def macrocells(self):
        mx, my, mz = self.mc_dims()
        ranges = np.empty((mz, my, mx, 2), np.float32)
        maj = np.empty((mz, my, mx), np.float32)
        load().ref_volume_get_mcs(self.handle, ranges.ctypes.data, maj.ctypes.data)
        return ranges, maj

    def sample(self, P):
        P = np.ascontiguousarray(P, np.float32)
        return load().ref_volume_sample(self.handle, P.ctypes.data)

    def tf_map(self, s):
        o = np.zeros(4, np.float32)
        load().ref_tf_map(self.handle, float(s), o.ctypes.data)
        return o

    def tf_majorant(self, lo, hi):
        return load().ref_tf_majorant(self.handle, float(lo), float(hi))

    def sample_rays(self, rays, seeds, exact_log=False, threads=0):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        seeds = np.ascontiguousarray(seeds, np.uint64)
        out = np.zeros((len(rays), 4), np.float32)
        ns = C.c_longlong(0)
        load().ref_volume_sample_rays(self.handle, rays.ctypes.data, seeds.ctypes.data, out.ctypes.data, len(rays),
                                      int(exact_log), threads, C.byref(ns))
        return out, ns.value

    def iso_rays(self, rays, iso, threads=0):
        """rays: [n, 8] (org, tmin, dir, tmax) -> [n, 4] (t or inf, normal)."""
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros((len(r), 4), np.float32)
        load().ref_volume_iso_rays(self.handle, r.ctypes.data, float(iso), out.ctypes.data, len(r), threads)
        return out

    def render(self, cam, params, y0=0, y1=None, threads=0, accum=None):
        y1 = params.height if y1 is None else y1
        acc = accum if accum is not None else np.zeros((params.height, params.width, 4), np.float32)
        ns = C.c_longlong(0)
        n = load().ref_render_volume(self.handle, C.byref(cam), C.byref(params), acc.ctypes.data, y0, y1, threads,
                                     C.byref(ns))
        return acc, n, ns.value

    def __del__(self):
        try:
            load().ref_volume_destroy(self.handle)
        except Exception:
            pass


class Amr:
    """Oracle block-structured AMR field: blocks (origins, dims int3; levels;
    offsets of their first scalar), per-level refinements, scalars; the
    macrocell grid, transfer function, point sampling and render."""

    def __init__(self, origins, dims, levels, offsets, refinements, scalars):
        self._o = np.ascontiguousarray(origins, np.int32).reshape(-1, 3)
        self._d = np.ascontiguousarray(dims, np.int32).reshape(-1, 3)
        self._l = np.ascontiguousarray(levels, np.int32).reshape(-1)
        self._off = np.ascontiguousarray(offsets, np.uint64).reshape(-1)
        self._r = np.ascontiguousarray(refinements, np.int32).reshape(-1)
        self._s = np.ascontiguousarray(scalars, np.float32).reshape(-1)
        self.handle = load().ref_amr_create(self._o.ctypes.data, self._d.ctypes.data, self._l.ctypes.data,
                                            self._off.ctypes.data, len(self._o), self._r.ctypes.data, len(self._r),
                                            self._s.ctypes.data, len(self._s))
        if not self.handle:
            raise ValueError("ref_amr_create: invalid block")

    def set_tf(self, values, domain=(0.0, 1.0), base_density=1.0):
        values = np.ascontiguousarray(values, np.float32).reshape(-1, 4)
        load().ref_amr_set_tf(self.handle, values.ctypes.data, len(values), float(domain[0]), float(domain[1]),
                              float(base_density))

    def mc_dims(self):
        d = np.zeros(3, np.int32)
        load().ref_amr_mc_dims(self.handle, d.ctypes.data)
        return tuple(int(x) for x in d)

    def macrocells(self):
        mx, my, mz = self.mc_dims()
        ranges = np.empty((mz, my, mx, 2), np.float32)
        maj = np.empty((mz, my, mx), np.float32)
        load().ref_amr_get_mcs(self.handle, ranges.ctypes.data, maj.ctypes.data)
        return ranges, maj

    def sample(self, points, brute=False, threads=0):
        P = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        out = np.empty(len(P), np.float32)
        load().ref_amr_sample(self.handle, P.ctypes.data, out.ctypes.data, len(P), int(bool(brute)), threads)
        return out

    def render(self, cam, params, y0=0, y1=None, threads=0, accum=None):
        y1 = params.height if y1 is None else y1
        acc = accum if accum is not None else np.zeros((params.height, params.width, 4), np.float32)
        ns = C.c_longlong(0)
        n = load().ref_amr_render(self.handle, C.byref(cam), C.byref(params), acc.ctypes.data, y0, y1, threads,
                                  C.byref(ns))
        return acc, n, ns.value

    def __del__(self):
        try:
            load().ref_amr_destroy(self.handle)
        except Exception:
            pass


class NanoVDB:
    """Oracle NanoVDB field over a grid buffer (bytes / uint8 array): the
    macrocell grid, transfer function, point sampling and render."""

    def __init__(self, grid):
        self._g = np.frombuffer(bytes(grid), np.uint8) if not isinstance(grid, np.ndarray) else \
            np.ascontiguousarray(grid, np.uint8)
        self.handle = load().ref_nvdb_create(self._g.ctypes.data, self._g.size)
        if not self.handle:
            raise ValueError("ref_nvdb_create: not a readable NanoVDB grid")

    def set_tf(self, values, domain=(0.0, 1.0), base_density=1.0):
        values = np.ascontiguousarray(values, np.float32).reshape(-1, 4)
        load().ref_nvdb_set_tf(self.handle, values.ctypes.data, len(values), float(domain[0]), float(domain[1]),
                               float(base_density))

    def mc_dims(self):
        d = np.zeros(3, np.int32)
        load().ref_nvdb_mc_dims(self.handle, d.ctypes.data)
        return tuple(int(x) for x in d)

    def boxes(self):
        w = np.zeros(6, np.float32)
        i = np.zeros(6, np.int32)
        m = np.zeros(6, np.float32)
        load().ref_nvdb_boxes(self.handle, w.ctypes.data, i.ctypes.data, m.ctypes.data)
        return dict(world_lo=w[:3], world_hi=w[3:], index_lo=i[:3], index_hi=i[3:], mc_origin=m[:3],
                    mc_spacing=m[3:])

    def macrocells(self):
        mx, my, mz = self.mc_dims()
        ranges = np.empty((mz, my, mx, 2), np.float32)
        maj = np.empty((mz, my, mx), np.float32)
        load().ref_nvdb_get_mcs(self.handle, ranges.ctypes.data, maj.ctypes.data)
        return ranges, maj

    def sample(self, points, threads=0):
        P = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        out = np.empty(len(P), np.float32)
        load().ref_nvdb_sample(self.handle, P.ctypes.data, out.ctypes.data, len(P), threads)
        return out

    def values(self, coords):
        c = np.ascontiguousarray(coords, np.int32).reshape(-1, 3)
        out = np.empty(len(c), np.float32)
        load().ref_nvdb_values(self.handle, c.ctypes.data, out.ctypes.data, len(c))
        return out

    def iso_rays(self, rays, iso, threads=0):
        """rays: [n, 8] (org, tmin, dir, tmax) -> [n, 4] (t or inf, normal)."""
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros((len(r), 4), np.float32)
        load().ref_nvdb_iso_rays(self.handle, r.ctypes.data, float(iso), out.ctypes.data, len(r), threads)
        return out

    def render(self, cam, params, y0=0, y1=None, threads=0, accum=None):
        y1 = params.height if y1 is None else y1
        acc = accum if accum is not None else np.zeros((params.height, params.width, 4), np.float32)
        ns = C.c_longlong(0)
        n = load().ref_nvdb_render(self.handle, C.byref(cam), C.byref(params), acc.ctypes.data, y0, y1, threads,
                                   C.byref(ns))
        return acc, n, ns.value

    def __del__(self):
        try:
            load().ref_nvdb_destroy(self.handle)
        except Exception:
            pass


class UMesh:
    """Oracle unstructured mesh: binned-SAH BVH over element boxes, macrocell
    grid (rasterTet<5> for tets, element boxes otherwise), transfer function,
    point sampling and render. Either tets (int4, positively oriented) or
    general cells: cells = (indices, offsets, types) in UMeshField's layout."""

    def __init__(self, vertices, scalars, tets=None, cells=None):
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        sc = np.ascontiguousarray(scalars, np.float32).reshape(-1)
        if cells is None:
            t = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
            self.handle = load().ref_umesh_create(v.ctypes.data, sc.ctypes.data, len(v), t.ctypes.data, len(t))
        else:
            ix = np.ascontiguousarray(cells[0], np.int32).reshape(-1)
            off = np.ascontiguousarray(cells[1], np.int32).reshape(-1)
            ty = np.ascontiguousarray(cells[2], np.uint8).reshape(-1)
            self.handle = load().ref_umesh_create_cells(v.ctypes.data, sc.ctypes.data, len(v), ix.ctypes.data,
                                                        len(ix), off.ctypes.data, ty.ctypes.data, len(off))
            if not self.handle:
                raise ValueError("ref_umesh_create_cells: unknown cell type or index out of range")

    def set_tf(self, values, domain=(0.0, 1.0), base_density=1.0):
        values = np.ascontiguousarray(values, np.float32).reshape(-1, 4)
        load().ref_umesh_set_tf(self.handle, values.ctypes.data, len(values), float(domain[0]), float(domain[1]),
                                float(base_density))

    def mc_dims(self):
        d = np.zeros(3, np.int32)
        load().ref_umesh_mc_dims(self.handle, d.ctypes.data)
        return tuple(int(x) for x in d)

    def macrocells(self):
        mx, my, mz = self.mc_dims()
        ranges = np.empty((mz, my, mx, 2), np.float32)
        maj = np.empty((mz, my, mx), np.float32)
        load().ref_umesh_get_mcs(self.handle, ranges.ctypes.data, maj.ctypes.data)
        return ranges, maj

    def sample(self, points, threads=0):
        p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        out = np.empty(len(p), np.float32)
        load().ref_umesh_sample(self.handle, p.ctypes.data, out.ctypes.data, len(p), threads)
        return out

    def sample_brute(self, points, threads=0):
        """Every element in index order; the first containing P gives the value."""
        p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
        out = np.empty(len(p), np.float32)
        load().ref_umesh_sample_brute(self.handle, p.ctypes.data, out.ctypes.data, len(p), threads)
        return out

    def iso_rays(self, rays, iso, threads=0):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.empty((len(r), 4), np.float32)
        load().ref_umesh_iso_rays(self.handle, r.ctypes.data, float(iso), out.ctypes.data, len(r), threads)
        return out

    def render(self, cam, params, volume=True, iso=None, iso_color=(0.8, 0.8, 0.8), y0=0, y1=None, threads=0,
               accum=None):
        y1 = params.height if y1 is None else y1
        acc = accum if accum is not None else np.zeros((params.height, params.width, 4), np.float32)
        col = np.ascontiguousarray(iso_color, np.float32)
        ns = C.c_longlong(0)
        n = load().ref_umesh_render(self.handle, C.byref(cam), C.byref(params), int(bool(volume)),
                                    int(iso is not None), float(iso if iso is not None else 0.0), col.ctypes.data,
                                    acc.ctypes.data, y0, y1, threads, C.byref(ns))
        return acc, n, ns.value

    def __del__(self):
        try:
            load().ref_umesh_destroy(self.handle)
        except Exception:
            pass


def camera(position, direction, up=(0, 1, 0), fovy=60.0):
    c = ref_camera()
    c.position[:] = [float(x) for x in position]
    c.direction[:] = [float(x) for x in direction]
    c.up[:] = [float(x) for x in up]
    c.fovy = float(fovy)
    return c


def render_params(width, height, spp=1, accum_id=0, ambient=1.0, bg=(0.0, 0.0, 0.0, 1.0), max_depth=10,
                  max_diffuse=1, russian_roulette=0, env=None, cut_plane=None, lights=None):
    """env: an EnvMap (HDRI environment light) or None; cut_plane: (nx, ny,
    nz, d) or None (disabled); lights: Lights (quad, directional and point
    lights) or None."""
    p = ref_render_params()
    p.width, p.height, p.accumID, p.pathsPerPixel = width, height, accum_id, spp
    p.ambientRadiance = float(ambient)
    p.bgColor[:] = [float(x) for x in bg]
    p.maxPathDepth, p.maxDiffuseBounces = max_depth, max_diffuse
    p.russianRoulette = russian_roulette
    p.envMap = env.handle if env is not None else None
    p._env = env  # keeps the map alive as long as the parameters
    p.cutPlane[:] = [float(x) for x in (cut_plane if cut_plane is not None else (0.0, 0.0, 0.0, -1e30))]
    p.lights = lights.handle if lights is not None else None
    p._lights = lights
    return p


def light_arrays(quads=(), dirs=(), points=()):
    """The three light lists as float32 arrays: quads (n, 12) = corner,
    edge0, edge1, emission; dirs (n, 7) = direction, color, radiance; points
    (n, 8) = position, color, power, intensity (NaN: color * power / 4pi)."""
    q = np.ascontiguousarray(np.asarray(quads, np.float32).reshape(-1, 12))
    d = np.ascontiguousarray(np.asarray(dirs, np.float32).reshape(-1, 7))
    pt = np.ascontiguousarray(np.asarray(points, np.float32).reshape(-1, 8))
    return q, d, pt


class Lights:
    """Quad, directional and point lights (light/QuadLight, DirLight,
    PointLight), lists as light_arrays takes them."""

    def __init__(self, quads=(), dirs=(), points=()):
        self.quads, self.dirs, self.points = light_arrays(quads, dirs, points)
        self.handle = load().ref_lights_create(self.quads.ctypes.data, len(self.quads), self.dirs.ctypes.data,
                                               len(self.dirs), self.points.ctypes.data, len(self.points))

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.ref_lights_destroy(self.handle)
            self.handle = None


def camera_ray(cam, params, ix, iy):
    """The primary ray of pixel (ix, iy): origin, direction, RNG state."""
    o = np.empty(3, np.float32)
    d = np.empty(3, np.float32)
    seed = np.zeros(1, np.uint64)
    load().ref_camera_ray(C.byref(cam), C.byref(params), ix, iy, o.ctypes.data, d.ctypes.data, seed.ctypes.data)
    return o, d, int(seed[0])


TEXEL_FORMATS = {"float4": 33, "float": 30, "uchar4": 1, "uchar": 0}
ADDRESS_MODES = {"wrap": 0, "clamp": 1, "border": 2, "mirror": 3}


class EnvMap:
    """The oracle's HDRI environment light: texels (H, W, 4) float32 /
    (H, W) float32 / (H, W, 4) or (H, W) uint8; filter 'linear' or 'point';
    address modes per axis."""

    def __init__(self, texels, direction=(1, 0, 0), up=(0, 0, 1), scale=1.0, filter="linear",
                 address=("wrap", "clamp")):
        t = np.ascontiguousarray(texels)
        self.h, self.w = t.shape[:2]
        if t.dtype == np.uint8:
            fmt = TEXEL_FORMATS["uchar4" if t.ndim == 3 else "uchar"]
        else:
            t = np.ascontiguousarray(t, np.float32)
            fmt = TEXEL_FORMATS["float4" if t.ndim == 3 else "float"]
        self._texels = t
        addr = (C.c_int * 2)(*[ADDRESS_MODES[a] for a in address])
        d = (C.c_float * 3)(*map(float, direction))
        u = (C.c_float * 3)(*map(float, up))
        self.handle = load().ref_envmap_create(t.ctypes.data, fmt, self.w, self.h,
                                               1 if filter == "linear" else 0, addr, d, u, float(scale))
        if not self.handle:
            raise ValueError("ref_envmap_create failed")

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.ref_envmap_destroy(self.handle)
            self.handle = None

    def cdfs(self):
        y = np.empty(self.h, np.float32)
        x = np.empty((self.h, self.w), np.float32)
        load().ref_envmap_get_cdfs(self.handle, y.ctypes.data, x.ctypes.data)
        return y, x

    def frame(self):
        w = np.empty(9, np.float32)
        l = np.empty(9, np.float32)
        load().ref_envmap_frame(self.handle, w.ctypes.data, l.ctypes.data)
        return w.reshape(3, 3), l.reshape(3, 3)  # rows = columns of the matrices

    def tex2d(self, uv):
        uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        out = np.empty((len(uv), 4), np.float32)
        load().ref_envmap_tex2d(self.handle, uv.ctypes.data, out.ctypes.data, len(uv))
        return out

    def eval(self, dirs):
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        pdf = np.empty(len(d), np.float32)
        rad = np.empty((len(d), 3), np.float32)
        load().ref_envmap_eval(self.handle, d.ctypes.data, pdf.ctypes.data, rad.ctypes.data, len(d))
        return pdf, rad

    def sample(self, n, seed=0):
        out = np.empty((n, 7), np.float32)
        load().ref_envmap_sample(self.handle, seed, out.ctypes.data, n)
        return out[:, :3], out[:, 3:6], out[:, 6]


def half_round(x):
    return load().ref_half_round(float(x))


def _material(m):
    g = ref_geom()
    g.matType = 1 if m.get("type", "matte") == "pbr" else 0
    g.metallic = float(m.get("metallic", 0.0))
    g.roughness = float(m.get("roughness", 0.5))
    g.transmission = float(m.get("transmission", 0.0))
    g.ior = float(m.get("ior", 1.45))
    return g


def bsdf_sample(material, color, Ng, wo, n, in_medium=False, seed=1):
    """n scatter samples: dict dir, f, pdf, type, offset, medium."""
    g = _material(material)
    col = np.ascontiguousarray(color, np.float32)
    N = np.ascontiguousarray(Ng, np.float32)
    w = np.ascontiguousarray(wo, np.float32)
    out = np.zeros((n, 10), np.float32)
    load().ref_bsdf_sample(C.byref(g), col.ctypes.data, N.ctypes.data, w.ctypes.data, int(in_medium), seed, n,
                           out.ctypes.data)
    return {"dir": out[:, 0:3], "f": out[:, 3:6], "pdf": out[:, 6], "type": out[:, 7].astype(int),
            "offset": out[:, 8], "medium": out[:, 9] > 0}


def bsdf_eval(material, color, Ng, wo, wi):
    """value[3], eval pdf, pdf() for one direction pair."""
    g = _material(material)
    a = [np.ascontiguousarray(x, np.float32) for x in (color, Ng, wo, wi)]
    out = np.zeros(5, np.float32)
    load().ref_bsdf_eval(C.byref(g), *[x.ctypes.data for x in a], out.ctypes.data)
    return out[:3].copy(), float(out[3]), float(out[4])


def dda_cells(org, dir, tmax, grid, max_cells=4096):
    org = np.ascontiguousarray(org, np.float32)
    dir = np.ascontiguousarray(dir, np.float32)
    g = np.ascontiguousarray(grid, np.int32)
    cells = np.zeros((max_cells, 3), np.int32)
    ts = np.zeros((max_cells, 2), np.float32)
    n = load().ref_dda_cells(org.ctypes.data, dir.ctypes.data, float(tmax), g.ctypes.data, cells.ctypes.data,
                             ts.ctypes.data, max_cells)
    return cells[:min(n, max_cells)], ts[:min(n, max_cells)], n


# ---------------------------------------------------------------- frame buffer
FB_FLOAT4, FB_RGBA8, FB_SRGB8 = 0, 1, 2
TILE = 32


def fb_make_8bit(f):
    return int(load().ref_fb_make_8bit(float(f)))


def fb_linear_to_srgb(x):
    return float(load().ref_fb_linear_to_srgb(float(x)))


def fb_tile_descs(width, height, rank=0, size=1):
    """int32 [numActive, 2] lower corners of the tiles device `rank` of
    `size` owns."""
    L = load()
    n = L.ref_fb_tile_descs(width, height, rank, size, None)
    out = np.zeros((max(n, 1), 2), np.int32)
    L.ref_fb_tile_descs(width, height, rank, size, out.ctypes.data)
    return out[:n]


def _out(width, height, fmt, fill):
    if fmt == FB_FLOAT4:
        return np.full((height, width, 4), fill, np.float32)
    return np.full((height, width), fill, np.uint32)


def fb_linearize(tiles, lower, width, height, scale, fmt, with_normal=False, fill=0):
    """tiles: float32 [numTiles, 1024*7] (accum float4 x 1024, normal float3
    x 1024); returns color [H, W, 4] float32 or [H, W] uint32 (and normal
    [H, W, 3]); pixels no tile covers keep `fill`."""
    tiles = np.ascontiguousarray(tiles, np.float32).reshape(-1, TILE * TILE * 7)
    lower = np.ascontiguousarray(lower, np.int32).reshape(-1, 2)
    out = _out(width, height, fmt, fill)
    nrm = np.full((height, width, 3), fill, np.float32) if with_normal else None
    load().ref_fb_linearize(tiles.ctypes.data, lower.ctypes.data, len(lower), width, height, float(scale), fmt,
                            out.ctypes.data, _ptr(nrm))
    return (out, nrm) if with_normal else out


def fb_linearize_aux(aux, lower, width, height, fill=0):
    aux = np.ascontiguousarray(aux).view(np.uint32).reshape(-1, TILE * TILE)
    lower = np.ascontiguousarray(lower, np.int32).reshape(-1, 2)
    out = np.full((height, width), fill, np.uint32)
    load().ref_fb_linearize_aux(aux.ctypes.data, lower.ctypes.data, len(lower), width, height, out.ctypes.data)
    return out


def fb_resolve(accum, scale, fmt):
    """accum float32 [..., 4] -> [..., 4] float32 or [...] uint32."""
    a = np.ascontiguousarray(accum, np.float32)
    shape = a.shape[:-1]
    out = np.zeros(shape + (4,), np.float32) if fmt == FB_FLOAT4 else np.zeros(shape, np.uint32)
    load().ref_fb_resolve(a.ctypes.data, a.size // 4, float(scale), fmt, out.ctypes.data)
    return out


def fb_to_fixed8(rgba, srgb):
    a = np.ascontiguousarray(rgba, np.float32)
    out = np.zeros(a.shape[:-1], np.uint32)
    load().ref_fb_to_fixed8(a.ctypes.data, a.size // 4, int(bool(srgb)), out.ctypes.data)
    return out


def fmath(x):
    """sin, cos and x^5 of the shared fmath.h sequences, elementwise."""
    x = np.ascontiguousarray(x, np.float32).ravel()
    s, c, p = (np.empty_like(x) for _ in range(3))
    load().ref_fmath_eval(x.ctypes.data, s.ctypes.data, c.ctypes.data, p.ctypes.data, len(x))
    return s, c, p


def fmath_inv(y, x):
    """atan2(y, x) and acos(x) of the shared fmath.h sequences, elementwise."""
    y = np.ascontiguousarray(y, np.float32).ravel()
    x = np.ascontiguousarray(x, np.float32).ravel()
    a, c = np.empty_like(x), np.empty_like(x)
    load().ref_fmath_eval_inv(y.ctypes.data, x.ctypes.data, a.ctypes.data, c.ctypes.data, len(x))
    return a, c

"""ctypes declarations for libbnamd.so (include/bnamd.h).

The library is the product: if it is missing or fails to load this module
raises, there is no fallback path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# BNAMD_LIBRARY selects another build of the same ABI (e.g. a debug build)
LIB_PATH = os.environ.get("BNAMD_LIBRARY", os.path.join(HERE, "libbnamd.so"))


class bnamd_geom(C.Structure):
    _fields_ = [("kind", C.c_int), ("numPrims", C.c_int), ("numVertices", C.c_int),
                ("materialID", C.c_int), ("vertices", C.c_void_p), ("indices", C.c_void_p),
                ("radii", C.c_void_p), ("colors", C.c_void_p), ("defaultRadius", C.c_float),
                ("userID", C.c_int), ("normals", C.c_void_p), ("programs", C.c_void_p)]


class bnamd_group(C.Structure):
    _fields_ = [("geoms", C.POINTER(bnamd_geom)), ("numGeoms", C.c_int)]


class bnamd_instance(C.Structure):
    _fields_ = [("group", C.c_int), ("instID", C.c_int), ("xfm", C.c_float * 12)]


class bnamd_material(C.Structure):
    _fields_ = [("type", C.c_int), ("color", C.c_float * 3), ("opacity", C.c_float),
                ("metallic", C.c_float), ("roughness", C.c_float), ("transmission", C.c_float),
                ("ior", C.c_float), ("specular", C.c_float)]


class bnamd_world_info(C.Structure):
    _fields_ = [("numPrims", C.c_longlong), ("numNodes", C.c_longlong), ("depth", C.c_int),
                ("buildMs", C.c_float), ("sceneLower", C.c_float * 3), ("sceneUpper", C.c_float * 3),
                ("numBlas", C.c_int)]


class bnamd_camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("direction", C.c_float * 3), ("up", C.c_float * 3),
                ("fovy", C.c_float)]


class bnamd_render_params(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("accumID", C.c_int),
                ("pathsPerPixel", C.c_int), ("ambientRadiance", C.c_float),
                ("bgColor", C.c_float * 4), ("maxPathDepth", C.c_int),
                ("maxDiffuseBounces", C.c_int), ("russianRoulette", C.c_int), ("envMap", C.c_void_p),
                ("cutPlane", C.c_float * 4), ("lights", C.c_void_p)]


class bnamd_texture2d_desc(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("format", C.c_int), ("texels", C.c_void_p),
                ("filterMode", C.c_int), ("addressMode", C.c_int * 2)]


class bnamd_envmap_desc(C.Structure):
    _fields_ = [("direction", C.c_float * 3), ("up", C.c_float * 3), ("scale", C.c_float)]


class bnamd_render_stats(C.Structure):
    _fields_ = [("raysTraced", C.c_longlong), ("primaryRays", C.c_longlong), ("ms", C.c_float),
                ("nodeVisits", C.c_longlong), ("primBytes", C.c_longlong)]


class bnamd_structured_desc(C.Structure):
    _fields_ = [("scalars", C.c_void_p), ("dims", C.c_int * 3), ("gridOrigin", C.c_float * 3),
                ("gridSpacing", C.c_float * 3), ("cellsPerMC", C.c_int)]


class bnamd_amr_desc(C.Structure):
    _fields_ = [("numBlocks", C.c_int), ("origins", C.c_void_p), ("dims", C.c_void_p), ("levels", C.c_void_p),
                ("offsets", C.c_void_p), ("numLevels", C.c_int), ("refinements", C.c_void_p),
                ("scalars", C.c_void_p), ("numScalars", C.c_ulonglong)]


class bnamd_amr_info(C.Structure):
    _fields_ = [("numBlocks", C.c_int), ("mcDims", C.c_int * 3), ("lookupDims", C.c_int * 3),
                ("worldLower", C.c_float * 3), ("worldUpper", C.c_float * 3), ("buildMs", C.c_float)]


class bnamd_nanovdb_desc(C.Structure):
    _fields_ = [("data", C.c_void_p), ("bytes", C.c_ulonglong), ("dataOnDevice", C.c_int), ("noLeafTable", C.c_int)]


class bnamd_nanovdb_info(C.Structure):
    _fields_ = [("gridType", C.c_int), ("indexLower", C.c_int * 3), ("indexUpper", C.c_int * 3),
                ("mcDims", C.c_int * 3), ("worldLower", C.c_float * 3), ("worldUpper", C.c_float * 3),
                ("mcOrigin", C.c_float * 3), ("mcSpacing", C.c_float * 3), ("voxelSize", C.c_float * 3),
                ("gridBytes", C.c_ulonglong), ("buildMs", C.c_float)]


class bnamd_volume_stats(C.Structure):
    _fields_ = [("samples", C.c_longlong), ("ddaSteps", C.c_longlong), ("raysTraced", C.c_longlong),
                ("ms", C.c_float), ("zeroMajorantSteps", C.c_longlong),
                ("nodeVisits", C.c_longlong), ("cellTests", C.c_longlong), ("cellBytes", C.c_longlong)]


class bnamd_umesh_desc(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("scalars", C.c_void_p), ("numVertices", C.c_int), ("tets", C.c_void_p),
                ("numTets", C.c_int), ("indices", C.c_void_p), ("numIndices", C.c_int), ("cellOffsets", C.c_void_p),
                ("cellTypes", C.c_void_p), ("numCells", C.c_int)]


class bnamd_umesh_info(C.Structure):
    _fields_ = [("numTets", C.c_int), ("numNodes", C.c_int), ("bvhDepth", C.c_int), ("mcDims", C.c_int * 3),
                ("worldLower", C.c_float * 3), ("worldUpper", C.c_float * 3), ("buildMs", C.c_float),
                ("mcMs", C.c_float)]


class bnamd_umesh_render_opts(C.Structure):
    _fields_ = [("volume", C.c_int), ("iso", C.c_int), ("isoValue", C.c_float), ("isoColor", C.c_float * 3),
                ("countTraversal", C.c_int)]


class bnamd_views(C.Structure):
    _fields_ = [("numViews", C.c_int), ("cameras", C.POINTER(bnamd_camera)), ("widths", C.POINTER(C.c_int)),
                ("heights", C.POINTER(C.c_int))]


class bnamd_batch_stats(C.Structure):
    _fields_ = [("raysTraced", C.c_longlong), ("primaryRays", C.c_longlong), ("shadowRays", C.c_longlong),
                ("ms", C.c_float), ("graphReplays", C.c_int), ("nodeVisits", C.c_longlong),
                ("primBytes", C.c_longlong)]


# name -> (restype, argtypes)
SIGNATURES = {
    "bnamd_init": (C.c_int, [C.c_int]),
    "bnamd_last_error": (C.c_char_p, []),
    "bnamd_device_malloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "bnamd_device_free": (C.c_int, [C.c_void_p]),
    "bnamd_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "bnamd_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "bnamd_memcpy_d2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "bnamd_memset": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t]),
    "bnamd_synchronize": (C.c_int, []),
    "bnamd_memcpy_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "bnamd_memset_async": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p]),
    "bnamd_host_malloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "bnamd_host_free": (C.c_int, [C.c_void_p]),
    "bnamd_get_device": (C.c_int, [C.POINTER(C.c_int)]),
    "bnamd_set_device": (C.c_int, [C.c_int]),
    "bnamd_copy_kernel": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "bnamd_ea_checkpoint_count": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]),
    "bnamd_ea_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "bnamd_ea_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p]),
    "bnamd_world_build": (C.c_int, [C.POINTER(bnamd_group), C.c_int, C.POINTER(bnamd_instance), C.c_int,
                                    C.POINTER(bnamd_material), C.c_int, C.c_void_p, C.POINTER(C.c_void_p)]),
    "bnamd_world_destroy": (C.c_int, [C.c_void_p]),
    "bnamd_blas_build": (C.c_int, [C.POINTER(bnamd_group), C.POINTER(bnamd_material), C.c_int, C.c_void_p,
                                   C.POINTER(C.c_void_p)]),
    "bnamd_blas_destroy": (C.c_int, [C.c_void_p]),
    "bnamd_world_build_instances": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.POINTER(bnamd_instance), C.c_int,
                                              C.POINTER(bnamd_material), C.c_int, C.c_void_p,
                                              C.POINTER(C.c_void_p)]),
    "bnamd_world_set_transforms": (C.c_int, [C.c_void_p, C.POINTER(C.c_float), C.c_int, C.c_void_p]),
    "bnamd_world_info_get": (C.c_int, [C.c_void_p, C.POINTER(bnamd_world_info)]),
    "bnamd_trace": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "bnamd_render_surface": (C.c_int, [C.c_void_p, C.POINTER(bnamd_camera), C.POINTER(bnamd_render_params),
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(bnamd_render_stats)]),
    "bnamd_render_mixed": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(bnamd_camera), C.POINTER(bnamd_render_params),
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(bnamd_render_stats),
                                     C.POINTER(C.c_longlong)]),
    "bnamd_volume_create": (C.c_int, [C.POINTER(bnamd_structured_desc), C.POINTER(C.c_void_p)]),
    "bnamd_volume_destroy": (C.c_int, [C.c_void_p]),
    "bnamd_volume_set_tf": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_float]),
    "bnamd_volume_mc_dims": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "bnamd_volume_get_mcs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "bnamd_volume_render": (C.c_int, [C.c_void_p, C.POINTER(bnamd_camera), C.POINTER(bnamd_render_params),
                                      C.c_void_p, C.c_void_p, C.POINTER(bnamd_volume_stats)]),
    "bnamd_amr_create": (C.c_int, [C.POINTER(bnamd_amr_desc), C.c_void_p, C.POINTER(C.c_void_p)]),
    "bnamd_amr_destroy": (C.c_int, [C.c_void_p]),
    "bnamd_amr_info_get": (C.c_int, [C.c_void_p, C.POINTER(bnamd_amr_info)]),
    "bnamd_amr_set_tf": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_float]),
    "bnamd_amr_get_mcs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "bnamd_amr_sample_points": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "bnamd_amr_render": (C.c_int, [C.c_void_p, C.POINTER(bnamd_camera), C.POINTER(bnamd_render_params),
                                   C.c_void_p, C.c_void_p, C.POINTER(bnamd_volume_stats)]),
    "bnamd_nanovdb_create": (C.c_int, [C.POINTER(bnamd_nanovdb_desc), C.c_void_p, C.POINTER(C.c_void_p)]),
    "bnamd_nanovdb_destroy": (C.c_int, [C.c_void_p]),
    "bnamd_nanovdb_info_get": (C.c_int, [C.c_void_p, C.POINTER(bnamd_nanovdb_info)]),
    "bnamd_nanovdb_set_tf": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_float]),
    "bnamd_nanovdb_get_mcs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "bnamd_nanovdb_sample_points": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "bnamd_nanovdb_render": (C.c_int, [C.c_void_p, C.POINTER(bnamd_camera), C.POINTER(bnamd_render_params),
                                       C.c_void_p, C.c_void_p, C.POINTER(bnamd_volume_stats)]),
    "bnamd_nanovdb_iso_rays": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.c_void_p, C.c_int, C.c_void_p]),
    "bnamd_volume_sample_rays": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                           C.c_void_p]),
    "bnamd_umesh_create": (C.c_int, [C.POINTER(bnamd_umesh_desc), C.c_void_p, C.POINTER(C.c_void_p)]),
    "bnamd_umesh_destroy": (C.c_int, [C.c_void_p]),
    "bnamd_umesh_info_get": (C.c_int, [C.c_void_p, C.POINTER(bnamd_umesh_info)]),
    "bnamd_umesh_set_tf": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_float]),
    "bnamd_umesh_get_mcs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "bnamd_umesh_sample_points": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "bnamd_volume_iso_rays": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.c_void_p, C.c_int, C.c_void_p]),
    "bnamd_umesh_iso_rays": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.c_void_p, C.c_int, C.c_void_p]),
    "bnamd_umesh_render": (C.c_int, [C.c_void_p, C.POINTER(bnamd_camera), C.POINTER(bnamd_render_params),
                                     C.POINTER(bnamd_umesh_render_opts), C.c_void_p, C.c_void_p,
                                     C.POINTER(bnamd_volume_stats)]),
    "bnamd_umesh_render_noalloc": (C.c_int, [C.c_void_p, C.POINTER(bnamd_camera), C.POINTER(bnamd_render_params),
                                             C.POINTER(bnamd_umesh_render_opts), C.c_void_p, C.c_void_p,
                                             C.POINTER(bnamd_volume_stats)]),
    "bnamd_render_batch": (C.c_int, [C.c_void_p, C.POINTER(bnamd_views), C.POINTER(bnamd_render_params),
                                     C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.POINTER(bnamd_batch_stats)]),
    "bnamd_batch_autotune": (C.c_int, [C.c_void_p, C.POINTER(bnamd_views), C.POINTER(bnamd_render_params),
                                       C.c_char_p]),
    "bnamd_fb_tile_descs": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_int),
                                      C.c_void_p]),
    "bnamd_fb_linearize": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int,
                                     C.c_void_p, C.c_void_p, C.c_void_p]),
    "bnamd_fb_linearize_aux": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                         C.c_void_p]),
    "bnamd_fb_resolve": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_int, C.c_void_p, C.c_void_p]),
    "bnamd_fb_to_fixed8": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "bnamd_world_validate": (C.c_int, [C.c_void_p]),
    "bnamd_world_device": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "bnamd_texture2d_create": (C.c_int, [C.POINTER(bnamd_texture2d_desc), C.c_void_p, C.POINTER(C.c_void_p)]),
    "bnamd_texture2d_destroy": (C.c_int, [C.c_void_p]),
    "bnamd_texture2d_sample": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "bnamd_envmap_create": (C.c_int, [C.c_void_p, C.POINTER(bnamd_envmap_desc), C.c_void_p,
                                      C.POINTER(C.c_void_p)]),
    "bnamd_envmap_destroy": (C.c_int, [C.c_void_p]),
    "bnamd_texture2d_device_view": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_int),
                                              C.POINTER(C.c_int)]),
    "bnamd_lights_create": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                      C.POINTER(C.c_void_p)]),
    "bnamd_lights_destroy": (C.c_int, [C.c_void_p]),
    "bnamd_envmap_get_cdfs": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "bnamd_envmap_get_frame": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "bnamd_envmap_eval": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "bnamd_envmap_sample": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_int, C.c_void_p]),
}

_lib = None


class BnamdError(RuntimeError):
    pass


def load():
    """Loads libbnamd.so and declares every entry point; raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BnamdError(f"{LIB_PATH} is not built (run python -m barney_amd.build)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status):
    if status != 0:
        raise BnamdError(load().bnamd_last_error().decode())

"""ctypes mirror of include/delta_ray.h and the loader of the HIP library.

The library is built in-tree by ``__graft_entry__.build()`` into
``delta_ray_amd/libdelta_ray_amd.so``.  There is no CPU fallback: when the
library (or a GPU) is missing, ``load_library`` raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdelta_ray_amd.so")
LTC_PATH = os.path.join(_HERE, "data", "ltc_tabM.f32")

DR_OK = 0
DR_ERROR_INVALID_ARGUMENT = 1
DR_ERROR_OUT_OF_MEMORY = 2
DR_ERROR_DEVICE = 3
DR_ERROR_UNSUPPORTED = 4

DR_LTC_FIT_CONSISTENT = 1

DR_EDGES_FILTERED = 0
DR_EDGES_ALL_CANDIDATES = 1

DR_QUIRK_JITTER_OVER_SPP = 1 << 0
DR_QUIRK_TEXEL_SWAP = 1 << 1
DR_QUIRK_NORMAL_DN0 = 1 << 2
DR_QUIRK_DP_LAST = 1 << 3
DR_QUIRK_PROJECT_SCALE = 1 << 4
DR_QUIRK_EDGE_BSDF_SUM = 1 << 5
DR_QUIRK_EDGE_HIT_JACOBIAN = 1 << 6
DR_QUIRK_CONE_PMF = 1 << 7
DR_SKIP_INTERIOR = 1 << 8
DR_SKIP_SECONDARY_EDGES = 1 << 9
DR_SKIP_PRIMARY_EDGES = 1 << 10
DR_WRITE_DX_IMAGE = 1 << 11
DR_LEARNED_EDGES = 1 << 12
DR_EDGE_STATS = 1 << 13
DR_QUIRK_CLIP_DERIVATIVE = 1 << 14
DR_QUIRK_EDGE_EMITTER_FULL = 1 << 15
DR_FLAGS_PARITY = (DR_QUIRK_JITTER_OVER_SPP | DR_QUIRK_TEXEL_SWAP |
                   DR_QUIRK_NORMAL_DN0 | DR_QUIRK_DP_LAST | DR_QUIRK_PROJECT_SCALE |
                   DR_QUIRK_EDGE_BSDF_SUM | DR_QUIRK_EDGE_HIT_JACOBIAN | DR_QUIRK_CONE_PMF |
                   DR_QUIRK_CLIP_DERIVATIVE | DR_QUIRK_EDGE_EMITTER_FULL)

_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)


class dr_camera(ctypes.Structure):
    _fields_ = [("cam_to_world", ctypes.c_float * 16),
                ("world_to_cam", ctypes.c_float * 16),
                ("sample_to_cam", ctypes.c_float * 16),
                ("cam_to_sample", ctypes.c_float * 16),
                ("fov_factor", ctypes.c_float),
                ("aspect_ratio", ctypes.c_float),
                ("clip_near", ctypes.c_float),
                ("fisheye", ctypes.c_int)]


class dr_texture(ctypes.Structure):
    _fields_ = [("data", _fp), ("height", ctypes.c_int), ("width", ctypes.c_int),
                ("channels", ctypes.c_int)]


class dr_material(ctypes.Structure):
    _fields_ = [("diffuse_reflectance", dr_texture),
                ("specular_reflectance", dr_texture),
                ("roughness", dr_texture),
                ("diffuse_uv_scale", ctypes.c_float * 2),
                ("specular_uv_scale", ctypes.c_float * 2),
                ("roughness_uv_scale", ctypes.c_float * 2),
                ("two_sided", ctypes.c_int)]


class dr_shape(ctypes.Structure):
    _fields_ = [("vertices", _fp), ("num_vertices", ctypes.c_int),
                ("indices", _ip), ("num_triangles", ctypes.c_int),
                ("uvs", _fp), ("normals", _fp),
                ("material_id", ctypes.c_int), ("light_id", ctypes.c_int)]


class dr_light(ctypes.Structure):
    _fields_ = [("shape_id", ctypes.c_int), ("intensity", ctypes.c_float * 3)]


class dr_scene_desc(ctypes.Structure):
    _fields_ = [("camera", dr_camera),
                ("shapes", ctypes.POINTER(dr_shape)), ("num_shapes", ctypes.c_int),
                ("materials", ctypes.POINTER(dr_material)), ("num_materials", ctypes.c_int),
                ("lights", ctypes.POINTER(dr_light)), ("num_lights", ctypes.c_int)]


class dr_render_options(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int),
                ("num_samples", ctypes.c_int), ("max_bounces", ctypes.c_int),
                ("seed", ctypes.c_uint64), ("flags", ctypes.c_uint32),
                ("num_primary_edge_samples", ctypes.c_int64)]


class dr_job(ctypes.Structure):
    _fields_ = [("scene", ctypes.c_void_p), ("view", ctypes.c_void_p), ("opts", dr_render_options),
                ("d_image", ctypes.c_void_p), ("image", ctypes.c_void_p), ("grads", ctypes.c_void_p),
                ("dx_image", ctypes.c_void_p)]


class dr_tuning(ctypes.Structure):
    _fields_ = [("pass_samples", ctypes.c_int64), ("min_threads", ctypes.c_int64), ("select_refill", ctypes.c_int32),
                ("best_ms", ctypes.c_double), ("candidates", ctypes.c_int)]


DR_BATCH_GRAPH = 1
DR_MLP_FEATURES = 24
DR_MLP_PARAMS = 5825

DR_STATS_COUNTERS = 1
DR_STATS_TIMING = 2


KERNELS = ["forward", "path", "primary", "select", "edge", "dp", "shade"]


class dr_stats(ctypes.Structure):
    _fields_ = [("rays", ctypes.c_uint64), ("nodes_visited", ctypes.c_uint64),
                ("tris_visited", ctypes.c_uint64), ("secondary_edges", ctypes.c_uint64),
                ("primary_edge_samples", ctypes.c_uint64),
                ("edge_nodes_visited", ctypes.c_uint64), ("edges_tested", ctypes.c_uint64),
                ("stack_overflows", ctypes.c_uint64),
                ("kernel_ms", ctypes.c_double * 7), ("kernel_launches", ctypes.c_uint64 * 7)]


class dr_scene_info(ctypes.Structure):
    _fields_ = [("num_triangles", ctypes.c_int), ("num_vertices", ctypes.c_int),
                ("bvh_nodes", ctypes.c_int), ("num_edges", ctypes.c_int),
                ("num_primary_edges", ctypes.c_int),
                ("silhouette_tree_nodes", ctypes.c_int),
                ("normal_tree_nodes", ctypes.c_int), ("build_ms", ctypes.c_double),
                ("refits", ctypes.c_int), ("rebuilds", ctypes.c_int), ("edges_switched_off", ctypes.c_int),
                ("bvh_depth", ctypes.c_int), ("edge_tree_depth", ctypes.c_int)]


class dr_shape_update(ctypes.Structure):
    _fields_ = [("shape_id", ctypes.c_int), ("vertices", ctypes.c_void_p), ("normals", ctypes.c_void_p)]


# entry points declared in include/delta_ray.h: (name, restype, argtypes)
class dr_mesh_data(ctypes.Structure):
    _fields_ = [("num_vertices", ctypes.c_int), ("num_triangles", ctypes.c_int),
                ("vertices", _fp), ("normals", _fp), ("uvs", _fp), ("colors", _fp),
                ("indices", ctypes.POINTER(ctypes.c_int))]


ENTRY_POINTS = [
    ("dr_scene_create", ctypes.c_int, [ctypes.POINTER(dr_scene_desc), ctypes.POINTER(ctypes.c_void_p)]),
    ("dr_scene_destroy", None, [ctypes.c_void_p]),
    ("dr_scene_set_camera", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(dr_camera)]),
    ("dr_scene_update_vertices", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(dr_shape_update), ctypes.c_int,
                                                ctypes.c_void_p]),
    ("dr_scene_set_edge_policy", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    ("dr_scene_update_materials", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(dr_material), ctypes.c_int]),
    ("dr_scene_update_lights", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(dr_light), ctypes.c_int]),
    ("dr_scene_grad_size", ctypes.c_size_t, [ctypes.c_void_p]),
    ("dr_render", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(dr_render_options), ctypes.c_void_p,
                                 ctypes.c_void_p]),
    ("dr_render_adjoint", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(dr_render_options),
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p]),
    ("dr_render_adjoint_noalloc", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(dr_render_options),
                                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                 ctypes.c_void_p, ctypes.c_void_p]),
    ("dr_intersect", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                    ctypes.c_int, ctypes.c_void_p]),
    ("dr_scene_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(dr_stats)]),
    ("dr_scene_get_info", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(dr_scene_info)]),
    ("dr_view_create", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(dr_camera), ctypes.POINTER(ctypes.c_void_p)]),
    ("dr_view_destroy", None, [ctypes.c_void_p]),
    ("dr_batch_create", ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p)]),
    ("dr_batch_destroy", None, [ctypes.c_void_p]),
    ("dr_render_batch", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(dr_job), ctypes.c_int, ctypes.c_int,
                                       ctypes.c_void_p]),
    ("dr_batch_autotune", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(dr_job), ctypes.c_int, ctypes.c_void_p,
                                         ctypes.POINTER(dr_tuning)]),
    ("dr_batch_set_tuning", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(dr_tuning)]),
    ("dr_scene_edge_count", ctypes.c_int64, [ctypes.c_void_p]),
    ("dr_scene_edge_features", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("dr_scene_set_edge_weights", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("dr_scene_edge_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    ("dr_scene_set_stat_shapes", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]),
    ("dr_scene_build_bvh_device", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("dr_mlp_create", ctypes.c_int, [ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
    ("dr_mlp_destroy", None, [ctypes.c_void_p]),
    ("dr_mlp_params", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("dr_mlp_infer", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                    ctypes.c_void_p]),
    ("dr_mlp_train", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_uint64,
                                    ctypes.POINTER(ctypes.c_float), ctypes.c_void_p]),
    ("dr_test_tree_stack", ctypes.c_int, [ctypes.c_int64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int64),
                                          ctypes.c_void_p]),
    ("dr_test_line_integral", ctypes.c_int, [ctypes.c_int64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double),
                                             ctypes.c_void_p]),
    ("dr_load_serialized", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(dr_mesh_data)]),
    ("dr_mesh_data_free", None, [ctypes.POINTER(dr_mesh_data)]),
    ("dr_ltc_fit", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp,
                                  ctypes.c_int, ctypes.c_int, _fp, _fp]),
    ("dr_last_error", ctypes.c_char_p, []),
]

_lib = None


def load_library(path=LIB_PATH):
    """Load the HIP library; raise if it is not built."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("delta_ray_amd: %s is not built (run __graft_entry__.build())" % path)
    lib = ctypes.CDLL(path)
    for name, res, args in ENTRY_POINTS:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib


def check(lib, status, what):
    if status != DR_OK:
        msg = lib.dr_last_error()
        raise RuntimeError("%s failed (%d): %s" % (what, status, msg.decode() if msg else "?"))


def fptr(a):
    return a.ctypes.data_as(_fp) if a is not None else ctypes.cast(None, _fp)


def iptr(a):
    return a.ctypes.data_as(_ip) if a is not None else ctypes.cast(None, _ip)


def f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))

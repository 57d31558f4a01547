"""ctypes access to the CPU oracle (oracle/liboracle.so) for the tests and
the CPU-baseline leg of bench.py.  The oracle is a checker: nothing in the
product imports this module."""
import ctypes
import os
import subprocess

import numpy as np

from delta_ray_amd import _abi
from delta_ray_amd import delta_ray as dr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# DR_ORACLE_SO: another build of the oracle (the planted-mistake builds of
# tests/test_oracle_mutants.py)
ORACLE_SO = os.environ.get("DR_ORACLE_SO") or os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO) and not os.environ.get("DR_ORACLE_SO"):
        build()
    lib = ctypes.CDLL(ORACLE_SO)
    lib.or_load_ltc.restype = ctypes.c_int
    lib.or_load_ltc.argtypes = [ctypes.c_char_p]
    lib.or_scene_create.restype = ctypes.c_void_p
    lib.or_scene_create.argtypes = [ctypes.POINTER(_abi.dr_scene_desc)]
    lib.or_scene_destroy.argtypes = [ctypes.c_void_p]
    lib.or_scene_use_bvh.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.or_last_render_rays.restype = ctypes.c_uint64
    lib.or_last_render_rays.argtypes = [ctypes.c_void_p]
    lib.or_scene_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    lib.or_grad_size.restype = ctypes.c_size_t
    lib.or_grad_size.argtypes = [ctypes.c_void_p]
    lib.or_render.restype = ctypes.c_int
    lib.or_render.argtypes = [ctypes.c_void_p, ctypes.POINTER(_abi.dr_render_options), ctypes.c_void_p,
                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    lib.or_intersect.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                 ctypes.c_int]
    if lib.or_load_ltc(_abi.LTC_PATH.encode()) != 0:
        raise RuntimeError("oracle: cannot load LTC table")
    _lib = lib
    return lib


class OracleScene:
    def __init__(self, camera, shapes, materials, lights):
        self.lib = load()
        self.sd = dr.SceneDesc(camera, shapes, materials, lights)
        self.h = self.lib.or_scene_create(ctypes.byref(self.sd.desc))
        self.grad_size = int(self.lib.or_grad_size(self.h))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.or_scene_destroy(self.h)
            self.h = None

    def info(self):
        out = (ctypes.c_int * 6)()
        self.lib.or_scene_info(self.h, out)
        keys = ["edges", "primary_edges", "silhouette_nodes", "normal_nodes", "bvh_nodes", "triangles"]
        return dict(zip(keys, list(out)))

    def use_bvh(self, on):
        self.lib.or_scene_use_bvh(self.h, 1 if on else 0)

    def last_render_rays(self):
        """ray queries (closest + any hit) of the last render"""
        return int(self.lib.or_last_render_rays(self.h))

    def render(self, width, height, spp, max_bounces, seed, flags=_abi.DR_FLAGS_PARITY, d_image=None,
               nthreads=0, num_primary_edge_samples=0, want_dx=False):
        o = dr.DeviceScene.options(width, height, spp, max_bounces, seed, flags, num_primary_edge_samples)
        img = np.zeros((height, width, 3), np.float32)
        grads = np.zeros(self.grad_size, np.float32) if d_image is not None else None
        dx = np.zeros_like(img) if want_dx else None
        dimg = None if d_image is None else np.ascontiguousarray(d_image, dtype=np.float32)
        rc = self.lib.or_render(self.h, ctypes.byref(o),
                                None if dimg is None else dimg.ctypes.data,
                                img.ctypes.data, None if grads is None else grads.ctypes.data,
                                None if dx is None else dx.ctypes.data, int(nthreads))
        if rc != 0:
            raise RuntimeError("or_render failed: %d" % rc)
        return img, grads, dx

    def intersect(self, rays, any_hit=False):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        hits = np.zeros((rays.shape[0], 4), np.int32)
        self.lib.or_intersect(self.h, rays.ctypes.data, rays.shape[0], hits.ctypes.data, 1 if any_hit else 0)
        return hits


def scene_from_workload(w):
    return OracleScene(w["camera"], w["shapes"], w["materials"], w["lights"])

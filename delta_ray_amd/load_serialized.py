"""Mitsuba ".serialized" meshes (reference load_serialized.cpp:1-277).

``load_serialized(filename, idx)`` returns ``(indices, vertices, uvs,
normals)`` as the reference's pybind11 module does: int32 [m, 3], float32
[n, 3], float32 [n, 2] and float32 [n, 3], the last two with zero rows when
the file holds none (the reference returns empty arrays there).  The file is
read by the library's C++ reader (dr_load_serialized).
"""
import ctypes

import numpy as np

from . import _abi


def _copy(ptr, shape, dtype):
    if not ptr:
        return np.zeros((0,) + tuple(shape[1:]), dtype)
    n = int(np.prod(shape))
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True).reshape(shape)


def load_serialized(filename, idx=0):
    lib = _abi.load_library()
    m = _abi.dr_mesh_data()
    _abi.check(lib, lib.dr_load_serialized(str(filename).encode(), int(idx), ctypes.byref(m)), "dr_load_serialized")
    try:
        nv, nt = m.num_vertices, m.num_triangles
        vertices = _copy(m.vertices, (nv, 3), np.float32)
        normals = _copy(m.normals, (nv, 3), np.float32)
        uvs = _copy(m.uvs, (nv, 2), np.float32)
        indices = _copy(m.indices, (nt, 3), np.int32)
    finally:
        lib.dr_mesh_data_free(ctypes.byref(m))
    return indices, vertices, uvs, normals

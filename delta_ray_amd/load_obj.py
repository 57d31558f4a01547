"""Wavefront OBJ meshes (reference load_obj.py:6-83).

``load_obj(filename)`` returns ``(indices, vertices, uvs, normals)``: int32
[m, 3] and float32 arrays, uvs / normals None when the file has none.  Same
rules as the reference:

* a face corner ``p/t/n`` (t and n optional) becomes one output vertex per
  distinct (position, uv, normal) triple, in order of first use;
* a quad ``f a b c d`` becomes the triangles (a, b, c) and (a, c, d); faces
  with more than four corners are rejected;
* 1-based indices; a negative index counts back from the number of OUTPUT
  vertices made so far (positions) or output uvs so far (texture
  coordinates), not from the number of ``v`` / ``vt`` lines read -- the
  reference's rule, kept so that files load identically; negative normal
  indices are taken as they are;
* fields are split on runs of spaces only (a tab does not separate them);
  lines of other kinds (``#``, ``o``, ``g``, ``usemtl`` ...) are ignored.
"""
import re

import numpy as np

_SPACES = re.compile(" +")


def load_obj(filename):
    vertices_pool, uvs_pool, normals_pool = [], [], []
    indices, vertices, uvs, normals = [], [], [], []
    vertex_ids = {}

    def index(corner, k):
        f = int(corner.split("/")[k])
        if f < 0:
            if k == 0:
                f += len(vertices)
            elif k == 1:
                f += len(uvs)
        else:
            f -= 1
        return f

    def vertex_id(corner):
        parts = corner.split("/")
        pi = index(corner, 0)
        uvi = index(corner, 1) if len(parts) > 1 and parts[1] != "" else None
        ni = index(corner, 2) if len(parts) > 2 and parts[2] != "" else None
        key = (pi, uvi, ni)
        vid = vertex_ids.get(key)
        if vid is not None:
            return vid
        vid = len(vertices)
        vertex_ids[key] = vid
        vertices.append(vertices_pool[pi])
        if uvi is not None:
            uvs.append(uvs_pool[uvi])
        if ni is not None:
            normals.append(normals_pool[ni])
        return vid

    with open(filename, "r") as fh:
        for line in fh:
            f = _SPACES.split(line.strip())
            tag = f[0]
            if tag == "v":
                vertices_pool.append([float(f[1]), float(f[2]), float(f[3])])
            elif tag == "vt":
                uvs_pool.append([float(f[1]), float(f[2])])
            elif tag == "vn":
                normals_pool.append([float(f[1]), float(f[2]), float(f[3])])
            elif tag == "f":
                if len(f) > 5:
                    raise ValueError("%s: faces with more than four corners are not supported" % filename)
                v0, v1, v2 = vertex_id(f[1]), vertex_id(f[2]), vertex_id(f[3])
                indices.append([v0, v1, v2])
                if len(f) == 5:
                    indices.append([v0, v2, vertex_id(f[4])])
    return (np.array(indices, dtype=np.int32).reshape(-1, 3), np.array(vertices, dtype=np.float32).reshape(-1, 3),
            np.array(uvs, dtype=np.float32) if uvs else None,
            np.array(normals, dtype=np.float32) if normals else None)

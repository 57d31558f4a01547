"""Scene description classes of the reference's Python layer, for
``render_pytorch.RenderFunction.serialize_scene``.

Same constructor arguments and attributes as camera.py:7-47, shape.py,
material.py and light.py of the reference, built on torch tensors, with the
matrix helpers of transform.py:12-49.  ``Material`` keeps the reference's
behaviour of using ``diffuse_uv_scale`` for all three uv scales
(material.py:18-19), so scripts written for the reference render the same.
"""
import math

import numpy as np
import torch


def radians(deg):
    return (math.pi / 180.0) * deg


def _t(x):
    return x if isinstance(x, torch.Tensor) else torch.tensor(np.asarray(x, np.float32))


def _normalize(v):
    return v / torch.norm(v)


def gen_look_at_matrix(pos, look, up):
    """transform.py:12-25: columns left, new_up, d, pos."""
    pos, look, up = _t(pos), _t(look), _t(up)
    d = _normalize(look - pos)
    left = _normalize(torch.cross(_normalize(up), d, dim=0))
    new_up = _normalize(torch.cross(d, left, dim=0))
    z = torch.zeros(1, dtype=pos.dtype)
    o = torch.ones(1, dtype=pos.dtype)
    return torch.stack([torch.cat([left, z]), torch.cat([new_up, z]), torch.cat([d, z]),
                        torch.cat([pos, o])]).t()


def gen_scale_matrix(scale):
    return torch.diag(torch.cat([_t(scale), torch.ones(1)]))


def gen_translate_matrix(t):
    t = _t(t)
    m = torch.eye(4)
    m[:3, 3] = t
    return m


def gen_perspective_matrix(fov, clip_near, clip_far):
    """transform.py:39-49"""
    fov, clip_near, clip_far = _t(fov), _t(clip_near), _t(clip_far)
    clip_dist = clip_far - clip_near
    cot = 1.0 / torch.tan(radians(fov / 2.0))
    m = torch.zeros(4, 4)
    m[0, 0] = cot
    m[1, 1] = cot
    m[2, 2] = 1.0 / clip_dist
    m[2, 3] = -clip_near / clip_dist
    m[3, 2] = 1.0
    return m


class Camera:
    """camera.py:7-47"""

    def __init__(self, position, look_at, up, cam_to_world, fov, clip_near, clip_far, resolution,
                 fisheye=False):
        self.resolution = resolution
        self.num_pixels = resolution[0] * resolution[1]
        if cam_to_world is None:
            self.position, self.look_at, self.up = _t(position), _t(look_at), _t(up)
            self.cam_to_world = gen_look_at_matrix(self.position, self.look_at, self.up)
        else:
            self.cam_to_world = _t(cam_to_world)
        self.world_to_cam = torch.inverse(self.cam_to_world)
        aspect_ratio = resolution[0] / resolution[1]
        s = gen_scale_matrix(np.array([-0.5, -0.5 * aspect_ratio, 1.0], np.float32))
        t = gen_translate_matrix(np.array([-1.0, -1.0 / aspect_ratio, 0.0], np.float32))
        p = gen_perspective_matrix(fov, clip_near, clip_far)
        self.cam_to_sample = s @ t @ p
        self.sample_to_cam = torch.inverse(self.cam_to_sample)
        self.fov = _t(fov)
        self.fov_factor = torch.tan(radians(0.5 * self.fov))
        self.aspect_ratio = aspect_ratio
        self.clip_near = _t(clip_near)
        self.clip_far = _t(clip_far)
        self.fisheye = fisheye


class Shape:
    """shape.py"""

    def __init__(self, vertices, indices, uvs, normals, mat_id):
        self.vertices = vertices
        self.indices = indices
        self.uvs = uvs
        self.normals = normals
        self.mat_id = mat_id


class Material:
    """material.py (specular and roughness uv scales alias the diffuse one,
    material.py:18-19)"""

    def __init__(self, diffuse_reflectance,
                 specular_reflectance=None, roughness=None, diffuse_uv_scale=None,
                 specular_uv_scale=None, roughness_uv_scale=None, two_sided=False):
        self.diffuse_reflectance = diffuse_reflectance
        self.specular_reflectance = (specular_reflectance if specular_reflectance is not None
                                     else torch.zeros(3))
        self.roughness = roughness if roughness is not None else torch.ones(1)
        self.diffuse_uv_scale = diffuse_uv_scale if diffuse_uv_scale is not None else torch.ones(2)
        self.specular_uv_scale = self.diffuse_uv_scale
        self.roughness_uv_scale = self.diffuse_uv_scale
        self.two_sided = two_sided


class Light:
    """light.py"""

    def __init__(self, shape_id, intensity):
        self.shape_id = shape_id
        self.intensity = intensity

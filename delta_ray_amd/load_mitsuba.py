"""Mitsuba 0.5 scene files, the subset the reference reads
(load_mitsuba.py:15-287): ``load_mitsuba(filename)`` returns ``(camera,
materials, shapes, lights, resolution)`` built from this package's scene
classes (scene.Camera, scene.Material, scene.Shape, scene.Light), ready for
``RenderFunction.serialize_scene``.

Read, as the reference:

* ``<sensor>``: ``fov`` (degrees; clip planes stay 0.01 / 10000 and
  ``fovAxis`` is not read), ``toWorld`` as a ``<lookat>`` or as a product of
  ``<matrix>`` / ``<translate>`` / ``<scale>`` (each applied after the ones
  before it), the film's ``width`` / ``height`` (default 256 x 256);
* ``<bsdf type="diffuse">`` with ``reflectance`` and ``specular`` given as
  ``<rgb>`` or as a bitmap ``<texture>`` (``filename``, ``uscale``,
  ``vscale``), and ``roughness``; ``type="twosided"`` around one of them
  (the two-sided material takes the outer id); materials without an id are
  dropped;
* ``<shape>`` of type ``obj``, ``serialized`` (``shapeIndex``) or
  ``rectangle`` (the quad [-1, 1]^2 at z = 0), a ``toWorld`` transform, a
  ``<ref>`` to a material (-1 when none) and an area ``<emitter>`` whose
  ``radiance`` (one or three values) makes the shape a light.  Vertices are
  transformed as points; normals are multiplied, as row vectors, by the
  inverse of the transposed matrix's upper 3 x 3 block and not renormalised
  (the reference's rule, kept so that scenes load identically).

File names are relative to the scene file's folder (the reference changes
the working directory there while it parses).  Where the reference fails --
a ``roughness`` value (it passes a Python float to torch.from_numpy), a bsdf
or shape type outside the subset -- this raises ValueError for the
unsupported types and reads the roughness as a one-element tensor.
"""
import os
import xml.etree.ElementTree as etree

import numpy as np
import torch

from . import image
from . import scene as sc
from .load_obj import load_obj
from .load_serialized import load_serialized


def parse_transform(node):
    ret = np.identity(4, dtype=np.float32)
    for child in node:
        if child.tag == "matrix":
            value = np.array(child.attrib["value"].split(), dtype=np.float32).reshape(4, 4)
            ret = value @ ret
        elif child.tag == "translate":
            t = np.array([float(child.attrib["x"]), float(child.attrib["y"]), float(child.attrib["z"])], np.float32)
            ret = sc.gen_translate_matrix(t).numpy() @ ret
        elif child.tag == "scale":
            s = np.array([float(child.attrib["x"]), float(child.attrib["y"]), float(child.attrib["z"])], np.float32)
            ret = sc.gen_scale_matrix(s).numpy() @ ret
    return ret


def parse_vector(text):
    """three (or one) numbers separated by commas, else by spaces"""
    parts = [p for p in text.split(",") if p.strip() != ""]
    if len(parts) != 3:
        parts = text.split()
    return np.array([float(p) for p in parts], dtype=np.float32)


def parse_camera(node):
    fov = torch.tensor([45.0])
    to_world = np.identity(4, dtype=np.float32)
    position = look_at = up = None
    resolution = [256, 256]
    for child in node:
        name = child.attrib.get("name")
        if name == "fov":
            fov = torch.tensor([float(child.attrib["value"])])
        elif name == "toWorld":
            has_lookat = False
            for g in child:
                if g.tag == "lookat":
                    has_lookat = True
                    position = torch.from_numpy(parse_vector(g.attrib["origin"]))
                    look_at = torch.from_numpy(parse_vector(g.attrib["target"]))
                    up = torch.from_numpy(parse_vector(g.attrib["up"]))
            if not has_lookat:
                to_world = parse_transform(child)
        if child.tag == "film":
            for g in child:
                if g.attrib.get("name") == "width":
                    resolution[0] = int(g.attrib["value"])
                elif g.attrib.get("name") == "height":
                    resolution[1] = int(g.attrib["value"])
    cam_to_world = None if position is not None else torch.from_numpy(to_world)
    camera = sc.Camera(position, look_at, up, cam_to_world, fov, torch.tensor([0.01]), torch.tensor([10000.0]),
                       resolution)
    return camera, resolution


def _texture(child, folder, uv_scale):
    value = None
    for g in child:
        n = g.attrib.get("name")
        if n == "filename":
            value = torch.from_numpy(image.imread(os.path.join(folder, g.attrib["value"])))
        elif n == "uscale":
            uv_scale[0] = float(g.attrib["value"])
        elif n == "vscale":
            uv_scale[1] = float(g.attrib["value"])
    return value


def parse_material(node, folder, two_sided=False):
    node_id = node.attrib.get("id")
    kind = node.attrib["type"]
    if kind == "diffuse":
        diffuse = torch.tensor([0.5, 0.5, 0.5])
        diffuse_uv_scale = torch.tensor([1.0, 1.0])
        specular = torch.tensor([0.0, 0.0, 0.0])
        specular_uv_scale = torch.tensor([1.0, 1.0])
        roughness = torch.tensor([1.0])
        for child in node:
            n = child.attrib["name"]
            if n == "reflectance":
                if child.tag == "texture":
                    t = _texture(child, folder, diffuse_uv_scale)
                    diffuse = t if t is not None else diffuse
                elif child.tag == "rgb":
                    diffuse = torch.from_numpy(parse_vector(child.attrib["value"]))
            elif n == "specular":
                if child.tag == "texture":
                    t = _texture(child, folder, specular_uv_scale)
                    specular = t if t is not None else specular
                elif child.tag == "rgb":
                    specular = torch.from_numpy(parse_vector(child.attrib["value"]))
            elif n == "roughness":
                roughness = torch.tensor([float(child.attrib["value"])])
        return node_id, sc.Material(diffuse, specular_reflectance=specular, roughness=roughness,
                                    diffuse_uv_scale=diffuse_uv_scale, specular_uv_scale=specular_uv_scale,
                                    two_sided=two_sided)
    if kind == "twosided":
        return node_id, parse_material(node[0], folder, True)[1]
    raise ValueError("bsdf type %r is not supported" % kind)


def _emitter(child):
    for g in child:
        if g.attrib.get("name") == "radiance":
            v = parse_vector(g.attrib["value"])
            if v.shape[0] == 1:
                v = np.array([v[0], v[0], v[0]], dtype=np.float32)
            return torch.from_numpy(v)
    return None


def _transformed(vertices, normals, to_world):
    v = np.concatenate([vertices, np.ones((vertices.shape[0], 1), np.float32)], 1) @ to_world.T
    v = (v / v[:, 3:4])[:, :3].astype(np.float32)
    if normals is not None:
        normals = (normals @ np.linalg.inv(to_world.T)[:3, :3]).astype(np.float32)
    return v, normals


def parse_shape(node, material_dict, shape_id, folder):
    kind = node.attrib["type"]
    if kind not in ("obj", "serialized", "rectangle"):
        raise ValueError("shape type %r is not supported" % kind)
    to_world = np.identity(4, dtype=np.float32)
    serialized_id = 0
    mat_id = -1
    intensity = None
    filename = ""
    for child in node:
        n = child.attrib.get("name")
        if n == "filename":
            filename = child.attrib["value"]
        elif n == "toWorld":
            to_world = parse_transform(child)
        elif n == "shapeIndex":
            serialized_id = int(child.attrib["value"])
        if child.tag == "ref":
            mat_id = material_dict[child.attrib["id"]]
        elif child.tag == "emitter":
            intensity = _emitter(child)
    if kind == "obj":
        indices, vertices, uvs, normals = load_obj(os.path.join(folder, filename))
    elif kind == "serialized":
        indices, vertices, uvs, normals = load_serialized(os.path.join(folder, filename), serialized_id)
        uvs = None if uvs.shape[0] == 0 else uvs
        normals = None if normals.shape[0] == 0 else normals
    else:
        indices = np.array([[0, 2, 1], [1, 2, 3]], dtype=np.int32)
        vertices = np.array([[-1, -1, 0], [-1, 1, 0], [1, -1, 0], [1, 1, 0]], dtype=np.float32)
        uvs = normals = None
    vertices, normals = _transformed(vertices, normals, to_world)
    light = sc.Light(shape_id, intensity) if intensity is not None else None
    shape = sc.Shape(torch.from_numpy(vertices), torch.from_numpy(np.asarray(indices, np.int32)),
                     None if uvs is None else torch.from_numpy(uvs),
                     None if normals is None else torch.from_numpy(normals), mat_id)
    return shape, light


def parse_scene(node, folder):
    camera = resolution = None
    materials, material_dict, shapes, lights = [], {}, [], []
    for child in node:
        if child.tag == "sensor":
            camera, resolution = parse_camera(child)
        elif child.tag == "bsdf":
            node_id, material = parse_material(child, folder)
            if node_id is not None:
                material_dict[node_id] = len(materials)
                materials.append(material)
        elif child.tag == "shape":
            shape, light = parse_shape(child, material_dict, len(shapes), folder)
            shapes.append(shape)
            if light is not None:
                lights.append(light)
    return camera, materials, shapes, lights, resolution


def load_mitsuba(filename):
    root = etree.parse(filename).getroot()
    return parse_scene(root, os.path.dirname(os.path.abspath(filename)))

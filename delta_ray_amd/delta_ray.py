"""Drop-in mirror of the reference's ``delta_ray`` extension module.

The reference exposes, from C++ through pybind11 (delta_ray.cpp:1838-1920):
``Material``, ``DMaterial``, ``Shape``, ``DShape``, ``Light``, ``DLight``,
``Camera``, ``DCamera``, ``RenderResult``, ``render`` and
``test_line_integral``.  This module keeps those names, constructor argument
orders and field names, and runs ``render`` on the MI355X through the C-ABI of
``libdelta_ray_amd.so`` (include/delta_ray.h).

Arrays are numpy float32 / int32 as in the reference.  ``render`` returns
numpy arrays (host copies), like the reference; ``render_device`` keeps
everything in device memory for callers that already hold torch tensors on
the GPU.
"""
import ctypes

import numpy as np

from . import _abi


class Material:
    """reference material.h:14-22 (pybind init order delta_ray.cpp:1841-1849)"""

    def __init__(self, diffuse_reflectance, specular_reflectance, roughness,
                 diffuse_uv_scale, specular_uv_scale, roughness_uv_scale, two_sided):
        self.diffuse_reflectance = _abi.f32(diffuse_reflectance)
        self.specular_reflectance = _abi.f32(specular_reflectance)
        self.roughness = _abi.f32(roughness)
        self.diffuse_uv_scale = _abi.f32(diffuse_uv_scale)
        self.specular_uv_scale = _abi.f32(specular_uv_scale)
        self.roughness_uv_scale = _abi.f32(roughness_uv_scale)
        self.two_sided = bool(two_sided)


class DMaterial:
    def __init__(self, diffuse_reflectance, specular_reflectance, roughness,
                 diffuse_uv_scale=None, specular_uv_scale=None, roughness_uv_scale=None):
        self.diffuse_reflectance = diffuse_reflectance
        self.specular_reflectance = specular_reflectance
        self.roughness = roughness
        # the reference never fills the uv-scale gradients (delta_ray.cpp:1127-1128)
        empty = np.zeros((0,), dtype=np.float32)
        self.diffuse_uv_scale = empty if diffuse_uv_scale is None else diffuse_uv_scale
        self.specular_uv_scale = empty if specular_uv_scale is None else specular_uv_scale
        self.roughness_uv_scale = empty if roughness_uv_scale is None else roughness_uv_scale


class Shape:
    """reference shape.h:11-18; ``light`` is read-write as there."""

    def __init__(self, vertices, indices, uvs, normals, material, light):
        self.vertices = _abi.f32(vertices)
        self.indices = _abi.i32(indices)
        self.uvs = None if uvs is None else _abi.f32(uvs)
        self.normals = None if normals is None else _abi.f32(normals)
        self.material = material
        self.light = light


class DShape:
    def __init__(self, vertices, normals):
        self.vertices = vertices
        self.normals = normals


class Light:
    """reference light.h:15-18"""

    def __init__(self, shape, intensity):
        self.shape = shape
        self.intensity = _abi.f32(intensity)


class DLight:
    def __init__(self, intensity):
        self.intensity = intensity


class Camera:
    """reference camera.h:12-21"""

    def __init__(self, cam_to_world, world_to_cam, sample_to_cam, cam_to_sample,
                 fov_factor, aspect_ratio, clip_near, fisheye):
        self.cam_to_world = _abi.f32(cam_to_world)
        self.world_to_cam = _abi.f32(world_to_cam)
        self.sample_to_cam = _abi.f32(sample_to_cam)
        self.cam_to_sample = _abi.f32(cam_to_sample)
        self.fov_factor = float(fov_factor)
        self.aspect_ratio = float(aspect_ratio)
        self.clip_near = float(clip_near)
        self.fisheye = bool(fisheye)


class DCamera:
    def __init__(self, d_cam_to_world, d_world_to_cam, d_sample_to_cam, d_cam_to_sample):
        self.d_cam_to_world = d_cam_to_world
        self.d_world_to_cam = d_world_to_cam
        self.d_sample_to_cam = d_sample_to_cam
        self.d_cam_to_sample = d_cam_to_sample


class RenderResult:
    def __init__(self):
        self.image = None
        self.dx_image = None
        self.dy_image = None
        self.d_camera = None
        self.d_shapes = []
        self.d_materials = []
        self.d_lights = []


# ----------------------------------------------------------------- desc ----
def _texture(arr, channels):
    a = _abi.f32(arr)
    t = _abi.dr_texture()
    t.data = _abi.fptr(a)
    t.channels = channels
    if a.ndim == 3:
        if a.shape[2] != channels:
            raise ValueError("texture must have %d channels, got shape %s" % (channels, a.shape))
        t.height, t.width = a.shape[0], a.shape[1]
    else:
        if a.size != channels:
            raise ValueError("constant must have %d values, got %d" % (channels, a.size))
        t.height = t.width = 0
    return t, a


class SceneDesc:
    """A dr_scene_desc plus the numpy arrays it points into (kept alive)."""

    def __init__(self, camera, shapes, materials, lights):
        keep = []
        d = _abi.dr_scene_desc()
        d.camera = SceneDesc.camera_struct(camera)
        mats, mkeep = _material_structs(materials)
        keep += mkeep
        mat_index = {id(m): i for i, m in enumerate(materials)}
        light_index = {id(l): i for i, l in enumerate(lights)}
        shape_index = {id(s): i for i, s in enumerate(shapes)}
        shs = (_abi.dr_shape * max(1, len(shapes)))()
        for i, s in enumerate(shapes):
            ss = shs[i]
            v = _abi.f32(s.vertices)
            ind = _abi.i32(s.indices)
            if v.ndim != 2 or v.shape[1] != 3 or ind.ndim != 2 or ind.shape[1] != 3:
                raise ValueError("shape %d: vertices [N,3] and indices [M,3] expected" % i)
            if ind.size and (ind.min() < 0 or ind.max() >= v.shape[0]):
                raise ValueError("shape %d: vertex index out of range" % i)
            keep += [v, ind]
            ss.vertices = _abi.fptr(v)
            ss.num_vertices = v.shape[0]
            ss.indices = _abi.iptr(ind)
            ss.num_triangles = ind.shape[0]
            uv = None if s.uvs is None else _abi.f32(s.uvs)
            nrm = None if s.normals is None else _abi.f32(s.normals)
            if uv is not None and uv.shape != (v.shape[0], 2):
                raise ValueError("shape %d: uvs must be [N,2]" % i)
            if nrm is not None and nrm.shape != (v.shape[0], 3):
                raise ValueError("shape %d: normals must be [N,3]" % i)
            keep += [uv, nrm]
            ss.uvs = _abi.fptr(uv)
            ss.normals = _abi.fptr(nrm)
            if id(s.material) not in mat_index:
                raise ValueError("shape %d: material not in the material list" % i)
            ss.material_id = mat_index[id(s.material)]
            ss.light_id = light_index.get(id(s.light), -1) if s.light is not None else -1
        lts = (_abi.dr_light * max(1, len(lights)))()
        for i, l in enumerate(lights):
            if id(l.shape) not in shape_index:
                raise ValueError("light %d: shape not in the shape list" % i)
            lts[i].shape_id = shape_index[id(l.shape)]
            lts[i].intensity[:] = _abi.f32(l.intensity).reshape(-1)[:3].tolist()
            shs[lts[i].shape_id].light_id = i
        d.shapes = shs
        d.num_shapes = len(shapes)
        d.materials = mats
        d.num_materials = len(materials)
        d.lights = lts
        d.num_lights = len(lights)
        self.desc = d
        self._keep = (keep, mats, shs, lts)
        self.shapes, self.materials, self.lights = shapes, materials, lights
        self.layout = grad_layout(shapes, materials, lights)


def _material_structs(materials):
    """dr_material array for Material objects, and the arrays it points into"""
    keep = []
    mats = (_abi.dr_material * max(1, len(materials)))()
    for i, m in enumerate(materials):
        mm = mats[i]
        mm.diffuse_reflectance, a = _texture(m.diffuse_reflectance, 3)
        keep.append(a)
        mm.specular_reflectance, a = _texture(m.specular_reflectance, 3)
        keep.append(a)
        mm.roughness, a = _texture(m.roughness, 1)
        keep.append(a)
        mm.diffuse_uv_scale[:] = _abi.f32(m.diffuse_uv_scale).reshape(-1)[:2].tolist()
        mm.specular_uv_scale[:] = _abi.f32(m.specular_uv_scale).reshape(-1)[:2].tolist()
        mm.roughness_uv_scale[:] = _abi.f32(m.roughness_uv_scale).reshape(-1)[:2].tolist()
        mm.two_sided = 1 if m.two_sided else 0
    return mats, keep


def _camera_struct(camera):
    c = _abi.dr_camera()
    for name in ("cam_to_world", "world_to_cam", "sample_to_cam", "cam_to_sample"):
        m = _abi.f32(getattr(camera, name))
        if m.shape != (4, 4):
            raise ValueError("%s must be 4x4" % name)
        getattr(c, name)[:] = m.reshape(-1).tolist()
    c.fov_factor = float(camera.fov_factor)
    c.aspect_ratio = float(camera.aspect_ratio)
    c.clip_near = float(camera.clip_near)
    c.fisheye = 1 if camera.fisheye else 0
    return c


SceneDesc.camera_struct = staticmethod(_camera_struct)


def grad_layout(shapes, materials, lights):
    """Offsets of every parameter gradient in the flat buffer, in the order of
    dr_grad_layout_compute (include/delta_ray.h)."""
    lay = {"cam_to_world": (0, (4, 4)), "world_to_cam": (16, (4, 4)),
           "sample_to_cam": (32, (4, 4)), "cam_to_sample": (48, (4, 4))}
    o = 64
    mats = []
    for m in materials:
        entry = {}
        for name in ("diffuse_reflectance", "specular_reflectance", "roughness"):
            a = np.asarray(getattr(m, name))
            shape = a.shape
            entry[name] = (o, shape)
            o += int(np.prod(shape))
        mats.append(entry)
    shs = []
    for s in shapes:
        nv = np.asarray(s.vertices).shape[0]
        entry = {"vertices": (o, (nv, 3))}
        o += nv * 3
        if s.normals is not None:
            entry["normals"] = (o, (nv, 3))
            o += nv * 3
        shs.append(entry)
    lts = []
    for _ in lights:
        lts.append({"intensity": (o, (3,))})
        o += 3
    lay["materials"] = mats
    lay["shapes"] = shs
    lay["lights"] = lts
    lay["total"] = o
    return lay


def split_grads(flat, layout):
    """Split a flat gradient vector (numpy or torch) into the reference's
    DCamera / DShape / DMaterial / DLight objects."""
    def view(off, shape):
        n = int(np.prod(shape))
        return flat[off:off + n].reshape(shape)

    d_camera = DCamera(*(view(*layout[k]) for k in
                         ("cam_to_world", "world_to_cam", "sample_to_cam", "cam_to_sample")))
    d_materials = [DMaterial(view(*e["diffuse_reflectance"]), view(*e["specular_reflectance"]),
                             view(*e["roughness"])) for e in layout["materials"]]
    d_shapes = [DShape(view(*e["vertices"]),
                       view(*e["normals"]) if "normals" in e else np.zeros((0,), np.float32))
                for e in layout["shapes"]]
    d_lights = [DLight(view(*e["intensity"])) for e in layout["lights"]]
    return d_camera, d_shapes, d_materials, d_lights


# --------------------------------------------------------------- device ----
class DeviceScene:
    """A scene resident on the GPU (dr_scene_create).  Keep it alive across
    forward and backward of one optimisation step."""

    def __init__(self, camera, shapes, materials, lights):
        self.lib = _abi.load_library()
        self.sd = SceneDesc(camera, shapes, materials, lights)
        h = ctypes.c_void_p()
        _abi.check(self.lib, self.lib.dr_scene_create(ctypes.byref(self.sd.desc), ctypes.byref(h)),
                   "dr_scene_create")
        self.handle = h
        self.grad_size = int(self.lib.dr_scene_grad_size(h))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self.lib.dr_scene_destroy(h)
            self.handle = None

    def build_bvh_device(self, stream=None):
        """Replace the triangle BVH with one built on the device
        (dr_scene_build_bvh_device); returns the device build's time in ms."""
        ms = ctypes.c_double(0.0)
        _abi.check(self.lib, self.lib.dr_scene_build_bvh_device(self.handle, ctypes.byref(ms), _stream_ptr(stream)),
                   "dr_scene_build_bvh_device")
        return ms.value

    def info(self):
        i = _abi.dr_scene_info()
        _abi.check(self.lib, self.lib.dr_scene_get_info(self.handle, ctypes.byref(i)), "dr_scene_get_info")
        return {k: getattr(i, k) for k, _ in _abi.dr_scene_info._fields_}

    @staticmethod
    def options(width, height, num_samples, max_bounces, seed, flags=_abi.DR_FLAGS_PARITY,
                num_primary_edge_samples=0):
        o = _abi.dr_render_options()
        o.width, o.height = int(width), int(height)
        o.num_samples, o.max_bounces = int(num_samples), int(max_bounces)
        o.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        o.flags = int(flags)
        o.num_primary_edge_samples = int(num_primary_edge_samples)
        return o

    def render(self, opts, image, stream=None):
        """image: torch float32 CUDA tensor [H, W, 3] (written)."""
        s = _stream_ptr(stream)
        _abi.check(self.lib, self.lib.dr_render(self.handle, ctypes.byref(opts),
                                                ctypes.c_void_p(image.data_ptr()), s), "dr_render")

    def render_adjoint(self, opts, d_image, image, grads, dx_image=None, stream=None):
        s = _stream_ptr(stream)
        _abi.check(self.lib, self.lib.dr_render_adjoint(
            self.handle, ctypes.byref(opts), ctypes.c_void_p(d_image.data_ptr()),
            ctypes.c_void_p(image.data_ptr()), ctypes.c_void_p(grads.data_ptr()),
            ctypes.c_void_p(dx_image.data_ptr() if dx_image is not None else None), s),
            "dr_render_adjoint")

    def set_edge_policy(self, policy):
        """dr_scene_set_edge_policy: _abi.DR_EDGES_FILTERED (trees over the
        edges the coplanar filter keeps, as the reference; a vertex update
        that flips a verdict rebuilds on the host) or
        _abi.DR_EDGES_ALL_CANDIDATES (trees over every mesh edge, the
        filter's verdicts as device flags: vertex updates never rebuild)."""
        _abi.check(self.lib, self.lib.dr_scene_set_edge_policy(self.handle, int(policy)),
                   "dr_scene_set_edge_policy")

    def update_vertices(self, updates, stream=None):
        """updates: [(shape_index, vertices, normals or None)], torch float32
        CUDA tensors [N, 3] of the shapes' sizes (dr_scene_update_vertices:
        device refit, or a host rebuild when the edge set changes)."""
        import torch
        arr = (_abi.dr_shape_update * max(1, len(updates)))()
        keep = []
        for i, (sid, v, n) in enumerate(updates):
            v = v.detach().to(torch.float32).contiguous()
            if not v.is_cuda:
                v = v.cuda()
            keep.append(v)
            arr[i].shape_id = int(sid)
            arr[i].vertices = v.data_ptr()
            if n is not None:
                n = n.detach().to(torch.float32).contiguous()
                if not n.is_cuda:
                    n = n.cuda()
                keep.append(n)
                arr[i].normals = n.data_ptr()
            else:
                arr[i].normals = None
        _abi.check(self.lib, self.lib.dr_scene_update_vertices(self.handle, arr, len(updates), _stream_ptr(stream)),
                   "dr_scene_update_vertices")
        del keep

    def update_materials(self, materials):
        """new parameters of the scene's materials (same texture sizes)"""
        mats, keep = _material_structs(materials)
        _abi.check(self.lib, self.lib.dr_scene_update_materials(self.handle, mats, len(materials)),
                   "dr_scene_update_materials")
        del keep

    def update_lights(self, shape_ids, intensities):
        """new intensities of the scene's lights (same shapes)"""
        lts = (_abi.dr_light * max(1, len(shape_ids)))()
        for i, (sid, inten) in enumerate(zip(shape_ids, intensities)):
            lts[i].shape_id = int(sid)
            lts[i].intensity[:] = _abi.f32(inten).reshape(-1)[:3].tolist()
        _abi.check(self.lib, self.lib.dr_scene_update_lights(self.handle, lts, len(shape_ids)),
                   "dr_scene_update_lights")

    def set_camera(self, camera):
        c = _camera_struct(camera)
        _abi.check(self.lib, self.lib.dr_scene_set_camera(self.handle, ctypes.byref(c)), "dr_scene_set_camera")

    def intersect(self, rays, hits, any_hit=False, stream=None):
        s = _stream_ptr(stream)
        _abi.check(self.lib, self.lib.dr_intersect(self.handle, ctypes.c_void_p(rays.data_ptr()),
                                                   rays.shape[0], ctypes.c_void_p(hits.data_ptr()),
                                                   1 if any_hit else 0, s), "dr_intersect")

    def stats(self, enable):
        """enable: bit mask of _abi.DR_STATS_COUNTERS / DR_STATS_TIMING (True =
        counters).  Returns the statistics gathered since the previous call."""
        st = _abi.dr_stats()
        mask = _abi.DR_STATS_COUNTERS if enable is True else int(enable)
        _abi.check(self.lib, self.lib.dr_scene_stats(self.handle, mask, ctypes.byref(st)),
                   "dr_scene_stats")
        out = {k: getattr(st, k) for k in ("rays", "nodes_visited", "tris_visited", "secondary_edges",
                                           "primary_edge_samples", "edge_nodes_visited", "edges_tested",
                                           "stack_overflows")}
        for i, k in enumerate(_abi.KERNELS):
            out[k + "_ms"] = st.kernel_ms[i]
            out[k + "_launches"] = st.kernel_launches[i]
        return out


class View:
    """The scene of a DeviceScene seen by another camera (dr_view_create); keep
    the DeviceScene alive while the view is used."""

    def __init__(self, scene, camera):
        self.scene = scene
        self.lib = scene.lib
        c = SceneDesc.camera_struct(camera)
        h = ctypes.c_void_p()
        _abi.check(self.lib, self.lib.dr_view_create(scene.handle, ctypes.byref(c), ctypes.byref(h)),
                   "dr_view_create")
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self.lib.dr_view_destroy(h)
            self.handle = None


class Batch:
    """Many renders in one launch sequence (dr_render_batch).  A job is a dict
    with scene (DeviceScene), optional view (View), options
    (DeviceScene.options), image, optional d_image / grads / dx_image (torch
    CUDA tensors); jobs with d_image are forward + adjoint."""

    def __init__(self):
        self.lib = _abi.load_library()
        h = ctypes.c_void_p()
        _abi.check(self.lib, self.lib.dr_batch_create(ctypes.byref(h)), "dr_batch_create")
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self.lib.dr_batch_destroy(h)
            self.handle = None

    @staticmethod
    def _jobs(jobs):
        arr = (_abi.dr_job * len(jobs))()
        for i, j in enumerate(jobs):
            a = arr[i]
            a.scene = j["scene"].handle
            a.view = j["view"].handle if j.get("view") is not None else None
            a.opts = j["options"]
            ptr = lambda t: t.data_ptr() if t is not None else None
            a.d_image = ptr(j.get("d_image"))
            a.image = ptr(j["image"])
            a.grads = ptr(j.get("grads"))
            a.dx_image = ptr(j.get("dx_image"))
        return arr

    def render(self, jobs, graph=False, stream=None):
        arr = self._jobs(jobs)
        _abi.check(self.lib, self.lib.dr_render_batch(self.handle, arr, len(jobs),
                                                      _abi.DR_BATCH_GRAPH if graph else 0, _stream_ptr(stream)),
                   "dr_render_batch")

    def autotune(self, jobs, stream=None):
        arr = self._jobs(jobs)
        t = _abi.dr_tuning()
        _abi.check(self.lib, self.lib.dr_batch_autotune(self.handle, arr, len(jobs), _stream_ptr(stream),
                                                        ctypes.byref(t)), "dr_batch_autotune")
        return {"pass_samples": t.pass_samples, "min_threads": t.min_threads, "select_refill": t.select_refill,
                "best_ms": t.best_ms, "candidates": t.candidates}

    def set_tuning(self, pass_samples, min_threads, select_refill=64):
        """select_refill: 1-64 = depth-first edge selection (the default 64),
        with the idle lanes a wave gathers before starting new shading
        points; 0 = breadth first."""
        t = _abi.dr_tuning(pass_samples, min_threads, select_refill, 0.0, 0)
        _abi.check(self.lib, self.lib.dr_batch_set_tuning(self.handle, ctypes.byref(t)), "dr_batch_set_tuning")


def _stream_ptr(stream):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def render(camera, shapes, materials, lights, resolution, d_image, num_samples, max_bounces,
           seed, print_progress=False, flags=_abi.DR_FLAGS_PARITY):
    """reference delta_ray.cpp:1017-1570.  ``d_image`` 0-d (or None): forward
    only; [H, W, 3]: forward + adjoint.  Returns a RenderResult of numpy
    arrays."""
    import torch
    width, height = int(resolution[0]), int(resolution[1])
    scene = DeviceScene(camera, shapes, materials, lights)
    opts = DeviceScene.options(width, height, num_samples, max_bounces, seed, flags)
    dev = torch.device("cuda")
    image = torch.empty((height, width, 3), dtype=torch.float32, device=dev)
    res = RenderResult()
    d_img = None if d_image is None else np.asarray(d_image, dtype=np.float32)
    if d_img is None or d_img.ndim != 3:
        scene.render(opts, image)
        res.image = image.cpu().numpy()
        res.dx_image = np.zeros_like(res.image)
        res.dy_image = np.zeros_like(res.image)
        z = np.zeros((4, 4), np.float32)
        res.d_camera = DCamera(z.copy(), z.copy(), z.copy(), z.copy())
        return res
    if d_img.shape != (height, width, 3):
        raise ValueError("d_image must be [H, W, 3] = %s" % ((height, width, 3),))
    opts.flags |= _abi.DR_WRITE_DX_IMAGE
    d_dev = torch.from_numpy(np.ascontiguousarray(d_img)).to(dev)
    grads = torch.empty((scene.grad_size,), dtype=torch.float32, device=dev)
    dx = torch.empty_like(image)
    scene.render_adjoint(opts, d_dev, image, grads, dx)
    res.image = image.cpu().numpy()
    res.dx_image = dx.cpu().numpy()
    res.dy_image = np.zeros_like(res.image)
    g = grads.cpu().numpy()
    res.d_camera, res.d_shapes, res.d_materials, res.d_lights = split_grads(g, scene.sd.layout)
    return res


def test_line_integral(num_samples=65536, seed=0, stream=None):
    """reference delta_ray.cpp:1572-1836 (run by unit_tests.py): uniform vs
    LTC-importance-sampled estimates of the line integral of a cosine and of a
    glossy lobe over the segment (-10,3,0)-(10,3,0), num_samples each, on the
    GPU with the device's line sampler (the one secondary edges use).  Prints
    the four means and variances with the reference's labels and returns them
    as a dict."""
    lib = _abi.load_library()
    out = (ctypes.c_double * 8)()
    st = _stream_ptr(stream)
    _abi.check(lib, lib.dr_test_line_integral(int(num_samples), int(seed), out, st), "dr_test_line_integral")
    labels = ["Uniform line", "Cosine line", "Uniform line glossy", "Cosine line glossy"]
    vlabels = ["Uniform variance", "Cosine variance", "Uniform line glossy variance",
               "Cosine line glossy variance"]
    res = {}
    for k in range(4):
        res[labels[k]] = out[k]
        res[vlabels[k]] = out[4 + k]
    for k in range(4):
        print("%s:%g" % (labels[k], out[k]))
        print("%s:%g" % (vlabels[k], out[4 + k]))
    return res

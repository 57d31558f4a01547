"""Synthetic scenes of the benchmark workloads, built with seeded numpy RNGs.

single_triangle  1 diffuse triangle (albedo 0.5), vertices (x, y) ~ U[-1,1]^2,
                 z = 0; camera at (0,0,5) looking at the origin, fov 45; one
                 triangle area light at z = 7 facing -z, intensity 20
                 (cf. the reference's test/test_single_triangle.py:28-41).
shadow           ground quad [-5,5]^2 at y = 0, a blocker triangle with
                 (x, z) ~ U[-1,1]^2, y ~ U[1,2]; a 0.2 x 0.2 light quad at
                 y = 5 facing -y, intensity 10; camera (0,2,-5) -> origin
                 (cf. test/test_shadow.py:12-47).
heightfield      128 x 128 grid on [-5,5]^2 with heights 0.2 * fBm value
                 noise (6 octaves, lacunarity 2, gain 0.5), angle-weighted
                 vertex normals, uvs = grid coordinates, Blinn-Phong
                 material with F0 ~ U[0.04, 1], diffuse 0.5 (1 - F0), a
                 128 x 128 roughness texture ~ U[0.05, 0.5]; two 1 x 1 light
                 quads (intensity 50) at (+-3, 6, 0) facing the origin;
                 camera (0,3,-6) -> origin, fov 45.

Each builder returns a dict with the reference-API objects (delta_ray.Camera,
Shape, Material, Light) plus the resolution / spp / bounce budget of the
workload.  Camera matrices follow the reference's camera.py / transform.py
(look-at, perspective, cam_to_sample = scale @ translate @ perspective).
"""
import math

import numpy as np

from . import delta_ray as dr


# ------------------------------------------------------------ camera ------
def look_at_matrix(pos, look, up):
    """transform.py:12-25: columns left, new_up, d, pos."""
    pos, look, up = (np.asarray(a, np.float64) for a in (pos, look, up))

    def nrm(v):
        return v / np.linalg.norm(v)
    d = nrm(look - pos)
    left = nrm(np.cross(nrm(up), d))
    new_up = nrm(np.cross(d, left))
    m = np.eye(4)
    m[:3, 0], m[:3, 1], m[:3, 2], m[:3, 3] = left, new_up, d, pos
    return m


def perspective_matrix(fov, clip_near, clip_far):
    """transform.py:39-49"""
    clip_dist = clip_far - clip_near
    cot = 1.0 / math.tan(math.radians(fov / 2.0))
    return np.array([[cot, 0, 0, 0], [0, cot, 0, 0],
                     [0, 0, 1.0 / clip_dist, -clip_near / clip_dist], [0, 0, 1, 0]], np.float64)


def make_camera(position, look_at, up, fov, resolution, clip_near=0.01, clip_far=10000.0,
                cam_to_world=None, fisheye=False):
    """camera.py:7-47 -> a delta_ray.Camera."""
    c2w = look_at_matrix(position, look_at, up) if cam_to_world is None else np.asarray(cam_to_world, np.float64)
    w2c = np.linalg.inv(c2w)
    aspect = resolution[0] / resolution[1]
    s = np.diag([-0.5, -0.5 * aspect, 1.0, 1.0])
    t = np.eye(4)
    t[:3, 3] = [-1.0, -1.0 / aspect, 0.0]
    p = perspective_matrix(fov, clip_near, clip_far)
    c2s = s @ t @ p
    s2c = np.linalg.inv(c2s)
    fov_factor = math.tan(math.radians(0.5 * fov))
    return dr.Camera(c2w.astype(np.float32), w2c.astype(np.float32), s2c.astype(np.float32),
                     c2s.astype(np.float32), np.float32(fov_factor), aspect, clip_near, fisheye)


def diffuse_material(albedo, specular=(0.0, 0.0, 0.0), roughness=(1.0,), two_sided=False):
    one = np.ones(2, np.float32)
    return dr.Material(np.asarray(albedo, np.float32), np.asarray(specular, np.float32),
                       np.asarray(roughness, np.float32), one, one, one, two_sided)


def quad(center, normal, size, up_hint=(0.0, 1.0, 0.0)):
    """Two triangles of a square of side `size` facing `normal`; the winding
    makes the geometric normal (cross(v1-v0, v2-v0)) equal `normal`."""
    n = np.asarray(normal, np.float64)
    n = n / np.linalg.norm(n)
    a = np.asarray(up_hint, np.float64)
    if abs(np.dot(a, n)) > 0.9:
        a = np.array([1.0, 0.0, 0.0])
    t = np.cross(a, n)
    t /= np.linalg.norm(t)
    b = np.cross(n, t)
    h = 0.5 * size
    c = np.asarray(center, np.float64)
    v = np.array([c - h * t - h * b, c + h * t - h * b, c - h * t + h * b, c + h * t + h * b])
    ind = np.array([[0, 1, 2], [1, 3, 2]], np.int32)
    gn = np.cross(v[1] - v[0], v[2] - v[0])
    if np.dot(gn, n) < 0:
        ind = ind[:, [0, 2, 1]]
    return v.astype(np.float32), ind


def _orient(v, ind, want):
    """Flip triangle windings so that geometric normals point along `want`."""
    ind = ind.copy()
    for k in range(ind.shape[0]):
        a, b, c = v[ind[k]]
        if np.dot(np.cross(b - a, c - a), want) < 0:
            ind[k] = ind[k][[0, 2, 1]]
    return ind


# ------------------------------------------------------------- scenes -----
def single_triangle(seed=1, resolution=(256, 256), num_samples=4, max_bounces=1):
    rng = np.random.default_rng(seed)
    xy = rng.uniform(-1.0, 1.0, size=(3, 2))
    v = np.concatenate([xy, np.zeros((3, 1))], axis=1).astype(np.float32)
    ind = _orient(v, np.array([[0, 1, 2]], np.int32), np.array([0, 0, 1.0]))
    cam = make_camera([0, 0, 5], [0, 0, 0], [0, 1, 0], 45.0, resolution)
    grey = diffuse_material([0.5, 0.5, 0.5])
    black = diffuse_material([0.0, 0.0, 0.0])
    lv = np.array([[-1, -1, 7], [-1, 1, 7], [1, -1, 7]], np.float32)
    li = np.array([[0, 1, 2]], np.int32)
    tri = dr.Shape(v, ind, None, None, grey, None)
    light_shape = dr.Shape(lv, li, None, None, black, None)
    light = dr.Light(light_shape, np.array([20, 20, 20], np.float32))
    light_shape.light = light
    return dict(name="single_triangle", camera=cam, shapes=[tri, light_shape], materials=[grey, black],
                lights=[light], resolution=tuple(resolution), num_samples=num_samples,
                max_bounces=max_bounces, params=[("shape", 0, "vertices")])


def _triangle_and_quad_light(tri, light_z, intensity, cam, resolution, num_samples, max_bounces, name):
    grey = diffuse_material([0.5, 0.5, 0.5])
    tv = np.asarray(tri, np.float32)
    lv = np.array([[-1, -1, light_z], [1, -1, light_z], [-1, 1, light_z], [1, 1, light_z]], np.float32)
    li = np.array([[0, 1, 2], [1, 3, 2]], np.int32)
    t = dr.Shape(tv, np.array([[0, 1, 2]], np.int32), None, None, grey, None)
    ls = dr.Shape(lv, li, None, None, grey, None)
    light = dr.Light(ls, np.full(3, intensity, np.float32))
    ls.light = light
    return dict(name=name, camera=cam, shapes=[t, ls], materials=[grey], lights=[light],
                resolution=tuple(resolution), num_samples=num_samples, max_bounces=max_bounces,
                params=[("shape", 0, "vertices")])


def single_triangle_fisheye(resolution=(256, 256), num_samples=4, max_bounces=1):
    """The reference's fisheye test scene (test_single_triangle_fisheye.py):
    a triangle at z = 0 seen from z = -1 through the equi-angular camera, a
    quad light at z = -9."""
    cam = make_camera([0, 0, -1], [0, 0, 0], [0, 1, 0], 45.0, resolution, fisheye=True)
    return _triangle_and_quad_light([[-1.7, 1.0, 0.0], [1.0, 1.0, 0.0], [-0.5, -1.0, 0.0]], -9.0, 30.0, cam,
                                    resolution, num_samples, max_bounces, "single_triangle_fisheye")


def single_triangle_clipped(resolution=(256, 256), num_samples=4, max_bounces=1):
    """The reference's near-plane test scene (test_single_triangle_clipped.py):
    one vertex of the triangle lies behind the camera, so two of its edges
    cross the near plane and are clipped before projection."""
    cam = make_camera([0, 0, -5], [0, 0, 0], [0, 1, 0], 45.0, resolution)
    return _triangle_and_quad_light([[-1.3, 1.0, 0.0], [1.0, 1.0, 0.0], [-0.5, -2.0, -7.0]], -7.0, 20.0, cam,
                                    resolution, num_samples, max_bounces, "single_triangle_clipped")


def shadow(seed=2, resolution=(256, 256), num_samples=16, max_bounces=1, light_intensity=10.0):
    rng = np.random.default_rng(seed)
    gv = np.array([[-5, 0, -5], [-5, 0, 5], [5, 0, -5], [5, 0, 5]], np.float32)
    gi = np.array([[0, 1, 2], [1, 3, 2]], np.int32)
    xz = rng.uniform(-1.0, 1.0, size=(3, 2))
    y = rng.uniform(1.0, 2.0, size=(3, 1))
    bv = np.concatenate([xz[:, :1], y, xz[:, 1:]], axis=1).astype(np.float32)
    bi = _orient(bv, np.array([[0, 1, 2]], np.int32), np.array([0, 1.0, 0]))
    lv = np.array([[-0.1, 5, -0.1], [-0.1, 5, 0.1], [0.1, 5, -0.1], [0.1, 5, 0.1]], np.float32)
    li = np.array([[0, 2, 1], [1, 2, 3]], np.int32)
    cam = make_camera([0, 2, -5], [0, 0, 0], [0, 1, 0], 45.0, resolution)
    grey = diffuse_material([0.5, 0.5, 0.5])
    black = diffuse_material([0.0, 0.0, 0.0])
    ground = dr.Shape(gv, gi, None, None, grey, None)
    blocker = dr.Shape(bv, bi, None, None, grey, None)
    light_shape = dr.Shape(lv, li, None, None, black, None)
    light = dr.Light(light_shape, np.full(3, light_intensity, np.float32))
    light_shape.light = light
    return dict(name="shadow", camera=cam, shapes=[ground, blocker, light_shape], materials=[grey, black],
                lights=[light], resolution=tuple(resolution), num_samples=num_samples,
                max_bounces=max_bounces, params=[("shape", 1, "vertices")])


def value_noise_fbm(n, rng, octaves=6, lacunarity=2.0, gain=0.5, base=4):
    """fBm of bilinearly interpolated value noise on an n x n grid in [0,1]^2."""
    xs = np.linspace(0.0, 1.0, n)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    out = np.zeros((n, n))
    amp, freq, norm = 1.0, float(base), 0.0
    for _ in range(octaves):
        g = int(math.ceil(freq)) + 2
        lattice = rng.uniform(-1.0, 1.0, size=(g, g))
        fx, fy = X * freq, Y * freq
        ix, iy = np.floor(fx).astype(int), np.floor(fy).astype(int)
        tx, ty = fx - ix, fy - iy
        sx, sy = tx * tx * (3 - 2 * tx), ty * ty * (3 - 2 * ty)
        a = lattice[ix, iy]
        b = lattice[ix + 1, iy]
        c = lattice[ix, iy + 1]
        d = lattice[ix + 1, iy + 1]
        out += amp * ((a * (1 - sx) + b * sx) * (1 - sy) + (c * (1 - sx) + d * sx) * sy)
        norm += amp
        amp *= gain
        freq *= lacunarity
    return out / norm


def vertex_normals(v, ind):
    """Angle-weighted vertex normals (Thurmer & Wuthrich 1998), the weighting
    of the reference's heightfield_gan.py:55-83."""
    v = v.astype(np.float64)
    nrm = np.zeros_like(v)
    p = [v[ind[:, 0]], v[ind[:, 1]], v[ind[:, 2]]]
    fn = np.cross(p[1] - p[0], p[2] - p[0])
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
    for i in range(3):
        a = p[(i + 1) % 3] - p[i]
        b = p[(i + 2) % 3] - p[i]
        a /= np.linalg.norm(a, axis=1, keepdims=True)
        b /= np.linalg.norm(b, axis=1, keepdims=True)
        ang = np.arccos(np.clip(np.sum(a * b, axis=1), -1.0, 1.0))
        np.add.at(nrm, ind[:, i], fn * ang[:, None])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    return nrm.astype(np.float32)


def heightfield(seed=3, grid=128, resolution=(512, 512), num_samples=64, max_bounces=2,
                amplitude=0.2, extent=5.0, light_intensity=50.0, light_size=1.0):
    rng = np.random.default_rng(seed)
    h = amplitude * value_noise_fbm(grid, rng)
    xs = np.linspace(-extent, extent, grid)
    X, Z = np.meshgrid(xs, xs, indexing="ij")
    v = np.stack([X, h, Z], axis=-1).reshape(-1, 3).astype(np.float32)
    uu = np.linspace(0.0, 1.0, grid)
    U, V = np.meshgrid(uu, uu, indexing="ij")
    uvs = np.stack([U, V], axis=-1).reshape(-1, 2).astype(np.float32)
    i = np.arange(grid - 1)
    I, J = np.meshgrid(i, i, indexing="ij")
    a = (I * grid + J).reshape(-1)
    b = a + 1
    c = a + grid
    d = c + 1
    ind = np.concatenate([np.stack([a, b, c], 1), np.stack([b, d, c], 1)], 0).astype(np.int32)
    ind = _orient_fast(v, ind, np.array([0.0, 1.0, 0.0]))
    nrm = vertex_normals(v, ind)
    f0 = rng.uniform(0.04, 1.0)
    rough = rng.uniform(0.05, 0.5, size=(grid, grid, 1)).astype(np.float32)
    one = np.ones(2, np.float32)
    mat = dr.Material(np.full(3, 0.5 * (1.0 - f0), np.float32), np.full(3, f0, np.float32), rough,
                      one, one, one, False)
    black = diffuse_material([0.0, 0.0, 0.0])
    hf = dr.Shape(v, ind, uvs, nrm, mat, None)
    shapes = [hf]
    lights = []
    for sx in (-3.0, 3.0):
        c = np.array([sx, 6.0, 0.0])
        lv, li = quad(c, -c, light_size)
        ls = dr.Shape(lv, li, None, None, black, None)
        l = dr.Light(ls, np.full(3, light_intensity, np.float32))
        ls.light = l
        shapes.append(ls)
        lights.append(l)
    cam = make_camera([0, 3, -6], [0, 0, 0], [0, 1, 0], 45.0, resolution)
    return dict(name="heightfield", camera=cam, shapes=shapes, materials=[mat, black], lights=lights,
                resolution=tuple(resolution), num_samples=num_samples, max_bounces=max_bounces,
                params=[("shape", 0, "vertices"), ("material", 0, "roughness")])


def _orient_fast(v, ind, want):
    p0, p1, p2 = v[ind[:, 0]], v[ind[:, 1]], v[ind[:, 2]]
    flip = np.sum(np.cross(p1 - p0, p2 - p0) * want, axis=1) < 0
    out = ind.copy()
    out[flip] = out[flip][:, [0, 2, 1]]
    return out


def icosphere(level):
    """Unit icosphere: 20 * 4**level triangles (outward winding), with its
    vertex normals (the positions)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    v = [np.asarray(p, np.float64) / np.linalg.norm(p) for p in v]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    for _ in range(level):
        mid = {}

        def midpoint(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                mid[k] = len(v) - 1
            return mid[k]
        nf = []
        for a, b, c in f:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = nf
    V = np.asarray(v, np.float32)
    F = np.asarray(f, np.int32)
    p0, p1, p2 = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    flip = np.sum(np.cross(p1 - p0, p2 - p0) * (p0 + p1 + p2), axis=1) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return V, F, V.copy()


def random_affine(rng):
    """Rotation * anisotropic scale with det in [0.5, 2], as a 3x3 matrix."""
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    while True:
        sc = rng.uniform(0.7, 1.3, size=3)
        if 0.5 <= np.prod(sc) <= 2.0:
            return R @ np.diag(sc)


def batched(seed=4, num_scenes=64, views=4, min_tris=1000, max_tris=200000, res_choices=(128, 192, 256, 384, 512),
            spp_choices=(16, 32, 64), max_bounces=2):
    """Uneven batch: scenes of log-uniform triangle counts (subdivided
    icospheres or fBm heightfields under random affine maps, on a ground
    quad), 1-4 area lights, diffuse or glossy materials, and `views` cameras
    per scene at random resolutions and sample counts.  Returns a list of
    scenes, each a dict of workload fields plus "views": [(camera,
    resolution, spp)]."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(num_scenes):
        target = float(np.exp(rng.uniform(np.log(min_tris), np.log(max_tris))))
        one = np.ones(2, np.float32)
        glossy = rng.uniform() < 0.5
        if glossy:
            f0 = rng.uniform(0.04, 1.0)
            mat = dr.Material(np.full(3, 0.5 * (1.0 - f0), np.float32), np.full(3, f0, np.float32),
                              np.array([rng.uniform(0.05, 0.5)], np.float32), one, one, one, False)
        else:
            mat = diffuse_material(rng.uniform(0.2, 0.8, size=3))
        A = random_affine(rng)
        if rng.uniform() < 0.5:
            # icosphere level closest to the target count (20 * 4**level)
            level = int(np.clip(np.round(np.log(target / 20.0) / np.log(4.0)), 3, 7))
            level = min(level, 6) if target < 3e5 else level
            V, F, N = icosphere(level)
            v = (V.astype(np.float64) @ A.T + np.array([0.0, 1.2, 0.0])).astype(np.float32)
            nrm = (N.astype(np.float64) @ np.linalg.inv(A)).astype(np.float32)
            nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
            obj = dr.Shape(v, F, None, nrm, mat, None)
        else:
            grid = int(np.clip(np.round(np.sqrt(target / 2.0)) + 1, 23, 317))
            h = heightfield(seed=int(rng.integers(1 << 30)), grid=grid, resolution=(64, 64), num_samples=1)
            hv = h["shapes"][0].vertices.astype(np.float64) * np.array([0.3, 1.0, 0.3])
            v = (hv @ A.T + np.array([0.0, 1.0, 0.0])).astype(np.float32)
            nrm = vertex_normals(v, h["shapes"][0].indices)
            obj = dr.Shape(v, h["shapes"][0].indices, h["shapes"][0].uvs, nrm, mat, None)
        grey = diffuse_material([0.5, 0.5, 0.5])
        gv, gi = quad([0.0, -1.0, 0.0], [0.0, 1.0, 0.0], 12.0)
        ground = dr.Shape(gv, gi, None, None, grey, None)
        black = diffuse_material([0.0, 0.0, 0.0])
        shapes, lights = [obj, ground], []
        for _ in range(int(rng.integers(1, 5))):
            ang = rng.uniform(0, 2 * np.pi)
            c = np.array([3.0 * np.cos(ang), rng.uniform(4.0, 7.0), 3.0 * np.sin(ang)])
            lv, li = quad(c, -c, rng.uniform(0.5, 1.5))
            ls = dr.Shape(lv, li, None, None, black, None)
            l = dr.Light(ls, np.full(3, rng.uniform(20.0, 60.0), np.float32))
            ls.light = l
            shapes.append(ls)
            lights.append(l)
        cams = []
        for _ in range(views):
            res = int(rng.choice(res_choices))
            spp = int(rng.choice(spp_choices))
            ang = rng.uniform(0, 2 * np.pi)
            pos = [6.0 * np.cos(ang), rng.uniform(2.0, 4.0), 6.0 * np.sin(ang)]
            cams.append((make_camera(pos, [0, 0.8, 0], [0, 1, 0], 45.0, (res, res)), (res, res), spp))
        out.append(dict(name="batched_%d" % k, camera=cams[0][0], shapes=shapes, materials=[mat, grey, black],
                        lights=lights, views=cams, max_bounces=max_bounces,
                        num_triangles=int(sum(np.asarray(s.indices).shape[0] for s in shapes))))
    return out


def from_scene(camera, materials, shapes, lights, resolution, num_samples=4, max_bounces=2, name="scene"):
    """A workload from scene classes (scene.Camera / Material / Shape / Light,
    as load_mitsuba returns them).  The camera's matrices depend on its
    aspect ratio only, so `resolution` may differ from the file's if the
    aspect ratio is kept."""
    def arr(t, dt=np.float32):
        return None if t is None else np.asarray(t.detach().cpu().numpy() if hasattr(t, "detach") else t, dt)
    if abs(resolution[0] / resolution[1] - camera.aspect_ratio) > 1e-6:
        raise ValueError("resolution %s does not keep the camera's aspect ratio %g" % (resolution,
                                                                                       camera.aspect_ratio))
    cam = dr.Camera(arr(camera.cam_to_world), arr(camera.world_to_cam), arr(camera.sample_to_cam),
                    arr(camera.cam_to_sample), float(arr(camera.fov_factor).reshape(-1)[0]),
                    float(camera.aspect_ratio), float(arr(camera.clip_near).reshape(-1)[0]), bool(camera.fisheye))
    mats = [dr.Material(arr(m.diffuse_reflectance), arr(m.specular_reflectance), arr(m.roughness),
                        arr(m.diffuse_uv_scale), arr(m.specular_uv_scale), arr(m.roughness_uv_scale),
                        bool(m.two_sided)) for m in materials]
    out = []
    for s in shapes:
        if not 0 <= int(s.mat_id) < len(mats):
            raise ValueError("a shape has no material (material id %d)" % int(s.mat_id))
        out.append(dr.Shape(arr(s.vertices), arr(s.indices, np.int32), arr(s.uvs), arr(s.normals),
                            mats[int(s.mat_id)], None))
    ls = []
    for l in lights:
        light = dr.Light(out[int(l.shape_id)], arr(l.intensity))
        out[int(l.shape_id)].light = light
        ls.append(light)
    return dict(name=name, camera=cam, shapes=out, materials=mats, lights=ls, resolution=tuple(resolution),
                num_samples=num_samples, max_bounces=max_bounces,
                params=[("shape", i, "vertices") for i in range(len(out))])


def mitsuba(path, resolution=None, num_samples=4, max_bounces=2):
    """A workload from a Mitsuba scene file (load_mitsuba), at the file's
    resolution or another one of the same aspect ratio."""
    from .load_mitsuba import load_mitsuba
    import os
    camera, materials, shapes, lights, res = load_mitsuba(path)
    return from_scene(camera, materials, shapes, lights, tuple(resolution or res), num_samples, max_bounces,
                      name=os.path.splitext(os.path.basename(path))[0])


WORKLOADS = {"single_triangle": single_triangle, "shadow": shadow, "heightfield": heightfield}

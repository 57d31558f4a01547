"""Closed-form quantities the oracle's edge terms are pinned against.

- ``pixel_area_loss``: an emitting triangle of radiance I seen by the
  camera renders, in pixel p, I * W * H * area(T_s ∩ pixel_p) where T_s is the
  triangle projected to the [0,1]^2 screen (the pixel filter is a box over the
  pixel).  L = sum d_p . I_p is therefore a closed form of the screen vertices,
  and its derivative is what the primary-edge term must estimate.
- ``occlusion_grad``: a diffuse ground point p under a square emitter of
  radiance I, partially shadowed by a triangle, has radiance
  rho/pi * I * (Omega(light) - Omega(light ∩ shadow(p))), Omega the projected
  solid angle (Lambert's polygon formula).  The blocker's vertices enter only
  through Omega(light ∩ shadow), whose derivative is what the secondary-edge
  term must estimate.  The derivative is split per blocker edge so that the
  reference estimator's per-edge Dirac scaling can be predicted exactly.

Everything is float64 NumPy; nothing here calls the oracle or the library.
"""
import numpy as np


def clip(poly, f):
    """Sutherland-Hodgman: keep the part of convex polygon `poly` (list of 2D
    points) where the affine function f >= 0."""
    out = []
    n = len(poly)
    for i in range(n):
        p, q = poly[i], poly[(i + 1) % n]
        fp, fq = f(p), f(q)
        if fp >= 0.0:
            out.append(p)
        if (fp >= 0.0) != (fq >= 0.0):
            t = fp / (fp - fq)
            out.append((p[0] + t * (q[0] - p[0]), p[1] + t * (q[1] - p[1])))
    return out


def polygon_area(poly):
    if len(poly) < 3:
        return 0.0
    a = 0.0
    for i in range(len(poly)):
        p, q = poly[i], poly[(i + 1) % len(poly)]
        a += p[0] * q[1] - q[0] * p[1]
    return abs(a) * 0.5


def screen_points(camera, verts):
    """World points -> [0,1]^2 screen with the camera model of camera.h
    (project_local): x = (-l.x / (l.z fov) + 1) / 2, y = (-l.y / (l.z fov
    aspect) + 1) / 2, l = world_to_cam * p."""
    w2c = np.asarray(camera.world_to_cam, np.float64)
    fov = float(camera.fov_factor)
    asp = float(camera.aspect_ratio)
    out = []
    for v in np.asarray(verts, np.float64):
        l = w2c @ np.r_[v, 1.0]
        out.append(((-l[0] / (l[2] * fov) + 1.0) * 0.5, (-l[1] / (l[2] * fov * asp) + 1.0) * 0.5))
    return out


def pixel_area_loss(camera, verts, d_image, radiance):
    """sum_p d_p . radiance * W * H * area(T_s ∩ pixel_p)."""
    H, W, _ = d_image.shape
    T = screen_points(camera, verts)
    xs = [p[0] for p in T]
    ys = [p[1] for p in T]
    L = 0.0
    for y in range(max(0, int(np.floor(min(ys) * H))), min(H, int(max(ys) * H) + 1)):
        for x in range(max(0, int(np.floor(min(xs) * W))), min(W, int(max(xs) * W) + 1)):
            poly = clip(T, lambda q: q[0] - x / W)
            poly = clip(poly, lambda q: (x + 1) / W - q[0])
            poly = clip(poly, lambda q: q[1] - y / H)
            poly = clip(poly, lambda q: (y + 1) / H - q[1])
            a = polygon_area(poly)
            if a > 0.0:
                L += a * W * H * float(np.dot(d_image[y, x].astype(np.float64), radiance))
    return L


def central_diff(fn, x, eps):
    """Central differences of scalar fn over every entry of array x."""
    x = np.array(x, np.float64)
    g = np.zeros(x.size)
    for i in range(x.size):
        xp, xm = x.copy(), x.copy()
        xp.flat[i] += eps
        xm.flat[i] -= eps
        g[i] = (fn(xp) - fn(xm)) / (2.0 * eps)
    return g.reshape(x.shape)


def projected_solid_angle(p, n, poly3):
    """Lambert's formula: (1/2) |sum_k angle(r_k, r_k+1) n . unit(r_k x r_k+1)|."""
    R = [(q - p) / np.linalg.norm(q - p) for q in poly3]
    s = 0.0
    for k in range(len(R)):
        a, b = R[k], R[(k + 1) % len(R)]
        c = np.cross(a, b)
        cn = np.linalg.norm(c)
        if cn > 0.0:
            s += np.arctan2(cn, np.dot(a, b)) * np.dot(n, c / cn)
    return 0.5 * abs(s)


BLOCKER_EDGES = ((0, 1), (1, 2), (2, 0))


def _shadow_omega(p, planes, inner, light_y, light_half):
    """Projected solid angle (normal +y) of the part of the square emitter
    {|x|,|z| <= light_half, y = light_y} inside every wedge plane (p, va, vb)
    on the side of `inner`."""
    poly = [(-light_half, -light_half), (light_half, -light_half), (light_half, light_half),
            (-light_half, light_half)]
    for va, vb in planes:
        N = np.cross(va - p, vb - p)
        sgn = np.sign(np.dot(inner - p, N))
        poly = clip(poly, lambda q, N=N, sgn=sgn: sgn * np.dot(np.array([q[0], light_y, q[1]]) - p, N))
        if len(poly) < 3:
            return 0.0
    return projected_solid_angle(p, np.array([0.0, 1.0, 0.0]),
                                 [np.array([x, light_y, z]) for x, z in poly])


def ground_points(camera, W, H, n):
    """Midpoint quadrature (n x n per pixel) of the ground plane y = 0 as seen
    through each pixel: camera.h sample_primary's ray through the screen point,
    intersected with y = 0.  Yields (x, y, point)."""
    c2w = np.asarray(camera.cam_to_world, np.float64)
    fov = float(camera.fov_factor)
    asp = float(camera.aspect_ratio)
    org = c2w[:3, 3]
    for y in range(H):
        for x in range(W):
            for j in range(n):
                for i in range(n):
                    sx = (x + (i + 0.5) / n) / W
                    sy = (y + (j + 0.5) / n) / H
                    dc = np.array([fov * (sx - 0.5) * -2.0, fov * (sy - 0.5) * -2.0 / asp, 1.0])
                    dw = c2w[:3, :3] @ dc
                    t = -org[1] / dw[1]
                    yield x, y, org + t * dw


def occlusion_grad(camera, W, H, blocker, d_image, rho, radiance, light_y, light_half, n=2, eps=1e-6):
    """dL/d(blocker vertices) of L = sum_p d_p . I_p for the diffuse ground
    under the partially blocked square emitter, split per blocker edge.

    Returns (exact, reference_dirac): the exact derivative, and the
    expectation of an edge estimator that divides by the full 3-D gradient
    |d0 x d1| of the edge-plane equation instead of its part tangential to
    the emitter (each edge scaled by sqrt(1 - (n_plane . y)^2))."""
    B = np.asarray(blocker, np.float64)
    inner = B.mean(axis=0)
    exact = np.zeros(9)
    biased = np.zeros(9)
    w_pix = [[float(np.dot(d_image[y, x].astype(np.float64), rho / np.pi * radiance)) / (n * n)
              for x in range(W)] for y in range(H)]
    for x, y, p in ground_points(camera, W, H, n):
        wp = w_pix[y][x]
        for k, (a, b) in enumerate(BLOCKER_EDGES):
            ne = np.cross(B[a] - p, B[b] - p)
            fac = np.sqrt(max(0.0, 1.0 - (ne[1] / np.linalg.norm(ne)) ** 2))
            for vi in (a, b):
                for c in range(3):
                    vals = []
                    for sg in (1.0, -1.0):
                        Bp = B.copy()
                        Bp[vi, c] += sg * eps
                        planes = [(B[aa], B[bb]) for aa, bb in BLOCKER_EDGES]
                        planes[k] = (Bp[a], Bp[b])
                        vals.append(_shadow_omega(p, planes, inner, light_y, light_half))
                    g = -wp * (vals[0] - vals[1]) / (2.0 * eps)
                    exact[vi * 3 + c] += g
                    biased[vi * 3 + c] += fac * g
    return exact, biased

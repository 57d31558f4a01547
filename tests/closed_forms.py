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


# ------------------------------------------------------- two-bounce light --
def gauss_panels(a, b, panels, order=8, grade=None):
    """Composite Gauss-Legendre nodes and weights on [a, b]: `panels` panels,
    uniform, or graded geometrically toward `grade` (= a or b) so that an
    integrand varying on the scale of the distance to that end is resolved."""
    x, w = np.polynomial.legendre.leggauss(order)
    if grade is None:
        edges = np.linspace(a, b, panels + 1)
    else:
        # panel ends at distances L r^k from the graded end, r = 0.6
        L = abs(b - a)
        d = np.concatenate([[0.0], L * 0.6 ** np.arange(panels - 1, -1, -1)])
        edges = grade + np.sign((a + b) / 2 - grade) * d
        edges = np.sort(edges)
    xs, ws = [], []
    for lo, hi in zip(edges[:-1], edges[1:]):
        xs.append(0.5 * (hi - lo) * x + 0.5 * (hi + lo))
        ws.append(0.5 * (hi - lo) * w)
    return np.concatenate(xs), np.concatenate(ws)


def rect_quadrature(origin, eu, ev, nu, nv, grade_u=None, grade_v=None, order=8):
    """Nodes (N, 3) and area weights (N,) of the rectangle origin + s eu + t ev,
    s, t in [0, 1] (composite Gauss-Legendre, optionally graded toward s or
    t = 0 / 1)."""
    su, wu = gauss_panels(0.0, 1.0, nu, order, grade_u)
    sv, wv = gauss_panels(0.0, 1.0, nv, order, grade_v)
    S, T = np.meshgrid(su, sv, indexing="ij")
    W = np.outer(wu, wv) * np.linalg.norm(np.cross(eu, ev))
    P = origin[None, :] + S.reshape(-1, 1) * eu[None, :] + T.reshape(-1, 1) * ev[None, :]
    return P, W.reshape(-1)


def lambert_direct(P, n, rho, light_nodes, light_w, light_n, I):
    """Reflected radiance of a Lambertian point set P (normal n) lit by an
    unoccluded one-sided emitter (nodes, area weights, normal, radiance I):
    rho/pi * sum_k I cos_p cos_l / r^2 w_k, with the BSDF's cut cos_p <= 5e-3
    (material.h:271-273)."""
    d = light_nodes[None, :, :] - P[:, None, :]
    r2 = np.sum(d * d, axis=2)
    r = np.sqrt(r2)
    cp = np.sum(d * n, axis=2) / r
    cl = -np.sum(d * light_n, axis=2) / r
    f = np.where((cp > 5e-3) & (cl > 0.0), cp * cl / r2, 0.0)
    return (rho / np.pi) * I * np.sum(f * light_w[None, :], axis=1)


def lambert_gather(X, nx, rho, P, nP, LP, wP):
    """One diffuse bounce: radiance at points X (normal nx, albedo rho) from
    surface nodes P (normal nP, area weights wP) of outgoing radiance LP,
    rho/pi sum L cos_x cos_p / r^2 w (cut cos_x <= 5e-3; P fully visible)."""
    d = P[None, :, :] - X[:, None, :]
    r2 = np.sum(d * d, axis=2)
    r = np.sqrt(r2)
    cx = np.sum(d * nx, axis=2) / r
    cp = -np.sum(d * nP, axis=2) / r
    f = np.where((cx > 5e-3) & (cp > 0.0), cx * cp / r2, 0.0)
    return (rho / np.pi) * np.sum(f * (LP * wP)[None, :], axis=1)


def camera_rays(camera, W, H, x, y, n, gauss=False):
    """n x n sub-pixel rays of pixel (x, y) with camera.h sample_primary's
    pinhole model: (origin, directions (n*n, 3)); midpoints, or with gauss
    the Gauss-Legendre nodes (equal weights per axis are then not equal:
    use gauss_pixel_weights)."""
    c2w = np.asarray(camera.cam_to_world, np.float64)
    fov = float(camera.fov_factor)
    asp = float(camera.aspect_ratio)
    if gauss:
        k = 0.5 * (np.polynomial.legendre.leggauss(n)[0] + 1.0)
    else:
        k = (np.arange(n) + 0.5) / n
    sx, sy = np.meshgrid((x + k) / W, (y + k) / H, indexing="xy")
    dc = np.stack([fov * (sx.ravel() - 0.5) * -2.0, fov * (sy.ravel() - 0.5) * -2.0 / asp,
                   np.ones(n * n)], axis=1)
    dw = dc @ c2w[:3, :3].T
    return c2w[:3, 3], dw / np.linalg.norm(dw, axis=1, keepdims=True)


def gauss_pixel_weights(n):
    """weights (n*n,) of camera_rays(..., gauss=True), summing to 1"""
    w = 0.5 * np.polynomial.legendre.leggauss(n)[1]
    return np.outer(w, w).ravel()


# ------------------------------------------------ glossy occluded emitter --
def triangle_rule(order=8):
    """Collapsed Gauss-Legendre rule on the unit triangle (barycentric a, b):
    nodes (N, 2) and weights summing to 1/2."""
    x, w = np.polynomial.legendre.leggauss(order)
    x = 0.5 * (x + 1.0)
    w = 0.5 * w
    U, V = np.meshgrid(x, x, indexing="ij")
    WU, WV = np.meshgrid(w, w, indexing="ij")
    a = U
    b = V * (1.0 - U)
    return np.stack([a.ravel(), b.ravel()], 1), (WU * WV * (1.0 - U)).ravel()


def polygon_nodes(poly3, rule):
    """Quadrature nodes / area weights of a convex planar polygon (fan from
    its first vertex)."""
    tn, tw = rule
    P, Wt = [], []
    for k in range(1, len(poly3) - 1):
        a, b, c = poly3[0], poly3[k], poly3[k + 1]
        area2 = np.linalg.norm(np.cross(b - a, c - a))
        P.append(a[None, :] + tn[:, :1] * (b - a)[None, :] + tn[:, 1:] * (c - a)[None, :])
        Wt.append(tw * area2)
    if not P:
        return np.zeros((0, 3)), np.zeros(0)
    return np.concatenate(P), np.concatenate(Wt)


def occluded_light_polygon(p, blocker, light_y, light_half):
    """The part of the square emitter {|x|, |z| <= light_half, y = light_y}
    that the blocker triangle hides from p (3-D points)."""
    B = np.asarray(blocker, np.float64)
    inner = B.mean(axis=0)
    poly = [(-light_half, -light_half), (light_half, -light_half), (light_half, light_half),
            (-light_half, light_half)]
    for a, b in BLOCKER_EDGES:
        N = np.cross(B[a] - p, B[b] - p)
        sgn = np.sign(np.dot(inner - p, N))
        poly = clip(poly, lambda q, N=N, sgn=sgn: sgn * np.dot(np.array([q[0], light_y, q[1]]) - p, N))
        if len(poly) < 3:
            return []
    return [np.array([x, light_y, z]) for x, z in poly]


def glossy_occlusion_grad(camera, W, H, blocker, d_image, bsdf_at, I, light_y, light_half, n=2, eps=1e-5,
                          order=8):
    """dL/d(blocker vertices) of L = sum_p d_p . I_p for a ground point seen by
    the camera under a square emitter (facing -y) partly hidden by the blocker:
    I_p = integral over the visible emitter of f(wi, w) I cos_l / r^2 dA, f the
    BSDF with its cosine (bsdf_at(p, wi, W) -> (N, 3)).  Only the hidden part
    depends on the blocker, so dL = - sum d_p . d/dB integral over the hidden
    polygon, by central differences of a fixed-order quadrature on it."""
    B0 = np.asarray(blocker, np.float64)
    rule = triangle_rule(order)
    c2w = np.asarray(camera.cam_to_world, np.float64)
    cam = c2w[:3, 3]
    grad = np.zeros(9)

    def hidden(p, wi, B):
        poly = occluded_light_polygon(p, B, light_y, light_half)
        if len(poly) < 3:
            return np.zeros(3)
        X, w = polygon_nodes(poly, rule)
        d = X - p[None, :]
        r2 = np.sum(d * d, axis=1)
        dirs = d / np.sqrt(r2)[:, None]
        cl = dirs[:, 1]  # emitter normal (0, -1, 0): cos = -n . (-dir)
        f = bsdf_at(p, wi, dirs)
        return np.sum(f * (I[None, :] * (cl / r2 * w)[:, None]), axis=0)

    for x, y, p in ground_points(camera, W, H, n):
        wi = (cam - p) / np.linalg.norm(cam - p)
        dw = d_image[y, x].astype(np.float64) / (n * n)
        for vi in range(3):
            for c in range(3):
                Bp, Bm = B0.copy(), B0.copy()
                Bp[vi, c] += eps
                Bm[vi, c] -= eps
                g = (hidden(p, wi, Bp) - hidden(p, wi, Bm)) / (2.0 * eps)
                grad[3 * vi + c] -= float(np.dot(dw, g))
    return grad


# ------------------------------------------------------------ node bound --
def point_node_bound(p, n, wi, q, dlum, slum, exponent):
    """edge_tree.cpp:345-480 for a node box shrunk to the point q, written
    without its half-vector frame: the box's polar bound degenerates to
    cos t' = sqrt((c + 1) / 2) with c = (q - p) . wi / |q - p|^2 (the
    reference divides by the squared length, edge_tree.cpp:425-447), its
    azimuth bound to the azimuth of q; the half-vector cosine bound is then
    min(|sin t' (n . u)| + cos t' (n . wi), 1), u the unit part of q - p
    orthogonal to wi."""
    v = np.asarray(q, np.float64) - p
    if np.dot(v, n) <= 0.0:
        return 0.0
    out = dlum / np.pi
    if slum > 0.0:
        c = np.dot(v, wi) / np.dot(v, v)
        ct = np.sqrt((c + 1.0) / 2.0)
        st = np.sqrt(max(1.0 - ct * ct, 0.0))
        u = v - np.dot(v, wi) * wi
        u = u / np.linalg.norm(u)
        h = min(abs(st * np.dot(n, u)) + ct * np.dot(n, wi), 1.0)
        out += slum * h ** exponent * (exponent + 2.0) / (2.0 * np.pi)
    return out

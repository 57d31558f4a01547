"""The CPU oracle's single functions vs the reference's own code (CPU only).

oracle/_ref/refprobe is compiled by oracle/Makefile.ref from the reference's
Embree-free headers (intersect.h, material.h + ltc.inc, camera.h, aabb.cpp,
line_clip.h, vector.h, distribution.h, sample.h) where they lie.  It exists
only in a container holding the reference sources; elsewhere these tests
skip.  Inputs are seeded random batches plus the special cases each function
has (parallel rays, below-horizon directions, specular roughness, clipped
segments, all-zero distributions, boxes containing the cone apex).
"""
import ctypes
import importlib
import os
import sys

import numpy as np
import pytest

from delta_ray_amd import workloads
from delta_ray_amd import delta_ray as dr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")


@pytest.fixture(scope="module")
def ref():
    if not os.path.isdir(REF_DIR):
        pytest.skip("oracle/_ref is not built (needs the reference sources)")
    sys.path.insert(0, REF_DIR)
    try:
        return importlib.import_module("refprobe")
    except ImportError:
        pytest.skip("oracle/_ref/refprobe is not importable")
    finally:
        sys.path.remove(REF_DIR)


@pytest.fixture(scope="module")
def olib(oracle):
    lib = oracle.load()
    fp = ctypes.POINTER(ctypes.c_float)
    for name, args, res in [
        ("or_probe_bsdf", [ctypes.c_void_p, ctypes.c_int, fp, fp, fp, fp], None),
        ("or_probe_sample_bsdf", [ctypes.c_void_p, ctypes.c_int, fp, fp, fp, fp], None),
        ("or_probe_ltc", [ctypes.c_void_p, ctypes.c_int, fp, fp, fp], None),
        ("or_probe_sample_primary", [ctypes.c_void_p, fp, fp], None),
        ("or_probe_project", [ctypes.c_void_p, fp, fp, fp], ctypes.c_int),
        ("or_probe_d_project", [ctypes.c_void_p, fp, fp, fp, fp], None),
        ("or_probe_camera", [ctypes.c_void_p, fp, fp, fp, fp], None),
        ("or_probe_cone_box", [fp, fp, fp, fp, ctypes.c_float], ctypes.c_int),
        ("or_probe_clip_line", [fp, fp, fp], ctypes.c_int),
        ("or_probe_jacobians", [fp, fp, fp, fp, fp], None),
        ("or_probe_mt", [fp, fp, fp, fp, fp, fp], None),
        ("or_probe_sample_discrete", [fp, ctypes.c_int, ctypes.c_float, fp], ctypes.c_int),
        ("or_probe_sample_triangle", [fp, fp, fp, fp, fp], None),
    ]:
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    return lib


def F(a):
    a = np.ascontiguousarray(a, np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def call(fn, *arrays, n_out=0, extra=()):
    keep = [F(a) for a in arrays]
    out, po = F(np.zeros(max(n_out, 1), np.float32))
    r = fn(*[k[1] for k in keep], *extra, po) if n_out else fn(*[k[1] for k in keep], *extra)
    return (r, out) if n_out else r


def unit(rng, n=None):
    v = rng.normal(size=(3,) if n is None else (n, 3))
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def frame_of(n):
    """An orthonormal frame (s, t, n) with the given normal (vector.h
    make_frame conventions are not needed: both sides get the same frame)."""
    a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    s = np.cross(a, n)
    s /= np.linalg.norm(s)
    t = np.cross(n, s)
    return np.stack([s, t, n])


def test_moller_trumbore_bit_exact(ref, olib):
    """intersect.h:71-90, including the divisor == 0 replacement."""
    rng = np.random.default_rng(1)
    for k in range(600):
        v = rng.normal(size=(3, 3)).astype(np.float32)
        org = rng.normal(size=3).astype(np.float32) * 3
        if k % 10 == 0:   # ray parallel to the triangle's plane: divisor 0
            d = (v[1] - v[0]).astype(np.float32)
        else:
            d = unit(rng).astype(np.float32)
        _, got = call(olib.or_probe_mt, v[0], v[1], v[2], org, d, n_out=3)
        exp = np.array(ref.mt(*[x.tolist() for x in (v[0], v[1], v[2], org, d)]), np.float32)
        np.testing.assert_array_equal(got.view(np.uint32), exp.view(np.uint32))


def _material_cases():
    one = np.ones(2, np.float32)
    rng = np.random.default_rng(7)
    cases = []
    for spec, rough, two in [((0.0, 0.0, 0.0), 1.0, False), ((0.5, 0.4, 0.3), 0.3, False),
                             ((0.8, 0.8, 0.8), 0.008, False), ((0.2, 0.6, 0.9), 0.05, True),
                             ((0.04, 0.04, 0.04), 0.9, True)]:
        cases.append(("const", np.array([0.5, 0.3, 0.2], np.float32), np.array(spec, np.float32),
                      np.array([rough], np.float32), one, one, one, two))
    tex_d = rng.uniform(0.0, 0.8, size=(6, 5, 3)).astype(np.float32)
    tex_s = rng.uniform(0.0, 0.8, size=(4, 7, 3)).astype(np.float32)
    tex_r = rng.uniform(0.02, 0.6, size=(5, 5, 1)).astype(np.float32)
    cases.append(("textured", tex_d, tex_s, tex_r, np.array([2.0, 3.0], np.float32),
                  np.array([1.5, 0.5], np.float32), np.array([1.0, 4.0], np.float32), False))
    return cases


@pytest.fixture(scope="module")
def material_scenes(oracle):
    """One oracle scene per material case (material id 0)."""
    out = []
    cam = workloads.make_camera([0, 0, 5], [0, 0, 0], [0, 1, 0], 45.0, (8, 8))
    for case in _material_cases():
        _, d, s, r, ud, us, ur, two = case
        m = dr.Material(d, s, r, ud, us, ur, two)
        tv = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
        ti = np.array([[0, 1, 2]], np.int32)
        uvs = np.array([[0, 0], [1, 0], [0, 1]], np.float32)
        tri = dr.Shape(tv, ti, uvs, None, m, None)
        black = workloads.diffuse_material([0.0, 0.0, 0.0])
        ls = dr.Shape(tv + np.float32(5.0), ti, None, None, black, None)
        light = dr.Light(ls, np.array([1, 1, 1], np.float32))
        ls.light = light
        out.append((case, oracle.OracleScene(cam, [tri, ls], [m, black], [light])))
    return out


def _sp(rng, two_sided_flip=False):
    n = unit(rng)
    fr = frame_of(n)
    pos = rng.normal(size=3)
    uv = rng.uniform(-1.5, 2.5, size=2)
    return np.concatenate([pos, n, fr.ravel(), uv]).astype(np.float32), n


def test_bsdf_and_pdf_match_reference(ref, olib, material_scenes):
    """material.h:257-351 (bsdf) and 414-462 (bsdf_pdf), constant and
    textured parameters, directions above and below the surface."""
    rng = np.random.default_rng(11)
    for case, osc in material_scenes:
        _, d, s, r, ud, us, ur, two = case
        for _ in range(150):
            sp, n = _sp(rng)
            wi = unit(rng).astype(np.float32)
            wo = unit(rng).astype(np.float32)
            if rng.uniform() < 0.7:   # mostly the upper hemisphere
                wi = np.where(np.dot(wi, n) < 0, -wi, wi).astype(np.float32)
                wo = np.where(np.dot(wo, n) < 0, -wo, wo).astype(np.float32)
            keep = [F(x) for x in (sp, wi, wo)]
            out, po = F(np.zeros(4, np.float32))
            olib.or_probe_bsdf(osc.h, 0, keep[0][1], keep[1][1], keep[2][1], po)
            eb = np.array(ref.bsdf(d, s, r, ud, us, ur, two, sp.tolist(), wi.tolist(), wo.tolist()), np.float32)
            ep = np.float32(ref.bsdf_pdf(d, s, r, ud, us, ur, two, sp.tolist(), wi.tolist(), wo.tolist()))
            np.testing.assert_allclose(out[:3], eb, rtol=2e-6, atol=1e-7)
            np.testing.assert_allclose(out[3], ep, rtol=2e-6, atol=1e-7)


def test_sample_bsdf_matches_reference(ref, olib, material_scenes):
    """material.h:353-412: direction and the is_specular flag."""
    rng = np.random.default_rng(12)
    for case, osc in material_scenes:
        _, d, s, r, ud, us, ur, two = case
        for _ in range(150):
            sp, n = _sp(rng)
            wi = unit(rng).astype(np.float32)
            wi = np.where(np.dot(wi, n) < 0, -wi, wi).astype(np.float32)
            bs = rng.uniform(size=3).astype(np.float32)
            keep = [F(x) for x in (sp, wi, bs)]
            out, po = F(np.zeros(4, np.float32))
            olib.or_probe_sample_bsdf(osc.h, 0, keep[0][1], keep[1][1], keep[2][1], po)
            e = np.array(ref.sample_bsdf(d, s, r, ud, us, ur, two, sp.tolist(), wi.tolist(), bs.tolist()),
                         np.float32)
            np.testing.assert_allclose(out, e, rtol=3e-6, atol=3e-7)


def test_ltc_lookup_matches_reference(ref, olib, material_scenes):
    """material.h:464-476: nearest table entry by (roughness, theta),
    including theta = 0 and roughness at both ends of the table."""
    rng = np.random.default_rng(13)
    for case, osc in material_scenes:
        _, d, s, r, ud, us, ur, two = case
        for k in range(100):
            sp, n = _sp(rng)
            if k == 0:
                wi = n.astype(np.float32)
            else:
                wi = unit(rng).astype(np.float32)
                wi = np.where(np.dot(wi, n) < 0, -wi, wi).astype(np.float32)
            keep = [F(x) for x in (sp, wi)]
            out, po = F(np.zeros(9, np.float32))
            olib.or_probe_ltc(osc.h, 0, keep[0][1], keep[1][1], po)
            e = np.array(ref.ltc(d, s, r, ud, us, ur, two, sp.tolist(), wi.tolist()), np.float32)
            np.testing.assert_array_equal(out, e)


def _camera_args(cam):
    return (np.asarray(cam.cam_to_world, np.float32), np.asarray(cam.world_to_cam, np.float32),
            np.asarray(cam.sample_to_cam, np.float32), np.asarray(cam.cam_to_sample, np.float32),
            float(cam.fov_factor), float(cam.aspect_ratio), float(cam.clip_near))


@pytest.fixture(scope="module")
def camera_scene(oracle):
    cam = workloads.make_camera([0.3, 1.0, 4.0], [0, 0.2, 0], [0, 1, 0], 50.0, (64, 64), clip_near=0.05)
    w = workloads.single_triangle(resolution=(64, 64))
    w["camera"] = cam
    return cam, oracle.scene_from_workload(w)


def test_camera_functions_match_reference(ref, olib, camera_scene):
    """camera.h: sample_primary (48-84), project with near-plane clipping
    (216-242) and d_project's world_to_cam gradient (244-313)."""
    cam, osc = camera_scene
    args = _camera_args(cam)
    rng = np.random.default_rng(14)
    for _ in range(300):
        scr = rng.uniform(size=2).astype(np.float32)
        out, po = F(np.zeros(6, np.float32))
        olib.or_probe_sample_primary(osc.h, F(scr)[1], po)
        e = np.array(ref.sample_primary(*args, scr.tolist()), np.float32)
        np.testing.assert_allclose(out, e, rtol=2e-6, atol=2e-7)
    n_clipped = 0
    for k in range(400):
        p0 = (rng.normal(size=3) * 2).astype(np.float32)
        p1 = (rng.normal(size=3) * 2).astype(np.float32)
        k0, k1 = F(p0), F(p1)
        out, po = F(np.zeros(4, np.float32))
        ok = olib.or_probe_project(osc.h, k0[1], k1[1], po)
        e = ref.project(*args, p0.tolist(), p1.tolist())
        assert bool(ok) == (len(e) > 0)
        if ok:
            np.testing.assert_allclose(out, np.array(e, np.float32), rtol=1e-5, atol=1e-6)
            dxy = rng.normal(size=4).astype(np.float32)
            g, pg = F(np.zeros(16, np.float32))
            olib.or_probe_d_project(osc.h, k0[1], k1[1], F(dxy)[1], pg)
            eg = np.array(ref.d_project(*args, p0.tolist(), p1.tolist(), dxy.tolist()), np.float32)
            np.testing.assert_allclose(g, eg, rtol=1e-5, atol=1e-5 * np.abs(eg).max())
            w2c = np.asarray(cam.world_to_cam, np.float64)
            z0 = (w2c @ np.r_[p0, 1])[2]
            z1 = (w2c @ np.r_[p1, 1])[2]
            n_clipped += (z0 < cam.clip_near) != (z1 < cam.clip_near)
    assert n_clipped > 10  # the clipped branches were exercised


@pytest.fixture(scope="module")
def fisheye_scene(oracle):
    cam = workloads.make_camera([0.0, 0.0, -1.0], [0, 0, 0], [0, 1, 0], 45.0, (64, 64), fisheye=True)
    w = workloads.single_triangle(resolution=(64, 64))
    w["camera"] = cam
    return cam, oracle.scene_from_workload(w)


@pytest.mark.parametrize("fisheye", [False, True])
def test_camera_screen_functions_match_reference(ref, olib, camera_scene, fisheye_scene, fisheye):
    """camera.h:86-214, 314-320: unproject, d_unproject, project_local,
    d_project_local and in_screen, pinhole and fisheye, bit for bit; and
    fisheye sample_primary (51-71), including screen points outside the
    disk (a ray without direction)."""
    cam, osc = fisheye_scene if fisheye else camera_scene
    args = _camera_args(cam)
    rng = np.random.default_rng(16 + fisheye)
    inside = 0
    for _ in range(400):
        scr = rng.uniform(-0.1, 1.1, size=2).astype(np.float32)
        loc = rng.normal(size=3).astype(np.float32)
        loc[2] = abs(loc[2]) + 0.1
        dxy = rng.normal(size=2).astype(np.float32)
        out, po = F(np.zeros(15, np.float32))
        keep = [F(scr), F(loc), F(dxy)]
        olib.or_probe_camera(osc.h, keep[0][1], keep[1][1], keep[2][1], po)
        e = np.array(ref.camera_functions(*args, fisheye, scr.tolist(), loc.tolist(), dxy.tolist()), np.float32)
        np.testing.assert_array_equal(out, e)
        inside += int(e[14])
        if fisheye:
            o6, p6 = F(np.zeros(6, np.float32))
            olib.or_probe_sample_primary(osc.h, keep[0][1], p6)
            er = np.array(ref.sample_primary_fisheye(*args, scr.tolist()), np.float32)
            np.testing.assert_array_equal(o6, er)
    assert 50 < inside < 400


def test_fisheye_maps_are_consistent(olib, fisheye_scene):
    """What the mathematics of the equi-angular map fixes: project_local
    inverts unproject inside the disk, the screen radius is linear in the
    angle to the view axis (pi/2 at the rim), and d_unproject / d_project_local
    match central differences of unproject / project_local."""
    cam, osc = fisheye_scene
    rng = np.random.default_rng(17)

    def probe(scr, loc=(0.0, 0.0, 1.0), dxy=(0.0, 0.0)):
        out, po = F(np.zeros(15, np.float64 if False else np.float32))
        keep = [F(scr), F(loc), F(dxy)]
        olib.or_probe_camera(osc.h, keep[0][1], keep[1][1], keep[2][1], po)
        return out.astype(np.float64)
    for _ in range(200):
        r, a = np.sqrt(rng.uniform(0.01, 0.9)), rng.uniform(0, 2 * np.pi)
        scr = np.array([0.5 + 0.5 * r * np.cos(a), 0.5 + 0.5 * r * np.sin(a)], np.float32)
        o = probe(scr)
        u = o[0:3]
        theta = np.arccos(u[2] / np.linalg.norm(u))
        assert abs(theta - r * np.pi / 2) < 2e-5
        back = probe(scr, loc=u)[9:11]
        np.testing.assert_allclose(back, scr, atol=2e-6)
        h = 1e-3
        for k, col in ((0, slice(3, 6)), (1, slice(6, 9))):
            ep = scr.astype(np.float64).copy()
            em = ep.copy()
            ep[k] += h
            em[k] -= h
            fd = (probe(ep.astype(np.float32))[0:3] - probe(em.astype(np.float32))[0:3]) / (2 * h)
            np.testing.assert_allclose(o[col], fd, rtol=2e-2, atol=2e-3)
        # d_project_local: the adjoint of project_local at a local direction
        loc = u * rng.uniform(0.5, 2.0)
        dxy = rng.normal(size=2)
        g = probe(scr, loc=loc, dxy=dxy)[11:14]
        fd = np.zeros(3)
        for k in range(3):
            lp, lm = loc.copy(), loc.copy()
            lp[k] += h
            lm[k] -= h
            fd[k] = dxy @ (probe(scr, loc=lp)[9:11] - probe(scr, loc=lm)[9:11]) / (2 * h)
        np.testing.assert_allclose(g, fd, rtol=2e-2, atol=2e-3)


def test_cone_box_matches_reference(ref, olib):
    """aabb.cpp:98-311 (the GTE cone-box test with its vertex/edge/face case
    table): identical decisions on random and apex-containing boxes."""
    rng = np.random.default_rng(15)
    hits = 0
    for k in range(4000):
        c = rng.normal(size=3) * 2
        h = rng.uniform(0.05, 1.5, size=3)
        lo, hi = (c - h).astype(np.float32), (c + h).astype(np.float32)
        if k % 20 == 0:
            org = c.astype(np.float32)   # apex inside the box
        else:
            org = (rng.normal(size=3) * 2).astype(np.float32)
        d = unit(rng).astype(np.float32)
        ca = np.float32(np.cos(rng.uniform(0.01, 1.4)))
        got = call(olib.or_probe_cone_box, lo, hi, org, d, extra=(ctypes.c_float(ca),))
        exp = ref.cone_box(lo.tolist(), hi.tolist(), org.tolist(), d.tolist(), float(ca))
        assert bool(got) == bool(exp), (lo, hi, org, d, ca)
        hits += bool(exp)
    assert 200 < hits < 3800


def test_clip_line_matches_reference(ref, olib):
    """line_clip.h: Cohen-Sutherland to [0,1]^2."""
    rng = np.random.default_rng(16)
    for _ in range(3000):
        a = rng.uniform(-1.0, 2.0, size=2).astype(np.float32)
        b = rng.uniform(-1.0, 2.0, size=2).astype(np.float32)
        ka, kb = F(a), F(b)
        out, po = F(np.zeros(4, np.float32))
        ok = olib.or_probe_clip_line(ka[1], kb[1], po)
        e = ref.clip_line(a.tolist(), b.tolist())
        assert bool(ok) == (len(e) > 0)
        if ok:
            np.testing.assert_array_equal(out, np.array(e, np.float32))


def test_jacobians_match_reference(ref, olib):
    """vector.h:690-733."""
    rng = np.random.default_rng(17)
    for k in range(500):
        x = rng.normal(size=3).astype(np.float32)
        d = unit(rng).astype(np.float32)
        p = rng.normal(size=3).astype(np.float32)
        n = unit(rng).astype(np.float32)
        if k % 50 == 0:
            n = np.cross(d, unit(rng)).astype(np.float32)   # dir . n ~ 0 branch
        keep = [F(v) for v in (x, d, p, n)]
        out, po = F(np.zeros(18, np.float32))
        olib.or_probe_jacobians(*[kk[1] for kk in keep], po)
        e = np.array(ref.jacobians(x.tolist(), d.tolist(), p.tolist(), n.tolist()), np.float32)
        np.testing.assert_allclose(out, e, rtol=2e-6, atol=1e-6 * max(1.0, np.abs(e).max()))


def test_distribution_matches_reference(ref, olib):
    """distribution.h:7-60: index and pmf, including zero entries, the
    all-zero fallback and u on CDF steps."""
    rng = np.random.default_rng(18)
    for k in range(300):
        n = int(rng.integers(1, 12))
        f = rng.uniform(size=n).astype(np.float32)
        f[rng.uniform(size=n) < 0.3] = 0.0
        if k % 25 == 0:
            f[:] = 0.0
        us = list(rng.uniform(size=5).astype(np.float32)) + [np.float32(0.0), np.float32(np.nextafter(1, 0))]
        for u in us:
            kf = F(f)
            pmf, pp = F(np.zeros(1, np.float32))
            i = olib.or_probe_sample_discrete(kf[1], n, ctypes.c_float(u), pp)
            ei, ep = ref.sample_discrete(f.tolist(), float(u))
            assert i == ei
            if np.isfinite(ep):
                assert pmf[0] == np.float32(ep)


def test_sample_triangle_matches_reference(ref, olib):
    """sample.h:6-21 (the light sampler's point on a triangle)."""
    rng = np.random.default_rng(19)
    for _ in range(500):
        v = rng.normal(size=(3, 3)).astype(np.float32)
        u = rng.uniform(size=2).astype(np.float32)
        keep = [F(x) for x in (v[0], v[1], v[2], u)]
        out, po = F(np.zeros(6, np.float32))
        olib.or_probe_sample_triangle(*[kk[1] for kk in keep], po)
        e = np.array(ref.sample_triangle(v[0].tolist(), v[1].tolist(), v[2].tolist(), u.tolist()), np.float32)
        np.testing.assert_allclose(out, e, rtol=1e-6, atol=1e-6)

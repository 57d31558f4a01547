"""Device refit (dr_scene_update_vertices): new vertex positions refitted on
the GPU must give what a scene built from those vertices gives.

* Ray queries and forward images do not depend on the trees' topology (the
  closest hit is the one with the least t, ties to the smaller triangle id),
  so a refitted scene must match a freshly built one bit for bit there.
* Scaling every shape of a scene by a power of two about the origin is exact
  in floating point and keeps every ordering and every comparison the
  builders make (median splits, SAH costs, the coplanar-edge filter), so the
  refitted trees have the fresh build's topology, and every adjoint term must
  match it up to the order of float atomics (and the last bit of acosf in a
  normal cone).  (Scaling one shape of several would reorder its edges
  against the others' in the median splits: a fresh build of that scene has
  other trees, and its Monte Carlo estimates differ from the refit's.)
* When the filter of degenerate / coplanar edges changes its verdict, the
  library rebuilds the structures on the host: the result is a fresh build.
"""
import numpy as np
import pytest

from delta_ray_amd import _abi, workloads
from delta_ray_amd import delta_ray as dr
from tests import helpers
from tests.helpers import gpu_render, gpu_scene

pytestmark = pytest.mark.gpu

TERMS = {
    "interior": _abi.DR_SKIP_SECONDARY_EDGES | _abi.DR_SKIP_PRIMARY_EDGES,
    "secondary": _abi.DR_SKIP_INTERIOR | _abi.DR_SKIP_PRIMARY_EDGES,
    "primary": _abi.DR_SKIP_INTERIOR | _abi.DR_SKIP_SECONDARY_EDGES,
}


def _with_vertices(w, verts, normals=None):
    """the workload with shape 0's vertices (and normals) replaced"""
    s0 = w["shapes"][0]
    n = s0.normals if normals is None else normals
    s = dr.Shape(verts, s0.indices, s0.uvs, n, s0.material, s0.light)
    return dict(w, shapes=[s] + list(w["shapes"][1:]))


def _scaled(w, k):
    """the workload with every shape scaled by k about the origin (lights
    follow their shapes)"""
    shapes, lights = [], []
    for s in w["shapes"]:
        ns = dr.Shape((k * np.asarray(s.vertices)).astype(np.float32), s.indices, s.uvs, s.normals,
                      s.material, None)
        if s.light is not None:
            ns.light = dr.Light(ns, np.asarray(s.light.intensity))
            lights.append(ns.light)
        shapes.append(ns)
    return dict(w, shapes=shapes, lights=lights)


def _update(scene, verts, normals=None):
    import torch
    v = torch.from_numpy(np.ascontiguousarray(verts, np.float32)).cuda()
    n = None if normals is None else torch.from_numpy(np.ascontiguousarray(normals, np.float32)).cuda()
    scene.update_vertices([(0, v, n)])


def _hits(scene, rays):
    import torch
    r = torch.from_numpy(rays).cuda()
    h = torch.zeros((len(rays), 4), dtype=torch.int32, device="cuda")
    scene.intersect(r, h)
    torch.cuda.synchronize()
    return h.cpu().numpy()


def _dimg(w):
    W, H = w["resolution"]
    return np.random.default_rng(4).normal(size=(H, W, 3)).astype(np.float32)


def _compare_scenes(wa, sa, wb, sb, rtol, floor):
    W, H = wa["resolution"]
    nb = wa["max_bounces"]
    ia, _, _ = gpu_render(sa, wa, 2, nb, 3)
    ib, _, _ = gpu_render(sb, wb, 2, nb, 3)
    np.testing.assert_array_equal(ia, ib)
    rays = helpers.random_rays(wb, 20000, np.random.default_rng(2))
    np.testing.assert_array_equal(_hits(sa, rays), _hits(sb, rays))
    d = _dimg(wa)
    for term, f in TERMS.items():
        print("term", term, flush=True)
        _, ga, _ = gpu_render(sa, wa, 2, nb, 5, flags=_abi.DR_FLAGS_PARITY | f, d_image=d)
        _, gb, _ = gpu_render(sb, wb, 2, nb, 5, flags=_abi.DR_FLAGS_PARITY | f, d_image=d)
        assert np.abs(gb).sum() > 0
        helpers.assert_groups(wb, ga, gb, rtol, floor)


def test_refit_back_to_built_vertices_reproduces_scene(gpu):
    print("case: refit identity", flush=True)
    w = helpers.small_workload("heightfield", res=48, grid=24)
    v0 = np.asarray(w["shapes"][0].vertices).copy()
    n0 = np.asarray(w["shapes"][0].normals).copy()
    ref = gpu_scene(w)
    sc = gpu_scene(w)
    # a move that keeps the coplanar-edge filter's verdicts (a generic one
    # flips the verdict of nearly coplanar faces of the heightfield, and the
    # library rebuilds), and its undo
    _update(sc, 4.0 * v0, n0)
    _update(sc, v0, n0)
    info = sc.info()
    assert info["refits"] == 2 and info["rebuilds"] == 0, info
    _compare_scenes(w, sc, w, ref, 1e-5, 1e-6)


def test_refit_matches_fresh_build(gpu):
    print("case: refit scaled", flush=True)
    w = helpers.small_workload("heightfield", res=48, grid=24)
    w2 = _scaled(w, 2.0)
    sc = gpu_scene(w)
    import torch
    sc.update_vertices([(i, torch.from_numpy(np.asarray(s.vertices)).cuda(), None)
                        for i, s in enumerate(w2["shapes"])])
    info = sc.info()
    assert info["refits"] == 1 and info["rebuilds"] == 0, info
    _compare_scenes(w2, sc, w2, gpu_scene(w2), 1e-5, 1e-6)


def test_refit_rebuilds_when_the_edge_set_changes(gpu):
    print("case: refit rebuild", flush=True)
    w = helpers.small_workload("heightfield", res=48, grid=24)
    v = np.asarray(w["shapes"][0].vertices)
    ind = np.asarray(w["shapes"][0].indices)
    flat = v.copy()
    flat[:, 1] = 0.0
    wf = _with_vertices(w, flat, workloads.vertex_normals(flat, ind))
    sc = gpu_scene(wf)
    e_flat = sc.info()["num_edges"]
    _update(sc, v, np.asarray(w["shapes"][0].normals))
    info = sc.info()
    assert info["rebuilds"] == 1 and info["num_edges"] > e_flat, (info, e_flat)
    _compare_scenes(w, sc, w, gpu_scene(w), 1e-5, 1e-6)


def test_refit_moves_a_light(gpu):
    """An emissive shape moved: the light distributions follow (area, CDF,
    bounding sphere) and the render equals a fresh build's."""
    print("case: refit light", flush=True)
    w = helpers.small_workload("shadow", res=32)
    li = len(w["shapes"]) - 1
    assert w["shapes"][li].light is not None
    lv = np.asarray(w["shapes"][li].vertices)
    lv2 = (lv + np.array([0.25, -0.5, 0.125], np.float32)).astype(np.float32)
    shapes = list(w["shapes"])
    s = shapes[li]
    ns = dr.Shape(lv2, s.indices, s.uvs, s.normals, s.material, None)
    lt = dr.Light(ns, np.asarray(w["lights"][0].intensity))
    ns.light = lt
    shapes[li] = ns
    w2 = dict(w, shapes=shapes, lights=[lt])
    sc = gpu_scene(w)
    import torch
    sc.update_vertices([(li, torch.from_numpy(lv2).cuda(), None)])
    W, H = w["resolution"]
    a, _, _ = gpu_render(sc, w, 4, 1, 3)
    b, _, _ = gpu_render(gpu_scene(w2), w2, 4, 1, 3)
    np.testing.assert_array_equal(a, b)


def _mean_secondary(scene, w, seeds, d):
    """mean and standard error (per entry) of the secondary-edge term over seeds"""
    W, H = w["resolution"]
    gs = []
    for sd in seeds:
        _, g, _ = gpu_render(scene, w, 16, w["max_bounces"], sd, flags=_abi.DR_FLAGS_PARITY | TERMS["secondary"],
                             d_image=d)
        gs.append(np.asarray(g, np.float64))
    gs = np.array(gs)
    return gs.mean(0), gs.std(0, ddof=1) / np.sqrt(len(gs))


def _same_in_expectation(sa, sb, w, d, name):
    ma, ea = _mean_secondary(sa, w, range(100, 132), d)
    mb, eb = _mean_secondary(sb, w, range(200, 232), d)
    se = np.sqrt(ea ** 2 + eb ** 2)
    z = np.linalg.norm(ma - mb) / np.linalg.norm(se)
    print("%s secondary: |mean difference| / |se| = %.3f, |mean| %.4g" % (name, z, np.linalg.norm(mb)), flush=True)
    assert np.linalg.norm(mb) > 0 and z < 3.0


def test_all_candidate_trees_refit_without_rebuild(gpu):
    """Trees over every mesh edge (DR_EDGES_ALL_CANDIDATES), the coplanar
    filter's verdicts as device flags: on the built vertices and after a move
    that flips verdicts (which rebuilds a filtered scene), the forward image,
    ray hits, interior and primary-edge terms equal a fresh filtered build's,
    and the secondary-edge term agrees with it in expectation (other trees,
    so other samples)."""
    print("case: all-candidate trees", flush=True)
    w = helpers.small_workload("heightfield", res=48, grid=24)
    v0 = np.asarray(w["shapes"][0].vertices)
    v1 = v0.copy()
    v1[:, 1] += 1e-4 * np.random.default_rng(3).standard_normal(len(v0)).astype(np.float32)
    w1 = _with_vertices(w, v1)
    filtered = gpu_scene(w)
    _update(filtered, v1)
    assert filtered.info()["rebuilds"] == 1
    allc = gpu_scene(w)
    n_kept = allc.info()["num_edges"]
    allc.set_edge_policy(_abi.DR_EDGES_ALL_CANDIDATES)
    info = allc.info()
    assert info["num_edges"] == n_kept + info["edges_switched_off"], (info, n_kept)
    d = _dimg(w)
    ref0 = gpu_scene(w)
    _same_in_expectation(allc, ref0, w, d, "built vertices")
    _update(allc, v1)
    info = allc.info()
    assert info["refits"] == 1 and info["rebuilds"] == 0, info
    fresh = gpu_scene(w1)
    assert info["num_edges"] == fresh.info()["num_edges"] + info["edges_switched_off"], (info, fresh.info())
    ia, _, _ = gpu_render(allc, w1, 2, 2, 3)
    ib, _, _ = gpu_render(fresh, w1, 2, 2, 3)
    np.testing.assert_array_equal(ia, ib)
    rays = helpers.random_rays(w1, 20000, np.random.default_rng(2))
    np.testing.assert_array_equal(_hits(allc, rays), _hits(fresh, rays))
    for term in ("interior", "primary"):
        _, ga, _ = gpu_render(allc, w1, 2, 2, 5, flags=_abi.DR_FLAGS_PARITY | TERMS[term], d_image=d)
        _, gb, _ = gpu_render(fresh, w1, 2, 2, 5, flags=_abi.DR_FLAGS_PARITY | TERMS[term], d_image=d)
        print("term", term, flush=True)
        helpers.assert_groups(w1, ga, gb, 1e-5, 1e-6)
    _same_in_expectation(allc, fresh, w1, d, "moved vertices")

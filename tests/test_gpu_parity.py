"""GPU parity: the HIP library against the CPU oracle on identical inputs.

Both sides draw the same counter-based random numbers, so a pixel sample
traces the same path on the CPU and on the GPU: images and gradients are
compared per pixel / per parameter, not only statistically.  Remaining
differences come from transcendental functions (ocml vs glibc ulps) that can
flip a rare near-tangent ray; tolerances below allow a small fraction of such
samples and say so.

Intersection is bit-exact: hit shape/triangle ids and the float bits of the
barycentrics u, v (reference intersect.h:71-90) must match the oracle's
brute force on every ray.
"""
import numpy as np
import pytest

from delta_ray_amd import _abi
from tests import helpers
from tests.helpers import gpu_render, gpu_scene, rel_err, small_workload

pytestmark = pytest.mark.gpu

# fisheye: the equi-angular camera; clipped: a triangle with a vertex behind
# the camera (primary edges clipped at the near plane), the reference's
# test_single_triangle_fisheye.py and test_single_triangle_clipped.py scenes
SCENES = ["single_triangle", "shadow", "heightfield", "glossy_plane", "fisheye", "clipped"]
NO_SECONDARY = ("single_triangle", "fisheye", "clipped")  # direct light, nothing to occlude it


def _print_case(*a):
    print("case:", *a, flush=True)


@pytest.mark.parametrize("name", SCENES)
def test_intersect_bitexact(gpu, oracle, name):
    import torch
    _print_case("intersect", name)
    w = small_workload(name)
    rng = np.random.default_rng(11)
    rays = helpers.random_rays(w, 20000, rng)
    osc = oracle.scene_from_workload(w)
    osc.use_bvh(False)
    ref = osc.intersect(rays)
    ref_any = osc.intersect(np.concatenate([rays[:, :7], np.full((len(rays), 1), 5.0, np.float32)], 1), any_hit=True)
    sc = gpu_scene(w)
    r = torch.from_numpy(rays).cuda()
    h = torch.zeros((len(rays), 4), dtype=torch.int32, device="cuda")
    sc.intersect(r, h)
    torch.cuda.synchronize()
    got = h.cpu().numpy()
    assert (got[:, 0] >= 0).sum() > len(rays) // 20, "too few hits to be a meaningful test"
    np.testing.assert_array_equal(got[:, 0], ref[:, 0])
    hitm = ref[:, 0] >= 0
    np.testing.assert_array_equal(got[hitm, 1:4], ref[hitm, 1:4])
    r2 = torch.from_numpy(np.concatenate([rays[:, :7], np.full((len(rays), 1), 5.0, np.float32)], 1)).cuda()
    sc.intersect(r2, h, any_hit=True)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(h.cpu().numpy()[:, 0], ref_any[:, 0])


@pytest.mark.parametrize("name", SCENES)
def test_forward_matches_oracle(gpu, oracle, name):
    _print_case("forward", name)
    w = small_workload(name)
    W, H = w["resolution"]
    nb = w["max_bounces"]
    img, _, _ = gpu_render(gpu_scene(w), w, 2, nb, 7)
    ref, _, _ = oracle.scene_from_workload(w).render(W, H, 2, nb, 7)
    assert ref.sum() > 0
    err = np.abs(img - ref)
    tol = 1e-4 * np.maximum(np.abs(ref), 1e-2)
    frac_bad = float(np.mean(np.any(err > tol, axis=2)))
    # a divergent path (different ulp in a sampled direction) may change a
    # pixel; allow 0.5% of pixels
    assert frac_bad <= 0.005, frac_bad
    assert rel_err(img, ref) < 2e-2


TERMS = {
    "interior": _abi.DR_SKIP_SECONDARY_EDGES | _abi.DR_SKIP_PRIMARY_EDGES,
    "secondary": _abi.DR_SKIP_INTERIOR | _abi.DR_SKIP_PRIMARY_EDGES,
    "primary": _abi.DR_SKIP_INTERIOR | _abi.DR_SKIP_SECONDARY_EDGES,
    # the secondary-edge term with every quirk bit clear (the exact
    # estimator: component-only BSDF, side-hit Jacobian from the shading
    # point, consistent cone-pass pmfs)
    "secondary_exact": _abi.DR_SKIP_INTERIOR | _abi.DR_SKIP_PRIMARY_EDGES,
}
BASE = {"secondary_exact": 0}


# Per parameter group (camera, each material texture, each shape's vertices
# and normals, each light): ||g - g_oracle|| <= RTOL ||g_oracle|| + FLOOR
# ||g_oracle (all)||.  Both sides trace the same paths with the same random
# numbers; what is left is float rounding of the sums (atomics in a different
# order) and the rare sample whose path flips on the last bit of a
# transcendental (ocml vs glibc), which the floor absorbs.
RTOL = {"interior": 1e-3, "secondary": 1e-3, "primary": 1e-3, "secondary_exact": 1e-3}
FLOOR = {"interior": 1e-4, "secondary": 2e-3, "primary": 1e-4, "secondary_exact": 2e-3}


def _adjoint_case(oracle, w, term, spp, seed):
    W, H = w["resolution"]
    nb = w["max_bounces"]
    rng = np.random.default_rng(3)
    d_img = rng.normal(size=(H, W, 3)).astype(np.float32)
    flags = BASE.get(term, _abi.DR_FLAGS_PARITY) | TERMS[term]
    img, g, _ = gpu_render(gpu_scene(w), w, spp, nb, seed, flags=flags, d_image=d_img)
    ref_img, gref, _ = oracle.scene_from_workload(w).render(W, H, spp, nb, seed, flags=flags, d_image=d_img)
    return g, gref


# The fisheye camera maps every screen point through atan2 / sin / cos
# (ocml on the GPU, glibc in the oracle), so every camera ray -- not a rare
# one -- differs in its last bits, and the secondary-edge estimator (heavy
# tailed, with visibility decisions at light borders) turns that into a few
# per cent per render (2-4 % measured at 48 x 48, 2 spp).  Its per-render bar
# is therefore loose; test_fisheye_secondary_agrees_in_expectation checks that
# the difference averages out (no bias), and the pinhole scenes keep the tight
# per-render bar.
LOOSE = {("fisheye", "secondary"): (0.1, 2e-2), ("fisheye", "secondary_exact"): (0.1, 2e-2)}


@pytest.mark.parametrize("name", SCENES)
@pytest.mark.parametrize("term", list(TERMS))
def test_adjoint_matches_oracle(gpu, oracle, name, term):
    _print_case("adjoint", name, term)
    w = small_workload(name)
    g, gref = _adjoint_case(oracle, w, term, 2, 5)
    assert np.abs(gref).sum() > 0 or term.startswith("secondary") and name in NO_SECONDARY, \
        "oracle gradient is zero"
    print("rel err (all)", rel_err(g, gref), "norm", float(np.linalg.norm(gref)))
    rtol, floor = LOOSE.get((name, term), (RTOL[term], FLOOR[term]))
    helpers.assert_groups(w, g, gref, rtol, floor)


def test_fisheye_secondary_agrees_in_expectation(gpu, oracle):
    """The fisheye secondary-edge term over 24 seeds: the GPU's and the
    oracle's means agree within their standard errors (|difference| / |se|
    below 1; measured 0.03), while single renders differ by a few per cent."""
    _print_case("fisheye secondary expectation")
    w = small_workload("fisheye")
    W, H = w["resolution"]
    d = np.random.default_rng(3).normal(size=(H, W, 3)).astype(np.float32)
    flags = _abi.DR_FLAGS_PARITY | TERMS["secondary"]
    sc = gpu_scene(w)
    osc = oracle.scene_from_workload(w)
    G, O = [], []
    for seed in range(100, 124):
        G.append(gpu_render(sc, w, 2, 1, seed, flags=flags, d_image=d)[1].astype(np.float64))
        O.append(osc.render(W, H, 2, 1, seed, flags=flags, d_image=d)[1].astype(np.float64))
    G, O = np.array(G), np.array(O)
    se = np.sqrt(G.var(0, ddof=1) / len(G) + O.var(0, ddof=1) / len(O))
    z = np.linalg.norm(G.mean(0) - O.mean(0)) / np.linalg.norm(se)
    print("|mean difference| / |se| = %.3f" % z, flush=True)
    assert np.linalg.norm(O.mean(0)) > 0 and z < 1.0


@pytest.mark.parametrize("term", list(TERMS))
def test_adjoint_matches_oracle_full_heightfield_mesh(gpu, oracle, term):
    """The benchmark's 128 x 128 heightfield mesh (32 258 triangles, normal
    edge tree of ~96 k nodes, 2 bounces) at 64 x 64 pixels, 1 sample."""
    _print_case("adjoint heightfield-128", term)
    w = small_workload("heightfield", res=64, grid=128)
    g, gref = _adjoint_case(oracle, w, term, 1, 9)
    assert np.abs(gref).sum() > 0
    print("rel err (all)", rel_err(g, gref), "norm", float(np.linalg.norm(gref)))
    helpers.assert_groups(w, g, gref, RTOL[term], FLOOR[term])


def _geometric_mesh(n=150):
    """n triangles along x at 0.8^k (k = 0 .. n-1), each of size 0.8^k / 8:
    most of them crowd one end of the binned-SAH build's bins, which then
    splits off a few triangles per level, far deeper than the traversal
    stack without the builder's depth bound"""
    from delta_ray_amd import workloads
    from delta_ray_amd import delta_ray as dr
    v, ind = [], []
    for k in range(n):
        c, h = 0.8 ** k, 0.8 ** k / 8
        v += [[c - h, -h, 0.0], [c + h, -h, 0.0], [c, h, 0.0]]
        ind.append([3 * k, 3 * k + 1, 3 * k + 2])
    v, ind = np.array(v, np.float32), np.array(ind, np.int32)
    grey = workloads.diffuse_material([0.5, 0.5, 0.5])
    mesh = dr.Shape(v, ind, None, None, grey, None)
    lv = np.array([[-1, -1, 5], [1, -1, 5], [-1, 1, 5]], np.float32)
    ls = dr.Shape(lv, np.array([[0, 1, 2]], np.int32), None, None, grey, None)
    light = dr.Light(ls, np.full(3, 10.0, np.float32))
    ls.light = light
    cam = workloads.make_camera([0.5, 0, 3], [0.5, 0, 0], [0, 1, 0], 45.0, (16, 16))
    return dict(name="geometric", camera=cam, shapes=[mesh, ls], materials=[grey], lights=[light],
                resolution=(16, 16), num_samples=1, max_bounces=1, params=[("shape", 0, "vertices")]), v


def test_deep_bvh_mesh_intersects_exactly(gpu, oracle):
    """Geometrically spaced triangles (a degenerate input for a SAH build):
    the builder's depth bound keeps the triangle BVH within the traversal
    stack, so every ray down onto each triangle finds it, bit for bit as the
    oracle's brute force, and no traversal stack overflows."""
    import torch
    _print_case("intersect", "geometric mesh")
    w, v = _geometric_mesh()
    n = len(v) // 3
    cen = v.reshape(n, 3, 3).mean(1)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:2] = cen[:, 0:2]
    rays[:, 2] = 1.0
    rays[:, 3] = 0.0
    rays[:, 6] = -1.0
    rays[:, 7] = 10.0
    osc = oracle.scene_from_workload(w)
    osc.use_bvh(False)
    ref = osc.intersect(rays)
    sc = gpu_scene(w)
    sc.stats(True)
    r = torch.from_numpy(rays).cuda()
    h = torch.zeros((n, 4), dtype=torch.int32, device="cuda")
    sc.intersect(r, h)
    torch.cuda.synchronize()
    got = h.cpu().numpy()
    assert (ref[:, 0] == 0).sum() >= n - 2, "rays should hit the mesh"
    np.testing.assert_array_equal(got, ref)
    st = sc.stats(False)
    assert st["stack_overflows"] == 0
    info = sc.info()
    print("bvh depth", info["bvh_depth"])
    assert info["bvh_depth"] <= 64

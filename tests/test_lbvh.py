"""The triangle BVH built on the device (dr_scene_build_bvh_device, a linear
BVH after Karras, HPG 2012) in place of the host's binned-SAH tree.  Ray hits
do not depend on the tree (closest t, ties to the smaller triangle id), so
they must equal the host-built tree's and the oracle's brute force bit for
bit; forward images equal; refits of the device-built tree track moved
vertices as a fresh scene does."""
import numpy as np
import pytest

from delta_ray_amd import _abi
from tests import helpers
from tests.test_gpu_parity import _geometric_mesh

pytestmark = pytest.mark.gpu


def _hits(torch, sc, rays):
    r = torch.from_numpy(rays).cuda()
    h = torch.zeros((len(rays), 4), dtype=torch.int32, device="cuda")
    sc.intersect(r, h)
    torch.cuda.synchronize()
    return h.cpu().numpy()


def _scaled(w, k):
    from delta_ray_amd import delta_ray as dr
    shapes, lights = [], []
    for s in w["shapes"]:
        shapes.append(dr.Shape((np.asarray(s.vertices) * k).astype(np.float32), s.indices, s.uvs, s.normals,
                               s.material, None))
    for l in w["lights"]:
        i = w["shapes"].index(l.shape)
        nl = dr.Light(shapes[i], l.intensity)
        shapes[i].light = nl
        lights.append(nl)
    return dict(w, shapes=shapes, lights=lights)


@pytest.mark.parametrize("name", ["shadow", "heightfield", "geometric"])
def test_device_bvh_gives_the_same_hits(gpu, oracle, name):
    import torch
    print("case: device bvh", name, flush=True)
    if name == "geometric":
        w, _ = _geometric_mesh()
        rng = np.random.default_rng(5)
        rays = helpers.random_rays(w, 20000, rng)
    else:
        w = helpers.small_workload(name)
        rays = helpers.random_rays(w, 20000, np.random.default_rng(7))
    osc = oracle.scene_from_workload(w)
    osc.use_bvh(False)
    ref = osc.intersect(rays)
    sc = helpers.gpu_scene(w)
    before = _hits(torch, sc, rays)
    W, H = w["resolution"]
    img0, _, _ = helpers.gpu_render(sc, w, 2, w["max_bounces"], 3)
    ms = sc.build_bvh_device()
    info = sc.info()
    print("device build %.3f ms, %d nodes, depth %d" % (ms, info["bvh_nodes"], info["bvh_depth"]))
    assert info["bvh_nodes"] == info["num_triangles"] - 1 and info["bvh_depth"] <= 64
    after = _hits(torch, sc, rays)
    np.testing.assert_array_equal(after, before)
    assert (ref[:, 0] >= 0).sum() > len(rays) // 20
    np.testing.assert_array_equal(after[:, 0], ref[:, 0])
    m = ref[:, 0] >= 0
    np.testing.assert_array_equal(after[m, 1:4], ref[m, 1:4])
    img1, _, _ = helpers.gpu_render(sc, w, 2, w["max_bounces"], 3)
    np.testing.assert_allclose(img1, img0, rtol=1e-6, atol=1e-7)
    # a refit of the device-built tree: every shape scaled by 4 (exact in
    # floating point) -- the hits are a fresh scene's
    w4 = _scaled(w, 4.0)
    sc.update_vertices([(i, torch.from_numpy(np.asarray(s.vertices)).cuda(), None)
                        for i, s in enumerate(w4["shapes"])])
    rays4 = rays.copy()
    rays4[:, 0:3] *= 4.0
    rays4[:, 7] = np.where(np.isfinite(rays4[:, 7]), rays4[:, 7] * 4.0, rays4[:, 7])
    o4 = oracle.scene_from_workload(w4)
    o4.use_bvh(False)
    np.testing.assert_array_equal(_hits(torch, sc, rays4), o4.intersect(rays4))

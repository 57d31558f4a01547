"""Batched renders (dr_render_batch): many scenes, views, resolutions and
sample counts in one launch sequence must give what separate
dr_render / dr_render_adjoint calls give (the random streams depend only on
the job's seed and sample index, so the results agree up to the order of float
atomics), also when captured and replayed as a hipGraph."""
import numpy as np
import pytest

from delta_ray_amd import _abi, workloads
from delta_ray_amd import delta_ray as dr
from tests import helpers

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12))


def _buffers(torch, scene, W, H, adjoint, seed, dev):
    rng = np.random.default_rng(seed)
    out = {"image": torch.zeros((H, W, 3), device=dev)}
    if adjoint:
        out["d_image"] = torch.from_numpy(rng.normal(size=(H, W, 3)).astype(np.float32)).to(dev)
        out["grads"] = torch.zeros((scene.grad_size,), device=dev)
    return out


def _single(torch, scene, o, bufs, dev):
    img = torch.zeros_like(bufs["image"])
    if "d_image" in bufs:
        g = torch.zeros_like(bufs["grads"])
        scene.render_adjoint(o, bufs["d_image"], img, g)
        torch.cuda.synchronize()
        return img.cpu().numpy(), g.cpu().numpy()
    scene.render(o, img)
    torch.cuda.synchronize()
    return img.cpu().numpy(), None


def _job_list(torch, dev):
    specs = [("single_triangle", 40, 3, 1, True), ("shadow", 24, 5, 1, True), ("glossy_plane", 32, 2, 2, True),
             ("heightfield", 20, 3, 2, True), ("single_triangle", 16, 7, 1, False), ("shadow", 12, 1, 1, False)]
    jobs = []
    for k, (name, res, spp, nb, adj) in enumerate(specs):
        w = helpers.small_workload(name, res=res)
        scene = helpers.gpu_scene(w)
        o = dr.DeviceScene.options(res, res, spp, nb, 100 + k)
        bufs = _buffers(torch, scene, res, res, adj, k, dev)
        jobs.append(dict(scene=scene, options=o, **bufs))
    return jobs


def test_batch_matches_single_renders(gpu):
    import torch
    jobs = _job_list(torch, gpu)
    batch = dr.Batch()
    batch.render(jobs)
    torch.cuda.synchronize()
    for j in jobs:
        ref_img, ref_g = _single(torch, j["scene"], j["options"], j, gpu)
        assert _rel(j["image"].cpu().numpy(), ref_img) < 1e-5
        if ref_g is not None:
            assert _rel(j["grads"].cpu().numpy(), ref_g) < 1e-4


@pytest.mark.parametrize("refill", [0, 1, 64])
def test_batch_small_passes_match(gpu, refill):
    """Jobs split over many passes (tiny pass budget) give the same result,
    whatever the edge selection: breadth first (refill 0) or depth first
    with 1 or 64 idle lanes gathered before new shading points start (the
    schedule changes, the selected edges and the estimate do not)."""
    import torch
    jobs = _job_list(torch, gpu)
    batch = dr.Batch()
    batch.set_tuning(pass_samples=1 << 12, min_threads=1 << 10, select_refill=refill)
    batch.render(jobs)
    torch.cuda.synchronize()
    for j in jobs:
        ref_img, ref_g = _single(torch, j["scene"], j["options"], j, gpu)
        assert _rel(j["image"].cpu().numpy(), ref_img) < 1e-5
        if ref_g is not None:
            assert _rel(j["grads"].cpu().numpy(), ref_g) < 1e-4


def test_views_and_shared_gradients(gpu):
    """A view equals a scene built with that camera; two views of one scene
    sharing a gradient buffer add their gradients."""
    import torch
    w = helpers.small_workload("single_triangle", res=32)
    scene = helpers.gpu_scene(w)
    cam2 = workloads.make_camera([0.5, 0.3, 4.5], [0, 0, 0], [0, 1, 0], 50.0, (32, 32))
    view = dr.View(scene, cam2)
    w2 = dict(w)
    w2["camera"] = cam2
    scene2 = helpers.gpu_scene(w2)
    o = dr.DeviceScene.options(32, 32, 4, 1, 9)
    a = _buffers(torch, scene, 32, 32, True, 1, gpu)
    b = _buffers(torch, scene, 32, 32, True, 2, gpu)
    shared = torch.zeros((scene.grad_size,), device=gpu)
    jobs = [dict(scene=scene, options=o, image=a["image"], d_image=a["d_image"], grads=shared),
            dict(scene=scene, view=view, options=o, image=b["image"], d_image=b["d_image"], grads=shared)]
    dr.Batch().render(jobs)
    torch.cuda.synchronize()
    img_a, g_a = _single(torch, scene, o, a, gpu)
    img_b, g_b = _single(torch, scene2, o, b, gpu)
    assert _rel(jobs[0]["image"].cpu().numpy(), img_a) < 1e-5
    assert _rel(jobs[1]["image"].cpu().numpy(), img_b) < 1e-5
    assert _rel(shared.cpu().numpy(), g_a + g_b) < 1e-4


def test_graph_replays_match_eager(gpu):
    """Capture once, replay with new seeds; eager work and synchronisation
    between replays; every replay equals an eager batch with that seed."""
    import torch
    jobs = _job_list(torch, gpu)
    graph_batch, eager_batch = dr.Batch(), dr.Batch()
    for rep in range(3):
        for k, j in enumerate(jobs):
            j["options"].seed = 1000 * rep + k
        graph_batch.render(jobs, graph=True)
        torch.cuda.synchronize()
        got = [(j["image"].cpu().numpy().copy(), None if "grads" not in j else j["grads"].cpu().numpy().copy())
               for j in jobs]
        _ = torch.ones(1 << 20, device=gpu).sum().item()   # unrelated eager work between replays
        eager_batch.render(jobs)
        torch.cuda.synchronize()
        for (img, g), j in zip(got, jobs):
            assert _rel(img, j["image"].cpu().numpy()) < 1e-5
            if g is not None:
                assert _rel(g, j["grads"].cpu().numpy()) < 1e-4


def test_autotune_picks_a_candidate(gpu):
    import torch
    jobs = _job_list(torch, gpu)[:3]
    batch = dr.Batch()
    t = batch.autotune(jobs)
    # 4 x 3 pass sizes, then 3 more refill thresholds and the breadth-first
    # selection at the best of them
    assert t["candidates"] == 16 and t["best_ms"] > 0
    assert t["pass_samples"] in (1 << 20, 1 << 21, 1 << 22, 1 << 23)
    assert t["select_refill"] in (0, 8, 16, 32, 64)
    batch.render(jobs)
    torch.cuda.synchronize()
    for j in jobs:
        ref_img, ref_g = _single(torch, j["scene"], j["options"], j, gpu)
        assert _rel(j["image"].cpu().numpy(), ref_img) < 1e-5
        assert _rel(j["grads"].cpu().numpy(), ref_g) < 1e-4


def test_batched_workload_matches_oracle(gpu, oracle):
    """Four scenes of the uneven batched workload (the largest icosphere, the
    largest heightfield and two small ones, 1 k - 144 k triangles, 2 bounces),
    two 64 x 64 views each, rendered in one captured batch: every job's image
    and every parameter group of its gradient against the oracle rendering
    that scene from that camera."""
    import torch
    scenes = workloads.batched()
    pick = [0, 62, 2, 6]
    jobs, refs = [], []
    for k in pick:
        s = scenes[k]
        print("case: batched scene", k, s["num_triangles"], "triangles", flush=True)
        w = dict(s, resolution=(64, 64))
        ds = helpers.gpu_scene(w)
        for v, (cam, _, _) in enumerate(s["views"][:2]):
            cam64 = workloads.make_camera(
                np.asarray(cam.cam_to_world)[:3, 3], [0, 0.8, 0], [0, 1, 0], 45.0, (64, 64))
            view = dr.View(ds, cam64)
            seed = 1000 * k + v
            o = dr.DeviceScene.options(64, 64, 2, 2, seed)
            bufs = _buffers(torch, ds, 64, 64, True, seed, gpu)
            jobs.append(dict(scene=ds, view=view, options=o, **bufs))
            wv = dict(w, camera=cam64)
            refs.append((wv, seed, bufs["d_image"].cpu().numpy()))
    batch = dr.Batch()
    batch.render(jobs, graph=True)
    torch.cuda.synchronize()
    for j, (wv, seed, d_img) in zip(jobs, refs):
        ref_img, gref, _ = oracle.scene_from_workload(wv).render(64, 64, 2, 2, seed, d_image=d_img)
        img = j["image"].cpu().numpy()
        g = j["grads"].cpu().numpy()
        assert ref_img.sum() > 0 and np.abs(gref).sum() > 0
        print("image rel err", _rel(img, ref_img), flush=True)
        assert _rel(img, ref_img) < 2e-3
        helpers.assert_groups(wv, g, gref, 1e-3, 2e-3)

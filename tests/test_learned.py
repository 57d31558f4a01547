"""The learned secondary-edge proposal (DR_LEARNED_EDGES).  Nothing in the
reference pins it; it is pinned by unbiasedness: whatever the per-edge
weights, the mean secondary-edge gradient equals the plain estimator's within
a few standard errors, because the two tree passes of the proposal a point
runs (0.8 learned, 0.2 plain) are weighted with the power heuristic and the
proposals with the balance heuristic.  With all weights 1 both proposals are
the plain one, and the mixture gives the plain estimate sample by sample."""
import numpy as np
import pytest

from delta_ray_amd import _abi
from delta_ray_amd import delta_ray as dr
from delta_ray_amd.learned import EdgeProposal
from tests import helpers

pytestmark = pytest.mark.gpu

SECONDARY = _abi.DR_SKIP_INTERIOR | _abi.DR_SKIP_PRIMARY_EDGES


def _grads(torch, scene, w, spp, seeds, flags, d, base=_abi.DR_FLAGS_PARITY):
    W, H = w["resolution"]
    out = []
    img = torch.zeros((H, W, 3), device="cuda")
    g = torch.zeros((scene.grad_size,), device="cuda")
    for s in seeds:
        o = dr.DeviceScene.options(W, H, spp, w["max_bounces"], s, base | flags)
        scene.render_adjoint(o, d, img, g)
        torch.cuda.synchronize()
        out.append(g.cpu().numpy().copy())
    return np.array(out)


@pytest.mark.parametrize("name", ["shadow", "heightfield"])
def test_learned_mixture_is_unbiased(gpu, name):
    import torch
    w = helpers.small_workload(name, res=32)
    scene = helpers.gpu_scene(w)
    W, H = w["resolution"]
    d = torch.from_numpy(np.random.default_rng(0).normal(size=(H, W, 3)).astype(np.float32)).to(gpu)
    prop = EdgeProposal(scene, seed=1)
    # arbitrary, far-from-uniform weights: the estimate must stay unbiased
    wts = np.exp(np.random.default_rng(1).normal(0.0, 1.5, size=prop.n)).astype(np.float32)
    prop.set_weights(wts)
    plain = _grads(torch, scene, w, 16, range(100, 164), SECONDARY, d)
    learned = _grads(torch, scene, w, 16, range(200, 264), SECONDARY | _abi.DR_LEARNED_EDGES, d)
    mp, ml = plain.mean(0), learned.mean(0)
    se = np.sqrt(plain.var(0) / len(plain) + learned.var(0) / len(learned))
    active = se > 0
    z = np.abs(mp - ml)[active] / se[active]
    # per parameter |z| is ~half-normal: few beyond 4, and the aggregate
    # distance is what independent noise gives
    assert np.mean(z > 4.0) < 0.01, np.sort(z)[-10:]
    assert np.linalg.norm(mp - ml) < 4.0 * np.linalg.norm(se) + 1e-6


def test_features_and_statistics(gpu):
    import torch
    w = helpers.small_workload("heightfield", res=32)
    scene = helpers.gpu_scene(w)
    prop = EdgeProposal(scene, seed=2)
    f = prop.features[:prop.n].cpu().numpy()
    assert f.shape == (prop.n, _abi.DR_MLP_FEATURES) and np.all(np.isfinite(f))
    np.testing.assert_allclose(np.linalg.norm(f[:, 5:8], axis=1), 1.0, rtol=1e-4)  # face normal n0
    assert np.all(f[:, 23] == 1.0) and np.all(f[:, 13] > 0)
    W, H = w["resolution"]
    d = torch.from_numpy(np.random.default_rng(0).normal(size=(H, W, 3)).astype(np.float32)).to(gpu)
    prop.reset_stats()
    _grads(torch, scene, w, 4, [5], SECONDARY | _abi.DR_EDGE_STATS, d)
    prop.collect()
    acc = prop.acc[:prop.n].cpu().numpy()
    assert acc[:, 1].sum() > 0 and np.all(acc >= 0) and np.all(np.isfinite(acc))
    # contributions only where there are samples
    assert np.all(acc[acc[:, 1] == 0, 0] == 0)


def test_train_and_apply_end_to_end(gpu):
    """Statistics from plain renders -> training -> weights -> learned
    renders that agree with plain ones in the mean."""
    import torch
    w = helpers.small_workload("shadow", res=32)
    scene = helpers.gpu_scene(w)
    W, H = w["resolution"]
    d = torch.from_numpy(np.random.default_rng(3).normal(size=(H, W, 3)).astype(np.float32)).to(gpu)
    prop = EdgeProposal(scene, seed=4)
    prop.reset_stats()
    _grads(torch, scene, w, 16, range(8), SECONDARY | _abi.DR_EDGE_STATS, d)
    prop.collect()
    loss = prop.train(steps=200, batch=256)
    assert np.isfinite(loss)
    wts = prop.apply()
    assert wts.shape == (prop.n,) and np.all(wts > 0)
    plain = _grads(torch, scene, w, 16, range(300, 340), SECONDARY, d)
    learned = _grads(torch, scene, w, 16, range(400, 440), SECONDARY | _abi.DR_LEARNED_EDGES, d)
    se = np.sqrt(plain.var(0) / len(plain) + learned.var(0) / len(learned))
    assert np.linalg.norm(plain.mean(0) - learned.mean(0)) < 4.0 * np.linalg.norm(se) + 1e-6


def test_uniform_weights_give_the_plain_estimate(gpu):
    """Weights all 1: the learned node lengths equal the plain ones (up to
    rounding), the balance heuristic over the proposals is 1, and the same
    seed gives the plain gradient."""
    import torch
    w = helpers.small_workload("heightfield", res=32)
    scene = helpers.gpu_scene(w)
    W, H = w["resolution"]
    d = torch.from_numpy(np.random.default_rng(4).normal(size=(H, W, 3)).astype(np.float32)).to(gpu)
    prop = EdgeProposal(scene, seed=3)
    prop.set_weights(np.ones(prop.n, np.float32))
    plain = _grads(torch, scene, w, 4, [11, 12], SECONDARY, d)
    learned = _grads(torch, scene, w, 4, [11, 12], SECONDARY | _abi.DR_LEARNED_EDGES, d)
    assert np.abs(plain).sum() > 0
    for a, b in zip(plain, learned):
        assert np.linalg.norm(a - b) <= 1e-3 * np.linalg.norm(a), np.linalg.norm(a - b) / np.linalg.norm(a)


def _param_entries(w):
    """gradient entries of the heightfield's vertices and roughness texture
    (the parameters an optimisation of it changes)"""
    lay = dr.grad_layout(w["shapes"], w["materials"], w["lights"])
    out = []
    for kind, i, name in w["params"]:
        o, shape = (lay["shapes"] if kind == "shape" else lay["materials"])[i][name]
        out.append(np.arange(o, o + int(np.prod(shape))))
    return np.concatenate(out)


def test_trained_proposal_reduces_variance(gpu):
    """What the trained network gains over the method without it: on the
    heightfield (grid 128, 32 x 32 pixels, 4 spp, exact derivative), two
    rounds of statistics and training, then 512 renders of each estimator;
    the variance is that of the gradient of the heightfield's parameters.
    The secondary-edge estimator is heavy-tailed, so the variance is the
    median over 32 groups of the groups' variances; the learned one is lower
    (1.2x at this size, about 3.8x on the full 512 x 512, 64 spp
    heightfield), and the means agree."""
    import torch
    from delta_ray_amd import workloads
    w = workloads.heightfield(resolution=(32, 32), num_samples=4)
    scene = helpers.gpu_scene(w)
    W, H = w["resolution"]
    d = torch.from_numpy(np.random.default_rng(6).normal(size=(H, W, 3)).astype(np.float32)).to(gpu)
    prop = EdgeProposal(scene, seed=7, stat_shapes=[0])
    for rnd in range(2):
        prop.reset_stats()
        flags = SECONDARY | _abi.DR_EDGE_STATS | (_abi.DR_LEARNED_EDGES if rnd else 0)
        _grads(torch, scene, w, 4, range(50 + 100 * rnd, 66 + 100 * rnd), flags, d, base=0)
        prop.collect()
        prop.train(steps=300, batch=1024)
        prop.apply()
    sel = _param_entries(w)
    plain = _grads(torch, scene, w, 4, range(1000, 1512), SECONDARY, d, base=0)[:, sel]
    learned = _grads(torch, scene, w, 4, range(5000, 5512), SECONDARY | _abi.DR_LEARNED_EDGES, d, base=0)[:, sel]

    def mom_var(a, groups=32):
        m = len(a) // groups
        return float(np.median([((a[i * m:(i + 1) * m] - a[i * m:(i + 1) * m].mean(0)) ** 2).sum(1).sum() / (m - 1)
                                for i in range(groups)]))
    ratio = mom_var(plain) / mom_var(learned)
    print("variance ratio plain / learned", ratio)
    assert ratio > 1.1, ratio
    se = np.sqrt(plain.var(0, ddof=1) / len(plain) + learned.var(0, ddof=1) / len(learned))
    assert np.linalg.norm(plain.mean(0) - learned.mean(0)) < 4.0 * np.linalg.norm(se) + 1e-6

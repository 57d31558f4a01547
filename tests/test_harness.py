"""The inverse-rendering harness (the reference's optimisation scripts) on
the GPU: Adam through RenderFunction brings vertex parameters toward the
ones that rendered the target, and the camera tasks' loss down (a camera's
look-at point can slide along the view direction without changing the
image, so their parameter distance is not a measure)."""
import numpy as np
import pytest

from delta_ray_amd import harness

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,iters,shrink", [("single_triangle", 60, 0.5), ("single_triangle_camera", 60, 0.7),
                                               ("single_triangle_fisheye", 60, 0.5), ("shadow", 80, 0.7)])
def test_optimisation_approaches_the_target(gpu, tmp_path, name, iters, shrink):
    print("case: harness", name, flush=True)
    task = harness.TASKS[name]((64, 64))
    out = str(tmp_path / name) if name == "single_triangle" else None
    h = harness.optimize(task, iters=iters, spp=4, target_spp=64, out_dir=out)
    d0, d1 = h["param_distance"][0], h["param_distance"][-1]
    first, last = np.mean(h["loss"][:3]), np.mean(h["loss"][-5:])
    print("%s: distance %.4f -> %.4f, loss %.4g -> %.4g" % (name, d0, d1, first, last), flush=True)
    if "camera" in name or "fisheye" in name:
        assert last < shrink * first
    else:
        assert d1 < shrink * d0 and last < first
    if out:
        from delta_ray_amd import image
        t = image.imread(out + "/target.exr")
        assert t.shape == (64, 64, 3) and t.sum() > 0


def test_heightfield_gan_steps(gpu, tmp_path):
    """Two WGAN-GP iterations (reference heightfield_gan.py): finite costs,
    and the generator's gradient arrives through the renderer (its weights
    move after the generator step)."""
    import torch
    from delta_ray_amd import heightfield_gan as hg
    print("case: heightfield gan", flush=True)
    torch.manual_seed(0)
    G = hg.Generator().cuda()
    render = hg.Renderer(torch.device("cuda"))
    z = torch.randn(2, hg.NOISE, device="cuda")
    imgs = render(G(z), [0.0, 1.0])
    assert imgs.shape == (2, 1, hg.RES, hg.RES) and float(imgs.sum()) > 0
    imgs.sum().backward()
    g = G.out32.weight.grad
    assert g is not None and float(g.abs().sum()) > 0
    h = hg.train(iters=2, batch=2, out_dir=str(tmp_path), log=None)
    assert all(np.isfinite(h["d_cost"])) and all(np.isfinite(h["g_cost"]))
    assert (tmp_path / "generated_000001.png").exists()

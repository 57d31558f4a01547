"""GPU vs oracle on materials the other parity scenes leave out.

* A glossy plane whose diffuse reflectance and roughness are 8 x 8 textures
  (bilinear lookups, material.h:108-239), under a light and a blocker: the
  interior term both with the reference's texel-corner swap
  (DR_QUIRK_TEXEL_SWAP, delta_ray.cpp:1364-1424, part of the parity flags)
  and with every quirk bit clear.
* A glossy panel seen and lit from below, its back side: black when
  one-sided, shaded when two-sided (material.h:257-351).

The forward image and every adjoint term are compared per parameter group."""
import numpy as np
import pytest

from delta_ray_amd import _abi, workloads
from delta_ray_amd import delta_ray as dr
from tests import helpers
from tests.helpers import gpu_render, gpu_scene, rel_err
from tests.test_gpu_parity import BASE, FLOOR, RTOL, TERMS, _print_case

pytestmark = pytest.mark.gpu

TEX_TERMS = dict(TERMS, interior_exact=TERMS["interior"])
TEX_BASE = dict(BASE, interior_exact=0)
TEX_RTOL = dict(RTOL, interior_exact=RTOL["interior"])
TEX_FLOOR = dict(FLOOR, interior_exact=FLOOR["interior"])


def test_textured_forward_matches_oracle(gpu, oracle):
    _print_case("forward textured glossy plane")
    w = helpers.glossy_plane(48, textured=True)
    W, H = w["resolution"]
    nb = w["max_bounces"]
    img, _, _ = gpu_render(gpu_scene(w), w, 2, nb, 7)
    ref, _, _ = oracle.scene_from_workload(w).render(W, H, 2, nb, 7)
    err = np.abs(img - ref)
    tol = 1e-4 * np.maximum(np.abs(ref), 1e-2)
    assert float(np.mean(np.any(err > tol, axis=2))) <= 0.005
    assert rel_err(img, ref) < 2e-2


@pytest.mark.parametrize("term", list(TEX_TERMS))
def test_textured_adjoint_matches_oracle(gpu, oracle, term):
    _print_case("adjoint textured glossy plane", term)
    w = helpers.glossy_plane(48, textured=True)
    W, H = w["resolution"]
    nb = w["max_bounces"]
    d_img = np.random.default_rng(3).normal(size=(H, W, 3)).astype(np.float32)
    flags = TEX_BASE.get(term, _abi.DR_FLAGS_PARITY) | TEX_TERMS[term]
    _, g, _ = gpu_render(gpu_scene(w), w, 2, nb, 5, flags=flags, d_image=d_img)
    _, gref, _ = oracle.scene_from_workload(w).render(W, H, 2, nb, 5, flags=flags, d_image=d_img)
    errs = helpers.assert_groups(w, g, gref, TEX_RTOL[term], TEX_FLOOR[term])
    if term.startswith("interior"):
        # the texture groups carry gradient and are compared
        assert errs["material0.roughness"][1] > 0 and errs["material0.diffuse_reflectance"][1] > 0


def back_lit_panel(two_sided, res=48):
    """a glossy panel facing +y, seen and lit from below; a blocker between
    the light and the panel"""
    pv, pi = workloads.quad([0, 0, 0], [0, 1, 0], 3.0)
    kv, ki = workloads.quad([0.4, -0.6, 0.3], [0.2, -1, 0.1], 0.7)
    lv, li = workloads.quad([-0.3, -2.0, 0.2], [0, 1, 0], 0.5)
    panel = workloads.diffuse_material([0.6, 0.55, 0.5], specular=(0.3, 0.3, 0.3), roughness=(0.15,),
                                       two_sided=two_sided)
    grey = workloads.diffuse_material([0.5] * 3)
    black = workloads.diffuse_material([0.0] * 3)
    ls = dr.Shape(lv, li, None, None, black, None)
    light = dr.Light(ls, np.array([15.0] * 3, np.float32))
    ls.light = light
    cam = workloads.make_camera([0.5, -2.2, -2.2], [0, 0, 0.3], [0, 1, 0], 55.0, (res, res))
    shapes = [dr.Shape(pv, pi, None, None, panel, None), dr.Shape(kv, ki, None, None, grey, None), ls]
    return dict(camera=cam, shapes=shapes, materials=[panel, grey, black], lights=[light],
                resolution=(res, res), max_bounces=2)


@pytest.mark.parametrize("two_sided", [False, True])
def test_back_side_forward_matches_oracle(gpu, oracle, two_sided):
    _print_case("forward back-lit panel, two-sided", two_sided)
    w = back_lit_panel(two_sided)
    W, H = w["resolution"]
    img, _, _ = gpu_render(gpu_scene(w), w, 2, 2, 7)
    ref, _, _ = oracle.scene_from_workload(w).render(W, H, 2, 2, 7)
    err = np.abs(img - ref)
    tol = 1e-4 * np.maximum(np.abs(ref), 1e-2)
    assert float(np.mean(np.any(err > tol, axis=2))) <= 0.005
    assert rel_err(img, ref) < 2e-2
    # the panel's back side is shaded only when two-sided
    one, _, _ = oracle.scene_from_workload(back_lit_panel(False)).render(W, H, 2, 2, 7)
    assert (ref.sum() > 5 * one.sum()) == two_sided


@pytest.mark.parametrize("two_sided", [False, True])
@pytest.mark.parametrize("term", list(TERMS))
def test_back_side_adjoint_matches_oracle(gpu, oracle, two_sided, term):
    _print_case("adjoint back-lit panel", term, "two-sided", two_sided)
    w = back_lit_panel(two_sided)
    W, H = w["resolution"]
    d_img = np.random.default_rng(3).normal(size=(H, W, 3)).astype(np.float32)
    flags = BASE.get(term, _abi.DR_FLAGS_PARITY) | TERMS[term]
    _, g, _ = gpu_render(gpu_scene(w), w, 2, 2, 5, flags=flags, d_image=d_img)
    _, gref, _ = oracle.scene_from_workload(w).render(W, H, 2, 2, 5, flags=flags, d_image=d_img)
    errs = helpers.assert_groups(w, g, gref, RTOL[term], FLOOR[term])
    if term == "interior":
        # the panel's material gets gradient exactly when its back is shaded
        assert (errs["material0.diffuse_reflectance"][1] > 0) == two_sided

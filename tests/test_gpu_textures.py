"""GPU vs oracle on textured materials: a glossy plane whose diffuse
reflectance and roughness are 8 x 8 textures (bilinear lookups,
material.h:108-239), under a light and a blocker.  The forward image and
every adjoint term are compared per parameter group, texture groups
included; the interior term both with the reference's texel-corner swap
(DR_QUIRK_TEXEL_SWAP, delta_ray.cpp:1364-1424, part of the parity flags)
and with every quirk bit clear."""
import numpy as np
import pytest

from delta_ray_amd import _abi
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

"""A reference scene read from its files (the reference's bunny_box.xml and
bunny_box.serialized: a Cornell box with the Stanford bunny, 14 416
triangles, one area light) rendered by the HIP library and by the oracle:
the same per-pixel and per-parameter-group bars as the procedural scenes."""
import os

import numpy as np
import pytest

from delta_ray_amd import _abi, workloads
from tests import helpers
from tests.helpers import gpu_render, gpu_scene, rel_err
from tests.test_gpu_parity import BASE, FLOOR, RTOL, TERMS

pytestmark = pytest.mark.gpu

SCENE = os.path.join(os.path.dirname(__file__), "golden", "scenes", "bunny_box.xml")


def _workload():
    return workloads.mitsuba(SCENE, resolution=(32, 32), num_samples=4, max_bounces=2)


def test_bunny_box_forward_matches_oracle(gpu, oracle):
    print("case: bunny_box forward", flush=True)
    w = _workload()
    img, _, _ = gpu_render(gpu_scene(w), w, 4, 2, 7)
    ref, _, _ = oracle.scene_from_workload(w).render(32, 32, 4, 2, 7)
    assert ref.sum() > 0
    err = np.abs(img - ref)
    tol = 1e-4 * np.maximum(np.abs(ref), 1e-2)
    assert float(np.mean(np.any(err > tol, axis=2))) <= 0.005
    assert rel_err(img, ref) < 2e-2


@pytest.mark.parametrize("term", ["interior", "secondary", "primary"])
def test_bunny_box_adjoint_matches_oracle(gpu, oracle, term):
    print("case: bunny_box adjoint", term, flush=True)
    w = _workload()
    d = np.random.default_rng(3).normal(size=(32, 32, 3)).astype(np.float32)
    flags = BASE.get(term, _abi.DR_FLAGS_PARITY) | TERMS[term]
    _, g, _ = gpu_render(gpu_scene(w), w, 4, 2, 5, flags=flags, d_image=d)
    _, gref, _ = oracle.scene_from_workload(w).render(32, 32, 4, 2, 5, flags=flags, d_image=d)
    assert np.abs(gref).sum() > 0
    helpers.assert_groups(w, g, gref, RTOL[term], FLOOR[term])

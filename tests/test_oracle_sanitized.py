"""The CPU oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
(oracle/Makefile, target `sanitize`) runs the renders the parity tests use --
forward and adjoint, every term, on the single triangle, the shadow, the
glossy plane and the small heightfield, plus ray queries with and without
its BVH -- in a child Python process with the sanitizer runtime preloaded:
no invalid access, leak-free teardown aside, and no undefined behaviour."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r'''
import sys
sys.path.insert(0, %r)
import numpy as np
from tests import oracle_lib, helpers
from delta_ray_amd import _abi
for name in ("single_triangle", "shadow", "glossy_plane", "heightfield"):
    w = helpers.small_workload(name, res=16)
    o = oracle_lib.scene_from_workload(w)
    W, H = w["resolution"]
    d = np.random.default_rng(0).normal(size=(H, W, 3)).astype(np.float32)
    for flags in (_abi.DR_FLAGS_PARITY, 0):
        img, g, _ = o.render(W, H, 2, w["max_bounces"], 7, flags=flags, d_image=d)
        assert np.isfinite(img).all() and np.isfinite(g).all()
    rays = helpers.random_rays(w, 2000, np.random.default_rng(1))
    a = o.intersect(rays)
    o.use_bvh(False)
    b = o.intersect(rays)
    assert (a == b).all()
    print("ok", name, flush=True)
'''


def _asan_runtime():
    try:
        out = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True, timeout=30)
    except (OSError, subprocess.TimeoutExpired):
        return None
    p = out.stdout.strip()
    return p if p and os.path.isabs(p) and os.path.exists(p) else None


def test_oracle_under_sanitizers(tmp_path):
    rt = _asan_runtime()
    if rt is None:
        pytest.skip("no AddressSanitizer runtime for g++ on this host")
    so = str(tmp_path / "liboracle_san.so")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize", "OUT=" + so],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    with open(so, "rb") as f:
        blob = f.read()
    assert b"__asan_" in blob and b"__ubsan_" in blob, "the oracle was built without the sanitizers"
    # the sanitizer runtime first, ahead of whatever the environment already
    # preloads (kept as it is)
    pre = os.environ.get("LD_PRELOAD", "")
    env = dict(os.environ, LD_PRELOAD=rt + (":" + pre if pre else ""), DR_ORACLE_SO=so,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:exitcode=23",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=24")
    r = subprocess.run([sys.executable, "-c", CHILD % ROOT], capture_output=True, text=True, timeout=900, env=env)
    log = r.stdout + r.stderr
    assert r.returncode == 0, log[-4000:]
    assert "AddressSanitizer" not in log and "runtime error:" not in log, log[-4000:]
    assert log.count("ok ") == 4, log[-2000:]

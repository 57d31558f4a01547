"""The oracle pins bite: each plants one plausible one-line mistake in the
oracle (oracle.cpp, OR_MUTANT) and checks that the pin of that function fails
on the planted build and passes on the real one (CPU only):

  1  the path throughput drops the BSDF pdf   -> two-bounce quadrature
  2  the specular edge sample uses the diffuse pmf -> glossy edge quadrature
  3  an edge sample's continuation drops the hit cosine -> continuation quadrature
  4  the node bound drops the absolute value of its azimuth term -> bound closed form
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")

MUTANTS = {
    1: "test_two_bounce_forward_quadrature",
    2: "test_secondary_edges_glossy_exact",
    3: "test_edge_continuation_paths",
    4: "test_node_bound_closed_forms",
}


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    d = tmp_path_factory.mktemp("mutants")

    def build(k):
        out = str(d / ("liboracle_m%d.so" % k))
        subprocess.check_call(["make", "-s", "-C", ORACLE_DIR, "mutant", "MUTANT=%d" % k, "OUT=" + out])
        return k, out

    with ThreadPoolExecutor(max_workers=4) as ex:
        return dict(ex.map(build, MUTANTS))


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_pin_fails_on_planted_mistake(mutant_libs, k):
    env = dict(os.environ, DR_ORACLE_SO=mutant_libs[k])
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        "tests/test_oracle_pins.py::" + MUTANTS[k]],
                       cwd=ROOT, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    print(r.stdout[-1500:])
    assert r.returncode == 1, "the pin did not fail on mutant %d" % k
    assert "AssertionError" in r.stdout and "1 failed" in r.stdout

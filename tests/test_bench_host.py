"""Host-side parts of bench.py: the CPU share it runs the CPU leg on, the
CPU leg itself on a tiny scene, and the writer of the results table."""
import json
import os

import numpy as np

import bench
from delta_ray_amd import workloads


def test_host_cpus_is_the_process_share():
    threads, host = bench.host_cpus()
    assert 1 <= threads <= host["affinity_cpus"] <= host["nproc"]
    if host["omp_num_threads"]:
        assert threads <= host["omp_num_threads"]


def test_cpu_leg_counts_samples_and_rays():
    w = workloads.single_triangle(resolution=(16, 16), num_samples=2)
    d = np.ones((16, 16, 3), np.float32)
    r = bench.cpu_leg(w, d, 2, 0.05)
    assert 1 <= r["spp"] <= 2
    assert r["samples_per_s"] > 0
    # at least the camera ray of every pixel sample
    assert r["rays"] >= r["spp"] * 16 * 16


def test_fill_baseline_writes_the_table(tmp_path):
    base = tmp_path / "results.md"
    base.write_text("# x\n\n### Results table (filled by the build's measurement harness, never by hand)\n\n"
                    "| config | mode | a |\n|---|---|---|\n| C1 | fwd | TBD |\n\n## 3. next\n")
    roof = {"kernel": "k_select", "frac": 0.1, "traffic": 3e9, "achieved": 800.0}
    row = lambda name: {"config": name, "workload": name, "spp": 4, "W": 8, "H": 8, "bounces": 1,
                        "fwd": {"gpu_samples_per_s": 1e9, "gpu_rays_per_s": 2e9, "cpu_samples_per_s": 1e6,
                                "cpu_rays_per_s": 2e6, "cpu_sample": "4 spp, 1.0 s"},
                        "adj": {"gpu_samples_per_s": 1e8, "gpu_rays_per_s": 5e8, "cpu_samples_per_s": 1e5,
                                "cpu_rays_per_s": 5e5, "cpu_sample": "4 spp, 2.0 s", "roofline": roof}}
    t = {"cpu_threads": 16, "host": {"model": "cpu", "nproc": 64},
         "rows": [row("single_triangle"), row("shadow"), row("heightfield")]}
    js = tmp_path / "t.json"
    js.write_text(json.dumps(t))
    bench.fill_baseline(str(js), str(base))
    text = base.read_text()
    assert "TBD" not in text and text.endswith("## 3. next\n")
    rows = [l for l in text.splitlines() if l.startswith("| C")]
    assert len(rows) == 6 and all("1000×" in l for l in rows)

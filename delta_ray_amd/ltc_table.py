"""Regenerate the LTC table the secondary-edge sampler reads (dr_ltc_fit, the
reference's fit/fitLTC.cpp) and compare it with the embedded one
(data/ltc_tabM.f32, the reference's ltc.inc).

    python -m delta_ray_amd.ltc_table --out ltc_tabM.f32 [--consistent] [--threads N]

--consistent fits the lobe the renderer evaluates at roughness r (Phong
exponent 2 / r) instead of the reference's alpha = r^2; the result is a table
of the same layout (128 x 128 entries of 9 floats, row-major 3x3, entry
a + 128 t for roughness index a and view angle index t).
"""
import argparse
import ctypes
import json
import time

import numpy as np

from . import _abi

N = 128


def fit(n=N, rows=None, ts=None, guess=None, consistent=False, threads=0):
    """fitted entries (tab [n*n, 9], amplitude [n*n]); rows / ts: ranges of
    roughness and view-angle indices (the whole table by default); guess: the
    table that gives first guesses to a partial fit"""
    lib = _abi.load_library()
    fp = ctypes.POINTER(ctypes.c_float)
    tab = np.zeros((n * n, 9), np.float32)
    amp = np.zeros(n * n, np.float32)
    r0, r1 = rows if rows else (0, n)
    t0, t1 = ts if ts else (0, n)
    g = None if guess is None else np.ascontiguousarray(guess, np.float32).reshape(n * n, 9)
    _abi.check(lib, lib.dr_ltc_fit(n, r0, r1, t0, t1, None if g is None else g.ctypes.data_as(fp),
                                   _abi.DR_LTC_FIT_CONSISTENT if consistent else 0, int(threads),
                                   tab.ctypes.data_as(fp), amp.ctypes.data_as(fp)), "dr_ltc_fit")
    return tab, amp


def embedded():
    return np.fromfile(_abi.LTC_PATH, np.float32).reshape(-1, 9)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--consistent", action="store_true")
    ap.add_argument("--threads", type=int, default=0)
    a = ap.parse_args()
    t = time.perf_counter()
    tab, amp = fit(consistent=a.consistent, threads=a.threads)
    dt = time.perf_counter() - t
    tab.tofile(a.out)
    d = np.abs(tab - embedded()).max(axis=1)
    print(json.dumps({"seconds": round(dt, 1), "consistent": a.consistent,
                      "max_abs_diff_vs_embedded": float(d.max()),
                      "entries_within_1e-4": float(np.mean(d <= 1e-4)),
                      "entries_within_1e-3": float(np.mean(d <= 1e-3)),
                      "median_max_abs_diff": float(np.median(d))}))


if __name__ == "__main__":
    main()

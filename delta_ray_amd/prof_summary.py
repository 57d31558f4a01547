"""Summarise rocprofv3 results of bench.py runs (rocpd SQLite databases).

    python -m delta_ray_amd.prof_summary --kt KT.db [--fetch F.db --write W.db] [--sq SQ.db] [--sq2 SQ2.db]
                                         [--warmup W --steps K] [--copy-bytes N]
    python -m delta_ray_amd.prof_summary --mfma MFMA.db     # the network kernels (python -m delta_ray_amd.learned)

bench.py renders, in order: two forward renders (the d_image of the loss),
`warmup` + `steps` forward+adjoint renders (the timed ones are the last
`steps`), then three counter renders.  Every forward+adjoint render ends with
one k_primary launch, so the timed renders are the dispatches after the
`warmup`-th k_primary up to and including the (`warmup` + `steps`)-th; the
summary is restricted to them, so its per-launch averages are comparable
with the ones bench.py measures with HIP events.

HBM traffic: FETCH_SIZE and WRITE_SIZE (KiB per dispatch) scaled by what the
same counters read for the `--calibrate` copy of bench.py (a device-to-device
copy of `copy-bytes` bytes: it reads and writes that many bytes).

Second SQ pass: active lanes per VALU instruction = SQ_THREAD_CYCLES_VALU /
SQ_INSTS_VALU; the fractions of wave-cycles issuing VALU (SQ_ACTIVE_INST_VALU)
and waiting (SQ_WAIT_ANY) over SQ_WAVE_CYCLES.

MFMA: per network kernel, bf16 matrix flops = SQ_INSTS_VALU_MFMA_MOPS_BF16 x
512, over the dispatch's duration; matrix-core busy fraction =
SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE x 1024 SIMDs).
"""
import argparse
import collections
import sqlite3


def _short(name):
    n = name.split("(")[0]
    return n[5:] if n.startswith("void ") else n


def _dispatches(db):
    """[(dispatch_id, kernel, start, end)] in launch order."""
    con = sqlite3.connect(db)
    rows = con.execute("select dispatch_id, name, start, end from kernels order by start").fetchall()
    con.close()
    return [(d, _short(n), s, e) for d, n, s, e in rows]


def _window(disp, warmup, steps):
    """Dispatch ids of the timed renders."""
    prim = [i for i, r in enumerate(disp) if r[1].startswith("k_primary")]
    if len(prim) < warmup + steps:
        raise SystemExit("found %d k_primary launches, expected at least %d" % (len(prim), warmup + steps))
    lo = prim[warmup - 1] + 1 if warmup > 0 else 0
    hi = prim[warmup + steps - 1]
    return {disp[i][0] for i in range(lo, hi + 1)}


def _counters(db, name):
    con = sqlite3.connect(db)
    rows = con.execute("select dispatch_id, kernel_name, value, start, end from counters_collection "
                       "where counter_name = ? order by start", (name,)).fetchall()
    con.close()
    per = collections.OrderedDict()
    for d, k, v, s, e in rows:
        if d not in per:
            per[d] = [_short(k), 0.0, s, e]
        per[d][1] += v
    return [(d, k, v, s, e) for d, (k, v, s, e) in per.items()]


def _pmc_window(rows, warmup, steps):
    disp = [(d, k, s, e) for d, k, v, s, e in rows]
    return _window(disp, warmup, steps)


def _copy_kib(rows, copy_bytes):
    """Counter value of the calibration copy: the dispatch whose value is
    closest to copy_bytes / 1024 among the copy kernels."""
    want = copy_bytes / 1024.0
    cands = [v for d, k, v, s, e in rows if "copyBuffer" in k or "copy" in k.lower()]
    if not cands:
        return None
    return min(cands, key=lambda v: abs(v - want))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--kt")
    ap.add_argument("--sq2")
    ap.add_argument("--mfma")
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--sq")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--pmc-warmup", type=int, default=0)
    ap.add_argument("--pmc-steps", type=int, default=1)
    ap.add_argument("--copy-bytes", type=int, default=(64 << 20) * 4)
    a = ap.parse_args()
    if a.mfma:
        mfma_summary(a.mfma)
        if not a.kt:
            return
    if not a.kt:
        raise SystemExit("--kt (or --mfma) is required")

    disp = _dispatches(a.kt)
    win = _window(disp, a.warmup, a.steps)
    agg = collections.OrderedDict()
    for d, k, s, e in disp:
        if d not in win:
            continue
        t = agg.setdefault(k, [0, 0.0])
        t[0] += 1
        t[1] += (e - s) / 1e3
    total = sum(v[1] for v in agg.values())
    print("kernel-trace, timed renders only (%d warm-up, %d timed):" % (a.warmup, a.steps))
    print("  %-14s %8s %12s %12s %7s" % ("kernel", "launches", "avg_us", "ms/render", "share"))
    for k, (n, us) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        if us / total < 1e-4:
            continue
        print("  %-14s %8d %12.1f %12.3f %6.1f%%" % (k, n, us / n, us / 1e3 / a.steps, 100.0 * us / total))

    traffic = {}
    if a.fetch and a.write:
        f_rows, w_rows = _counters(a.fetch, "FETCH_SIZE"), _counters(a.write, "WRITE_SIZE")
        f_cal, w_cal = _copy_kib(f_rows, a.copy_bytes), _copy_kib(w_rows, a.copy_bytes)
        want = a.copy_bytes / 1024.0
        print("\ncalibration copy of %d bytes: FETCH_SIZE %.0f KiB (x%.3f), WRITE_SIZE %.0f KiB (x%.3f)" %
              (a.copy_bytes, f_cal, f_cal / want, w_cal, w_cal / want))
        for name, rows, cal in (("FETCH", f_rows, f_cal), ("WRITE", w_rows, w_cal)):
            wset = _pmc_window(rows, a.pmc_warmup, a.pmc_steps)
            per = collections.defaultdict(list)
            for d, k, v, s, e in rows:
                if d in wset:
                    per[k].append(v)
            for k, vals in per.items():
                t = traffic.setdefault(k, {"FETCH": 0.0, "WRITE": 0.0, "n": len(vals)})
                t[name] = sum(vals) / len(vals) * 1024.0 * (want / cal)
        print("HBM traffic per launch (calibrated), timed render of the counter runs:")
        print("  %-14s %8s %14s %14s %14s" % ("kernel", "launches", "fetch_bytes", "write_bytes", "total_bytes"))
        for k, t in sorted(traffic.items(), key=lambda kv: -(kv[1]["FETCH"] + kv[1]["WRITE"])):
            tot = t["FETCH"] + t["WRITE"]
            if tot < 1e6:
                continue
            print("  %-14s %8d %14.4g %14.4g %14.4g" % (k, t["n"], t["FETCH"], t["WRITE"], tot))

    if a.sq:
        con = sqlite3.connect(a.sq)
        rows = con.execute("select dispatch_id, kernel_name, counter_name, value, start from counters_collection "
                           "order by start").fetchall()
        con.close()
        disp_sq = collections.OrderedDict()
        for d, k, c, v, s in rows:
            cs = disp_sq.setdefault(d, [_short(k), s, {}])[2]
            cs[c] = cs.get(c, 0.0) + v
        dl = [(d, k, s, 0) for d, (k, s, _) in disp_sq.items()]
        wset = _window(dl, a.pmc_warmup, a.pmc_steps)
        per = collections.defaultdict(lambda: collections.defaultdict(float))
        for d, (k, s, cs) in disp_sq.items():
            if d in wset:
                for c, v in cs.items():
                    per[k][c] += v
        print("\nSQ instruction mix, timed render of the counter run (per wave):")
        print("  %-14s %10s %12s %10s %10s %10s %10s" % ("kernel", "waves", "VALU", "SALU", "VMEM_RD", "VMEM_WR",
                                                         "LDS"))
        for k, cs in sorted(per.items(), key=lambda kv: -kv[1].get("SQ_INSTS_VALU", 0)):
            w = cs.get("SQ_WAVES", 0)
            if w <= 0 or cs.get("SQ_INSTS_VALU", 0) < 1e6:
                continue
            print("  %-14s %10.0f %12.0f %10.0f %10.0f %10.0f %10.0f" % (
                k, w, cs.get("SQ_INSTS_VALU", 0) / w, cs.get("SQ_INSTS_SALU", 0) / w, cs.get("SQ_INSTS_VMEM_RD", 0) / w,
                cs.get("SQ_INSTS_VMEM_WR", 0) / w, cs.get("SQ_INSTS_LDS", 0) / w))
    if a.sq2:
        sq2_summary(a.sq2, a.pmc_warmup, a.pmc_steps)


def _per_dispatch(db):
    con = sqlite3.connect(db)
    rows = con.execute("select dispatch_id, kernel_name, counter_name, value, start, end from counters_collection "
                       "order by start").fetchall()
    con.close()
    disp = collections.OrderedDict()
    for d, k, c, v, s, e in rows:
        x = disp.setdefault(d, [_short(k), s, e, {}])
        x[3][c] = x[3].get(c, 0.0) + v
    return disp


def sq2_summary(db, warmup, steps):
    disp = _per_dispatch(db)
    wset = _window([(d, k, s, e) for d, (k, s, e, _) in disp.items()], warmup, steps)
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    for d, (k, s, e, cs) in disp.items():
        if d in wset:
            for c, v in cs.items():
                per[k][c] += v
    print("\nSecond SQ pass (fractions of wave-cycles; active lanes per VALU instruction, of 64):")
    print("  %-14s %10s %12s %9s" % ("kernel", "lanes/VALU", "VALU-active", "waiting"))
    for k, cs in sorted(per.items(), key=lambda kv: -kv[1].get("SQ_INSTS_VALU", 0)):
        if cs.get("SQ_INSTS_VALU", 0) < 1e8 or cs.get("SQ_WAVE_CYCLES", 0) <= 0:
            continue
        print("  %-14s %10.1f %12.3f %9.3f" % (k, cs["SQ_THREAD_CYCLES_VALU"] / cs["SQ_INSTS_VALU"],
                                               cs["SQ_ACTIVE_INST_VALU"] / cs["SQ_WAVE_CYCLES"],
                                               cs["SQ_WAIT_ANY"] / cs["SQ_WAVE_CYCLES"]))


def mfma_summary(db):
    disp = _per_dispatch(db)
    per = collections.defaultdict(lambda: [0, 0.0, 0.0, 0.0, 0.0])
    for d, (k, s, e, cs) in disp.items():
        if "mlp" not in k:
            continue
        t = per[k]
        t[0] += 1
        t[1] += (e - s) * 1e-9
        t[2] += cs.get("SQ_INSTS_VALU_MFMA_MOPS_BF16", 0.0) * 512
        t[3] += cs.get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0)
        t[4] += cs.get("GRBM_GUI_ACTIVE", 0.0)
    print("MFMA, network kernels (bf16; peak 2.5 PFLOP/s dense):")
    print("  %-14s %8s %10s %12s %12s %10s" % ("kernel", "launches", "avg_us", "MFLOP/launch", "TFLOP/s", "MFMA busy"))
    for k, (n, sec, fl, busy, gui) in sorted(per.items()):
        print("  %-14s %8d %10.1f %12.3f %12.3f %10.4f" % (k, n, sec / n * 1e6, fl / n / 1e6,
                                                          fl / sec / 1e12 if sec > 0 else 0.0,
                                                          busy / (gui * 1024) if gui > 0 else 0.0))


if __name__ == "__main__":
    main()

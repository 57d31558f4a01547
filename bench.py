"""Benchmark: forward + adjoint samples/s of the differentiable path tracer.

Workload (default): the procedural glossy heightfield, 128 x 128 grid
(32 258 triangles + 2 light quads), 512 x 512 pixels, 64 samples per pixel,
2 bounces, fp32.  One step = one forward+adjoint render (the reference's
RenderFunction.backward call: pixel samples with interior adjoint and
secondary edges, then spp*W*H primary-edge samples), with scene, d_image and
buffers resident on the GPU before the timed region.

    python bench.py [--steps K] [--warmup W]
                    [--workload heightfield|shadow|single_triangle|batched|learned|torch_step]
    python bench.py --table OUT.json        # the three single-scene configs, fwd and fwd+adjoint, GPU and CPU
    python bench.py --fill-baseline OUT.json RESULTS.md  # writes that table into RESULTS.md (no GPU)

`--workload batched` times the uneven batch instead (64 scenes of 1k-200k
triangles x 4 views, forward + adjoint each, one graph-captured batch per
step, pass sizes from the autotuner).  `--workload learned` measures the time
to a gradient relative RMSE of 1e-2 with the learned edge proposal and with
the plain one (see learned_leg) at a size of its own.

Prints one JSON line.  `roofline` is for the dominant kernel (the one with
the most device time per step): algorithmic bytes per launch (BVH traversal +
edge sampling + image/d_image I/O, from the kernels' own counters in separate
untimed runs) divided by its average launch time measured with HIP events on
its launch stream during the timed steps; `traffic` is that kernel's HBM bytes
per launch from the committed rocprofv3 summary of the same workload
(profiles/), null when none covers it.  `cpu_baseline` times the CPU
restatement of the reference (oracle/, 8x8 pixel tiles from a shared work
list as the reference's thread pool) on the host's CPU share on a bounded
sample of the same workload.  `extra.learned_vs_plain` is the learned
proposal against the plain one on the same workload (a bounded number of
renders; see learned_leg).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "fwd+adjoint samples/s; time to gradient rel-RMSE 1e-2, learned vs plain"

# algorithmic bytes: the records each unit of work must read or write once
B_RAY_IO = 32 + 16      # ray in (org, dir, tmin, tmax) + hit out (ids, u, v)
B_NODE = 64             # one BVH2 node (two child boxes)
B_TRI = 48              # one triangle (v0, e1, e2)
B_PIXEL_IO = 12 + 12    # d_image read + image write per pixel sample
B_EDGE = 24 + 24 + 24 + 24 + 36  # edge v0/v1, faces, non-shared verts, normals, LTC matrix
B_POINT = 108                   # one shading-point record (read by selection / edge samples)
B_TREE_EXPAND = 2 * 56          # an expanded edge-tree node reads its two children (DrEdgeNode: 56 B)
B_EDGE_TEST = 20 + 24 + 24 + 24  # edge record, its endpoints, faces' third vertices + indices
B_QUEUE = 16                    # one selected-edge queue entry written


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="heightfield")
    ap.add_argument("--scene", metavar="XML", help="time a Mitsuba scene file instead (load_mitsuba), at its "
                    "own resolution, --scene-spp samples per pixel and --scene-bounces bounces")
    ap.add_argument("--scene-spp", type=int, default=16)
    ap.add_argument("--scene-bounces", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="CPU work of the bounded CPU sample (its spp is chosen from a 1-spp run)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-plain leg")
    ap.add_argument("--table", metavar="OUT_JSON", help="time the three single-scene configs (fwd and "
                    "fwd+adjoint, GPU and CPU) and write them to OUT_JSON")
    ap.add_argument("--fill-baseline", nargs=2, metavar=("TABLE_JSON", "MARKDOWN"),
                    help="rewrite the results table of MARKDOWN from a --table run (no GPU)")
    ap.add_argument("--batch-scenes", type=int, default=64, help="scenes of the batched workload")
    ap.add_argument("--no-autotune", action="store_true", help="batched workload: default pass sizes")
    ap.add_argument("--lv-res", type=int, default=128, help="learned workload: image size")
    ap.add_argument("--lv-spp", type=int, default=16, help="learned workload: samples per pixel")
    ap.add_argument("--lv-renders", type=int, default=512, help="learned workload: renders per estimator")
    ap.add_argument("--lv-train-renders", type=int, default=8, help="learned: statistics renders per round")
    ap.add_argument("--lv-train-steps", type=int, default=300, help="learned: MLP training steps per round")
    ap.add_argument("--lv-rounds", type=int, default=2, help="learned: rounds of statistics + training")
    ap.add_argument("--lv-eps", type=float, default=1e-2, help="learned: target relative RMSE")
    ap.add_argument("--lv-bench-renders", type=int, default=96,
                    help="renders per estimator of the learned-vs-plain leg of the default run")
    ap.add_argument("--lv-blocks", type=int, default=1,
                    help="learned: disjoint seed blocks of renders per estimator (each of --lv-renders, or "
                         "--lv-bench-renders in the default and --table runs); variances pool the blocks' "
                         "groups, and each block's ratios are reported too")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run one 256 MiB device-to-device copy (a kernel of known HBM traffic: "
                         "256 MiB read + 256 MiB written) to scale hardware byte counters")
    return ap.parse_args()


def host_cpus():
    """(threads for the CPU leg, what the host reports).  The threads are the
    CPUs this process may use: its affinity mask, bounded by a cgroup CPU
    quota and by OMP_NUM_THREADS when the host sets one (the CPU share of
    this process, on a machine shared with other jobs)."""
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = nproc
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                quota = max(1, -(-int(q) // int(per)))
    except (OSError, ValueError):
        pass
    omp = None
    try:
        omp = int(os.environ.get("OMP_NUM_THREADS", ""))
    except ValueError:
        pass
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = min(x for x in (aff, quota, omp) if x)
    return threads, {"nproc": nproc, "affinity_cpus": aff, "cgroup_quota_cpus": quota,
                     "omp_num_threads": omp, "model": model}


def make_workload(name):
    from delta_ray_amd import workloads
    return workloads.WORKLOADS[name]()


def param_index(w):
    """Indices, in the flat gradient buffer, of the parameters the workload
    optimises (w["params"]: ("shape", i, "vertices" | "normals") or
    ("material", i, "diffuse_reflectance" | "specular_reflectance" |
    "roughness")); the rest of the scene (lights' quads, camera) is fixed."""
    import numpy as np
    from delta_ray_amd import delta_ray as dr
    lay = dr.grad_layout(w["shapes"], w["materials"], w["lights"])
    idx = []
    for kind, i, name in w["params"]:
        o, shape = (lay["shapes"] if kind == "shape" else lay["materials"])[i][name]
        idx.append(np.arange(o, o + int(np.prod(shape))))
    return np.concatenate(idx).astype(np.int64)


def d_image_for(w, scene, dev):
    """dLoss/dImage of an L2 loss against a target rendered from perturbed
    parameters (as the reference's optimisation scripts do)."""
    import numpy as np
    import torch
    from delta_ray_amd import delta_ray as dr
    W, H = w["resolution"]
    spp, nb = w["num_samples"], w["max_bounces"]
    rng = np.random.default_rng(101)
    shapes = []
    pert = {i for kind, i, _ in w["params"] if kind == "shape"}
    for i, s in enumerate(w["shapes"]):
        v = np.asarray(s.vertices).copy()
        if i in pert:
            v += rng.normal(0.0, 0.05, size=v.shape).astype(np.float32)
        ns = dr.Shape(v, s.indices, s.uvs, s.normals, s.material, None)
        shapes.append(ns)
    lights = []
    for l in w["lights"]:
        k = w["shapes"].index(l.shape)
        nl = dr.Light(shapes[k], l.intensity)
        shapes[k].light = nl
        lights.append(nl)
    tscene = dr.DeviceScene(w["camera"], shapes, w["materials"], lights)
    o = dr.DeviceScene.options(W, H, spp, nb, 12345)
    target = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
    tscene.render(o, target)
    img = torch.empty_like(target)
    scene.render(dr.DeviceScene.options(W, H, spp, nb, 777), img)
    torch.cuda.synchronize()
    return (2.0 * (img - target)).contiguous()


def profiled_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 summaries (a `traffic_bytes_per_launch <kernel> <workload>: N`
    line in profiles/*.md, FETCH_SIZE + WRITE_SIZE scaled by a copy of known
    size); (None, None) when no summary covers them."""
    import glob
    import re
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(f) as fh:
            for line in fh:
                m = re.match(r"traffic_bytes_per_launch (\S+) (.+): (\d+)\s*$", line)
                if m and m.group(1) == kernel and m.group(2) == workload:
                    return int(m.group(3)), os.path.relpath(f, ROOT)
    return None, None


def cpu_leg(w, d_img, threads, seconds, adjoint=True):
    """The oracle (8x8 pixel tiles from a shared work list, as the
    reference's thread pool) on `threads` host threads: a 1-spp render of
    the whole image sizes the sample, then a render of about `seconds` of CPU
    work at spp_cpu <= the workload's spp.  Returns samples/s, rays/s and what
    the sample was."""
    from tests import oracle_lib
    osc = oracle_lib.scene_from_workload(w)
    W, H = w["resolution"]
    spp, nb = w["num_samples"], w["max_bounces"]
    dimg = d_img if adjoint else None
    t0 = time.perf_counter()
    osc.render(W, H, 1, nb, 3, d_image=dimg, nthreads=threads)
    t1 = time.perf_counter() - t0
    cspp = int(max(1, min(spp, seconds / max(t1, 1e-6))))
    t0 = time.perf_counter()
    osc.render(W, H, cspp, nb, 5, d_image=dimg, nthreads=threads)
    dt = time.perf_counter() - t0
    rays = osc.last_render_rays()
    return {"samples_per_s": cspp * W * H / dt, "rays_per_s": rays / dt, "spp": cspp, "seconds": dt,
            "rays": rays}


def bench_batched(args):
    """Uneven batch: scenes x 4 views, each view forward + adjoint, all in one
    graph-captured launch sequence per step (pass sizes from the autotuner)."""
    import numpy as np
    import torch
    from delta_ray_amd import delta_ray as dr
    from delta_ray_amd import workloads
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    scenes = workloads.batched(num_scenes=args.batch_scenes)
    gen_s = time.perf_counter() - t0
    print("batched: %d scenes generated in %.1f s" % (len(scenes), gen_s), file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    keep, jobs = [], []
    rng = np.random.default_rng(5)
    samples = 0
    for si, sc in enumerate(scenes):
        if si % 8 == 0:
            print("batched: building scene %d" % si, file=sys.stderr, flush=True)
        ds = dr.DeviceScene(sc["views"][0][0], sc["shapes"], sc["materials"], sc["lights"])
        grads = torch.zeros((ds.grad_size,), dtype=torch.float32, device=dev)
        keep.append(ds)
        for vi, (cam, (W, H), spp) in enumerate(sc["views"]):
            view = None if vi == 0 else dr.View(ds, cam)
            keep.append(view)
            d_img = torch.from_numpy(rng.normal(size=(H, W, 3)).astype(np.float32)).to(dev)
            img = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
            jobs.append(dict(scene=ds, view=view, options=dr.DeviceScene.options(W, H, spp, sc["max_bounces"], 0),
                             image=img, d_image=d_img, grads=grads))
            samples += W * H * spp
    build_s = time.perf_counter() - t0
    print("batched: %d jobs built in %.1f s" % (len(jobs), build_s), file=sys.stderr, flush=True)
    batch = dr.Batch()
    # the autotuner times 12 candidates: run it on every 8th job (a
    # representative eighth of the batch) and keep its choice for all jobs
    tune = None if args.no_autotune else batch.autotune(jobs[::8])
    print("batched: tuning %s" % (tune,), file=sys.stderr, flush=True)

    def step(i):
        for k, j in enumerate(jobs):
            j["options"].seed = 1000003 * i + k
        batch.render(jobs, graph=True)
        print("batched: step %d enqueued" % i, file=sys.stderr, flush=True)
    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(args.warmup + i)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    tris = [sc["num_triangles"] for sc in scenes]
    out = {"metric": METRIC, "value": round(samples / dt, 1), "unit": "samples/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1000.0, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded procedural scenes, random d_image)",
           "config": {"workload": "batched: %d scenes (%d-%d triangles, %d total) x %d views, 128-512 px, "
                                  "16-64 spp, 2 bounces, forward + adjoint per view, one hipGraph per step" %
                                  (len(scenes), min(tris), max(tris), sum(tris), len(scenes[0]["views"])),
                      "samples_per_step": samples, "jobs": len(jobs), "parallelism": "single GPU"},
           "extra": {"tuning": tune, "scene_generation_s": round(gen_s, 2), "scene_build_s": round(build_s, 2)}}
    print(json.dumps(out), flush=True)


def learned_leg(w, n, train_renders, train_steps, rounds, eps, terms=("secondary", "full"), log="",
                ablation=True, flags=0, blocks=1):
    """Time to a gradient relative RMSE of `eps`, learned edge proposal vs
    the plain one, on workload `w`.  Both estimators are unbiased, so for
    each one the per-render variance V of the gradient (sum over parameters)
    and the time per render t give the time to reach relative RMSE eps as
    t * V / (eps^2 |g|^2)  (n renders average to variance V / n).

    The gradient is the one an optimisation of the workload uses: its
    parameters' entries (w["params"]; on the heightfield its vertices and
    roughness texture, not the fixed light quads, whose vertex adjoints would
    otherwise dominate |g|), and the proposal's statistics measure
    contributions to those shapes.  Renders use `flags` (default 0: the exact
    derivative).  With the reference's quirks (DR_FLAGS_PARITY) the
    heightfield's vertex gradient is dominated by DR_QUIRK_EDGE_HIT_JACOBIAN,
    whose factor (x . n) / ((x - p) . n) is unbounded where a side ray hits
    the surface near the height of its origin: the estimate's variance is
    then not finite and no relative RMSE is reached at the 1/sqrt(n) rate.

    The estimator is still heavy-tailed (a rare sample through a tiny tree
    pmf carries a large weight), so V is the median over 16 groups of
    renders of the groups' sample variances, each around its own group's
    mean (median of means: what a run of typical length sees), with a
    bootstrap interval over the groups; |g| is the norm of the per-parameter
    median of all groups' means.  The plain sample variance over all renders
    and tail percentiles of the per-render squared error are reported
    beside it.  The learned time includes its training: `rounds` rounds of
    DR_EDGE_STATS renders (statistics of the plain estimator, then of the
    learned mixture), MLP training and applying the weights.  Measured on
    the secondary-edge term, the one the proposal changes, and (terms) on
    the full gradient (interior + secondary + primary edges).  `blocks`:
    each estimator renders that many disjoint seed blocks of n renders, cut
    into 16 groups each; V pools all blocks' groups, and each block's own
    variance ratios are reported beside it.  `ablation`:
    on the secondary-edge term, the same with weights from the per-edge
    statistics alone (no network), trained by the same rounds."""
    import numpy as np
    import torch
    from delta_ray_amd import _abi
    from delta_ray_amd import delta_ray as dr
    from delta_ray_amd.learned import EdgeProposal
    dev = torch.device("cuda:0")
    W, H = w["resolution"]
    spp, nb = w["num_samples"], w["max_bounces"]
    scene = dr.DeviceScene(w["camera"], w["shapes"], w["materials"], w["lights"])
    d_img = d_image_for(w, scene, dev)
    img = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
    g = torch.empty((scene.grad_size,), dtype=torch.float32, device=dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    sel = torch.from_numpy(param_index(w)).to(dev)

    def renders(fl, seeds):
        out = torch.empty((len(seeds), len(sel)), dtype=torch.float64, device=dev)
        ms = 0.0
        for k, sd in enumerate(seeds):
            o = dr.DeviceScene.options(W, H, spp, nb, sd, flags | fl)
            e0.record()
            scene.render_adjoint(o, d_img, img, g)
            e1.record()
            torch.cuda.synchronize()
            ms += e0.elapsed_time(e1)
            out[k] = g[sel].double()
            if (k + 1) % 32 == 0:
                print("learned%s: %d of %d renders" % (log, k + 1, len(seeds)), file=sys.stderr, flush=True)
        return out, ms / len(seeds)

    groups = 16 if n >= 64 else 8
    n = n - n % groups
    per_block = groups
    groups = per_block * blocks
    rng = np.random.default_rng(0)

    def group_stats(a):
        k = a.shape[0] // groups
        v = [float(((a[i * k:(i + 1) * k] - a[i * k:(i + 1) * k].mean(0)) ** 2).sum() / (k - 1))
             for i in range(groups)]
        means = torch.stack([a[i * k:(i + 1) * k].mean(0) for i in range(groups)])
        return np.array(v), means

    def full_var(a):
        return float(((a - a.mean(0)) ** 2).sum() / (a.shape[0] - 1))

    def tail(a):
        e = ((a - a.mean(0)) ** 2).sum(1).cpu().numpy()
        return [float("%.4g" % x) for x in np.percentile(e, [50, 90, 99, 100])]

    def ratio_ci(va, vb):
        r = [np.median(va[rng.integers(0, groups, groups)]) / np.median(vb[rng.integers(0, groups, groups)])
             for _ in range(2000)]
        return [round(float(x), 3) for x in np.percentile(r, [10, 90])]

    secondary = _abi.DR_SKIP_INTERIOR | _abi.DR_SKIP_PRIMARY_EDGES
    renders(secondary, [1])  # warm-up
    param_shapes = sorted({i for kind, i, _ in w["params"] if kind == "shape"})

    def train(use_network):
        """rounds of statistics + training + applying the weights, timed as
        a whole; round 0 gathers statistics of the plain estimator, later
        rounds of the learned mixture itself"""
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        prop = EdgeProposal(scene, seed=7, use_network=use_network, stat_shapes=param_shapes)
        loss, wts = 0.0, None
        for rnd in range(rounds):
            prop.reset_stats()
            fl = secondary | _abi.DR_EDGE_STATS | (_abi.DR_LEARNED_EDGES if rnd > 0 else 0)
            renders(fl, range(50 + 100 * rnd, 50 + 100 * rnd + train_renders))
            prop.collect()
            loss = prop.train(steps=train_steps, batch=1024)
            wts = prop.apply()
        torch.cuda.synchronize()
        return prop, wts, loss, time.perf_counter() - t0

    prop, wts, loss, train_s = train(True)
    print("learned%s: trained in %.2f s" % (log, train_s), file=sys.stderr, flush=True)
    out = {}
    for term in terms:
        fl = secondary if term == "secondary" else 0
        prop.set_weights(wts)
        gp, tp = renders(fl, range(1000, 1000 + blocks * n))
        gl, tl = renders(fl | _abi.DR_LEARNED_EDGES, range(100000, 100000 + blocks * n))
        print("learned%s: %s term measured" % (log, term), file=sys.stderr, flush=True)
        vpg, mp = group_stats(gp)
        vlg, ml = group_stats(gl)
        vp, vl = float(np.median(vpg)), float(np.median(vlg))
        g2 = max(float((torch.cat([mp, ml]).median(0).values ** 2).sum()), 1e-30)
        t_plain = tp / 1e3 * vp / (eps ** 2 * g2)
        t_learned = tl / 1e3 * vl / (eps ** 2 * g2) + train_s
        se = torch.sqrt(gp.var(0) / gp.shape[0] + gl.var(0) / gl.shape[0])
        diff = (gp.mean(0) - gl.mean(0)).abs()
        act = se > 0
        out[term] = {"time_to_eps_s_plain": round(t_plain, 3), "time_to_eps_s_learned": round(t_learned, 3),
                     "plain_over_learned_time": round(t_plain / t_learned, 4),
                     "variance_ratio_plain_over_learned": round(vp / vl, 4),
                     "variance_ratio_10_90": ratio_ci(vpg, vlg),
                     "ms_per_render_plain": round(tp, 3), "ms_per_render_learned": round(tl, 3),
                     "rel_rmse_per_render_plain": round(float(np.sqrt(vp / g2)), 5),
                     "rel_rmse_per_render_learned": round(float(np.sqrt(vl / g2)), 5),
                     "sample_variance_plain": float("%.4g" % full_var(gp)),
                     "sample_variance_learned": float("%.4g" % full_var(gl)),
                     "sq_error_p50_p90_p99_max_plain": tail(gp), "sq_error_p50_p90_p99_max_learned": tail(gl),
                     "mean_difference_over_se": round(float(diff.norm() / se.norm()), 3),
                     "params_beyond_4se": round(float((diff[act] / se[act] > 4.0).double().mean()), 5)}
        if blocks > 1:
            out[term]["variance_ratio_per_block"] = [
                round(float(np.median(vpg[b * per_block:(b + 1) * per_block]) /
                            np.median(vlg[b * per_block:(b + 1) * per_block])), 3) for b in range(blocks)]
        if term == "secondary" and ablation:
            # the per-edge statistics alone, no network: trained the same way
            _, _, _, ta_s = train(False)
            ga, ta = renders(fl | _abi.DR_LEARNED_EDGES, range(200000, 200000 + blocks * n))
            vag, _ = group_stats(ga)
            va = float(np.median(vag))
            out[term]["ablation_stats_only"] = {
                "time_to_eps_s": round(ta / 1e3 * va / (eps ** 2 * g2) + ta_s, 3),
                "training_s": round(ta_s, 3),
                "variance_ratio_plain_over_stats_only": round(vp / va, 4),
                "variance_ratio_stats_only_over_learned": round(va / vl, 4),
                "variance_ratio_stats_only_over_learned_10_90": ratio_ci(vag, vlg),
                "sample_variance_stats_only": float("%.4g" % full_var(ga)),
                "ms_per_render": round(ta, 3)}
            if blocks > 1:
                out[term]["ablation_stats_only"]["variance_ratio_plain_over_stats_only_per_block"] = [
                    round(float(np.median(vpg[b * per_block:(b + 1) * per_block]) /
                                np.median(vag[b * per_block:(b + 1) * per_block])), 3) for b in range(blocks)]
            print("learned%s: ablation measured" % log, file=sys.stderr, flush=True)
    return {"workload": "%s %dx%d %dspp %d-bounce, %d renders per estimator" % (w["name"], W, H, spp, nb,
                                                                                   blocks * n)
            + (" (%d seed blocks of %d)" % (blocks, n) if blocks > 1 else ""),
            "eps": eps, "flags": "exact derivative" if flags == 0 else "0x%x" % flags,
            "gradient": "parameters " + ", ".join("%s %d %s" % p for p in w["params"]),
            "training_s": round(train_s, 3),
            "training": {"rounds": rounds, "stats_renders_per_round": train_renders,
                         "mlp_steps_per_round": train_steps, "final_loss": round(float(loss), 5),
                         "weight_p5_p50_p95": [round(float(x), 4) for x in np.percentile(wts, [5, 50, 95])]},
            **out}


def bench_learned(args):
    """`--workload learned`: learned_leg on the heightfield at --lv-res,
    --lv-spp with --lv-renders renders per estimator, as its own JSON line."""
    from delta_ray_amd import workloads
    w = workloads.heightfield(resolution=(args.lv_res, args.lv_res), num_samples=args.lv_spp)
    r = learned_leg(w, args.lv_renders, args.lv_train_renders, args.lv_train_steps, args.lv_rounds, args.lv_eps,
                    blocks=args.lv_blocks)
    sec = r["secondary"]
    res_line = {"metric": "time to gradient rel-RMSE 1e-2 (secondary-edge term), learned proposal",
                "value": sec["time_to_eps_s_learned"], "unit": "s", "n_gpus": 1, "steps": args.lv_renders,
                "warmup": 1, "ms_per_step": sec["ms_per_render_learned"], "higher_is_better": False,
                "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural scene)",
                "config": {"workload": r["workload"], "parallelism": "single GPU"}, "extra": r}
    print(json.dumps(res_line), flush=True)


def bench_torch_step(args):
    """`--workload torch_step`: optimisation steps on the heightfield (512 x
    512, 64 spp, 2 bounces) through the torch front end with the reference's
    calling convention (RenderFunction.apply on serialize_scene's argument
    list, loss.backward(), Adam on the heightfield's vertices, as the
    reference's test_heightfield.py / test_single_triangle.py loops), the
    vertices a CUDA tensor.  Reports ms per step, the render kernels' launches
    per step (one backward = one adjoint render; its edge selection runs once
    per pass), and how often the scene was built, refitted or rebuilt."""
    import numpy as np
    import torch
    from delta_ray_amd import _abi
    from delta_ray_amd import scene as sc
    from delta_ray_amd import workloads
    from delta_ray_amd.render_pytorch import RenderFunction, _SceneCache
    dev = torch.device("cuda:0")
    w = workloads.heightfield()
    W, H = w["resolution"]
    spp, nb = w["num_samples"], w["max_bounces"]
    cam = sc.Camera(np.array([0, 3, -6], np.float32), np.array([0, 0, 0], np.float32),
                    np.array([0, 1, 0], np.float32), None, 45.0, 0.01, 10000.0, (W, H))
    mats = []
    for m in w["materials"]:
        mats.append(sc.Material(torch.from_numpy(np.asarray(m.diffuse_reflectance)),
                                torch.from_numpy(np.asarray(m.specular_reflectance)),
                                torch.from_numpy(np.asarray(m.roughness)), two_sided=bool(m.two_sided)))
    shapes = []
    for s in w["shapes"]:
        mid = w["materials"].index(s.material)
        v = torch.from_numpy(np.asarray(s.vertices)).to(dev)
        shapes.append(sc.Shape(v, torch.from_numpy(np.asarray(s.indices)).to(dev),
                               None if s.uvs is None else torch.from_numpy(np.asarray(s.uvs)).to(dev),
                               None if s.normals is None else torch.from_numpy(np.asarray(s.normals)).to(dev), mid))
    lights = [sc.Light(w["shapes"].index(l.shape), torch.from_numpy(np.asarray(l.intensity))) for l in w["lights"]]
    render = RenderFunction.apply
    args0 = RenderFunction.serialize_scene(cam, mats, shapes, lights, (W, H), spp, nb)
    target = render(1, *args0).detach()
    rng = np.random.default_rng(7)
    v0 = shapes[0].vertices
    shapes[0].vertices = (v0 + torch.from_numpy(rng.normal(0.0, 0.02, size=tuple(v0.shape)).astype(np.float32))
                          .to(dev)).requires_grad_(True)
    opt = torch.optim.Adam([shapes[0].vertices], lr=1e-3)
    losses = []

    def step(i):
        opt.zero_grad()
        a = RenderFunction.serialize_scene(cam, mats, shapes, lights, (W, H), spp, nb)
        img = render(100 + i, *a)
        loss = (img - target).pow(2).sum()
        loss.backward()
        opt.step()
        losses.append(float(loss.item()))
    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    scene = _SceneCache.scene
    info0 = scene.info()
    scene.stats(_abi.DR_STATS_TIMING)
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(args.warmup + i)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    kt = scene.stats(0)
    info1 = scene.info()
    out = {"metric": "optimisation steps/s (RenderFunction forward + backward + Adam, vertices on the GPU)",
           "value": round(1.0 / dt, 4), "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(dt * 1000.0, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic (seeded procedural scene, target rendered from the start vertices)",
           "config": {"workload": "heightfield %dx%d %dspp %d-bounce, %d triangles, Adam on %d vertex coordinates" %
                                  (W, H, spp, nb, info1["num_triangles"], v0.numel()),
                      "parallelism": "single GPU"},
           "extra": {"launches_per_step": {k: kt[k + "_launches"] // args.steps
                                           for k in ("path", "select", "edge", "dp", "primary", "forward")},
                     "kernel_ms_per_step": {k: round(kt[k + "_ms"] / args.steps, 3)
                                            for k in ("path", "select", "edge", "dp", "primary", "forward")},
                     "scene_builds_during_steps": 0 if _SceneCache.scene is scene else 1,
                     "refits_during_steps": info1["refits"] - info0["refits"],
                     "rebuilds_during_steps": info1["rebuilds"] - info0["rebuilds"],
                     "losses": [round(x, 3) for x in losses]}}
    print(json.dumps(out), flush=True)


def gpu_leg(w, steps, warmup, calibrate=False):
    """Forward + adjoint steps of workload `w` (the timed figure), forward
    steps, per-kernel counters and the roofline of the dominant kernel."""
    import torch
    from delta_ray_amd import _abi
    from delta_ray_amd import delta_ray as dr
    dev = torch.device("cuda:0")
    W, H = w["resolution"]
    spp, nb = w["num_samples"], w["max_bounces"]
    t0 = time.perf_counter()
    scene = dr.DeviceScene(w["camera"], w["shapes"], w["materials"], w["lights"])
    build_s = time.perf_counter() - t0
    info = scene.info()
    d_img = d_image_for(w, scene, dev)
    image = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
    grads = torch.empty((scene.grad_size,), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()

    def step(i):
        o = dr.DeviceScene.options(W, H, spp, nb, 1000 + i)
        scene.render_adjoint(o, d_img, image, grads, None, stream)

    for i in range(warmup):
        step(i)
    torch.cuda.synchronize()
    scene.stats(_abi.DR_STATS_TIMING)  # clear, start kernel timing
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        step(warmup + i)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kt = scene.stats(0)
    ms_per_step = (t1 - t0) * 1000.0 / steps
    samples = spp * W * H
    value = samples / (ms_per_step / 1000.0)

    # forward only (the reference's render() with a 0-d d_image), same
    # workload, after the timed adjoint steps
    fo = dr.DeviceScene.options(W, H, spp, nb, 4242)
    scene.render(fo, image, stream)
    torch.cuda.synchronize()
    tf0 = time.perf_counter()
    for i in range(steps):
        fo.seed = 5000 + i
        scene.render(fo, image, stream)
    torch.cuda.synchronize()
    fwd_ms = (time.perf_counter() - tf0) * 1000.0 / steps
    scene.stats(_abi.DR_STATS_COUNTERS)
    scene.render(fo, image, stream)
    cf = scene.stats(0)

    # per-kernel counters from untimed runs: path only, path + secondary
    # edges, path + primary edges; differences attribute rays to kernels
    def counters(extra_flags):
        scene.stats(_abi.DR_STATS_COUNTERS)
        o = dr.DeviceScene.options(W, H, spp, nb, 999, _abi.DR_FLAGS_PARITY | extra_flags)
        scene.render_adjoint(o, d_img, image, grads, None, stream)
        return scene.stats(0)
    ca = counters(_abi.DR_SKIP_SECONDARY_EDGES | _abi.DR_SKIP_PRIMARY_EDGES)
    cb = counters(_abi.DR_SKIP_PRIMARY_EDGES)
    cc = counters(_abi.DR_SKIP_SECONDARY_EDGES)

    def diff(x, y):
        return {k: x[k] - y[k] for k in ("rays", "nodes_visited", "tris_visited")}
    per_kernel = {"path": dict(ca), "edge": diff(cb, ca), "primary": diff(cc, ca)}
    per_kernel["edge"]["edges"] = cb["secondary_edges"]
    per_kernel["select"] = {"rays": 0, "nodes_visited": 0, "tris_visited": 0,
                            "edge_nodes_visited": cb["edge_nodes_visited"], "edges_tested": cb["edges_tested"]}
    bytes_k = {}
    for k, c in per_kernel.items():
        bytes_k[k] = B_RAY_IO * c["rays"] + B_NODE * c["nodes_visited"] + B_TRI * c["tris_visited"]
    bytes_k["path"] += B_PIXEL_IO * samples
    bytes_k["edge"] += (B_EDGE + B_POINT) * cb["secondary_edges"]
    n_points = samples * nb
    bytes_k["select"] += (B_POINT * n_points + B_TREE_EXPAND * cb["edge_nodes_visited"] +
                          B_EDGE_TEST * cb["edges_tested"] + B_QUEUE * cb["secondary_edges"])
    # the edge samples run in two launches per pass (k_edge, then k_edge_shade
    # on the samples with continuation paths): their work counters are one
    # figure, so their times are taken together as "edge"
    kt["edge_shade_ms"] = kt["shade_ms"]
    kt["edge_ms"] += kt["shade_ms"]
    kms = {k: kt[k + "_ms"] / max(kt[k + "_launches"], 1) for k in ("path", "select", "edge", "dp", "primary")}
    kms_step = {k: kt[k + "_ms"] / steps for k in ("path", "select", "edge", "dp", "primary", "edge_shade")}
    dom = max(("path", "select", "edge", "primary"), key=lambda k: kms_step[k])
    launches_per_step = max(kt[dom + "_launches"] // steps, 1)
    bytes_launch = bytes_k[dom] / launches_per_step
    achieved = bytes_launch / (kms[dom] / 1000.0) / 1e9
    rays_total = cb["rays"] + per_kernel["primary"]["rays"]
    wl_key = "%s %dx%d %dspp %d-bounce" % (w["name"], W, H, spp, nb)
    traffic, traffic_src = profiled_traffic("k_" + dom, wl_key)

    def per_ray(c):
        return {"nodes": round(c["nodes_visited"] / max(c["rays"], 1), 2),
                "tris": round(c["tris_visited"] / max(c["rays"], 1), 2), "rays": int(c["rays"])}
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "traffic_source": traffic_src,
                "kernel": "k_" + dom, "kernel_ms_per_launch": round(kms[dom], 3),
                "bytes_per_launch": int(bytes_launch),
                "per_ray": {k: per_ray(per_kernel[k]) for k in ("path", "edge", "primary")},
                "per_kernel_gbs": {k: round(bytes_k[k] / max(kt[k + "_launches"] // steps, 1)
                                            / max(kms[k] / 1000.0, 1e-12) / 1e9, 2) for k in bytes_k},
                "per_kernel_frac": {k: round(bytes_k[k] / max(kt[k + "_launches"] // steps, 1)
                                             / max(kms[k] / 1000.0, 1e-12) / 1e9 / HBM_PEAK_GBS, 5)
                                    for k in bytes_k}}
    r = {"value": value, "ms_per_step": ms_per_step, "samples": samples, "roofline": roofline,
         "workload": "%s %dx%d %dspp %d-bounce, %d triangles" % (w["name"], W, H, spp, nb, info["num_triangles"]),
         "extra": {"forward_samples_per_s": round(samples / (fwd_ms / 1000.0), 1),
                   "forward_ms_per_step": round(fwd_ms, 3),
                   "forward_rays_per_s": round(cf["rays"] / (fwd_ms / 1000.0), 1),
                   "forward_per_ray": per_ray(cf),
                   "rays_per_s": round(rays_total / (ms_per_step / 1000.0), 1),
                   "rays_per_step": int(rays_total),
                   "secondary_edges_per_step": int(cb["secondary_edges"]),
                   "edge_tree_nodes_per_point": round(cb["edge_nodes_visited"] / max(n_points, 1), 2),
                   "edges_tested_per_point": round(cb["edges_tested"] / max(n_points, 1), 2),
                   "stack_overflows": int(cb.get("stack_overflows", 0)),
                   "kernel_ms_per_step": {k: round(v, 3) for k, v in kms_step.items()},
                   "kernel_launches_per_step": {k: kt[k + "_launches"] // steps
                                                for k in ("path", "select", "edge", "dp", "primary")},
                   "scene_build_s": round(build_s, 3), "scene": info},
         "d_img": d_img}
    if calibrate:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        r["extra"]["calibration_copy_bytes"] = 2 * src.numel() * 4
    return r


def bench_table(args):
    """The three single-scene configurations (single triangle 256x256 4 spp,
    shadow 256x256 16 spp, heightfield 512x512 64 spp), forward and forward +
    adjoint, on the GPU and on the CPU oracle (bounded samples), as JSON."""
    from delta_ray_amd import workloads
    threads, host = host_cpus()
    rows = []
    for name in ("single_triangle", "shadow", "heightfield"):
        w = workloads.WORKLOADS[name]()
        W, H = w["resolution"]
        steps = args.steps if name == "heightfield" else max(args.steps, 20)
        g = gpu_leg(w, steps, args.warmup)
        print("table: %s GPU done" % name, file=sys.stderr, flush=True)
        d = g["d_img"].cpu().numpy()
        cf = cpu_leg(w, d, threads, args.cpu_seconds / 2, adjoint=False)
        ca = cpu_leg(w, d, threads, args.cpu_seconds / 2, adjoint=True)
        print("table: %s CPU done" % name, file=sys.stderr, flush=True)
        e = g["extra"]
        rows.append({"config": name, "workload": g["workload"], "spp": w["num_samples"], "W": W, "H": H,
                     "bounces": w["max_bounces"],
                     "fwd": {"gpu_samples_per_s": e["forward_samples_per_s"], "gpu_rays_per_s": e["forward_rays_per_s"],
                             "cpu_samples_per_s": round(cf["samples_per_s"], 1),
                             "cpu_rays_per_s": round(cf["rays_per_s"], 1),
                             "cpu_sample": "%d spp, %.1f s" % (cf["spp"], cf["seconds"])},
                     "adj": {"gpu_samples_per_s": round(g["value"], 1), "gpu_rays_per_s": e["rays_per_s"],
                             "cpu_samples_per_s": round(ca["samples_per_s"], 1),
                             "cpu_rays_per_s": round(ca["rays_per_s"], 1),
                             "cpu_sample": "%d spp, %.1f s" % (ca["spp"], ca["seconds"]),
                             "roofline": {k: g["roofline"][k] for k in ("kernel", "frac", "traffic", "achieved")}}})
    out = {"cpu_threads": threads, "host": host, "rows": rows}
    if not args.no_learned:
        w = workloads.heightfield()
        out["learned_vs_plain"] = learned_leg(w, args.lv_bench_renders, args.lv_train_renders, args.lv_train_steps,
                                              args.lv_rounds, args.lv_eps, terms=("secondary",), log=" (C3)",
                                              blocks=args.lv_blocks)
    with open(args.table, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out), flush=True)


def fill_baseline(path, base):
    """Rewrite the results table of the markdown file `base` (the table
    after its "### Results table" heading) from a --table JSON."""
    with open(path) as f:
        t = json.load(f)
    cores = "%d (%s; nproc %s)" % (t["cpu_threads"], t["host"].get("model") or "?", t["host"].get("nproc"))
    label = {"single_triangle": "C1 256²×4spp, B=1", "shadow": "C2 256²×16spp, B=1",
             "heightfield": "C3 512²×64spp, B=2"}
    lines = ["| config | mode | CPU cores | CPU samples/s | CPU rays/s | GPU samples/s | GPU rays/s | GPU ÷ CPU | "
             "roofline fraction (HBM algorithmic / HBM measured / MFMA) |",
             "|---|---|---|---|---|---|---|---|---|"]

    def f(x):
        return "%.4g" % x
    for r in t["rows"]:
        for mode in ("fwd", "adj"):
            m = r[mode]
            roof = "—"
            if mode == "adj":
                rf = m["roofline"]
                roof = "%s: %.3f / %s / —" % (rf["kernel"], rf["frac"],
                                              "%.2f GB per launch" % (rf["traffic"] / 1e9) if rf["traffic"] else "—")
            lines.append("| %s | %s | %s | %s (%s) | %s | %s | %s | %.0f× | %s |" % (
                label[r["config"]] if mode == "fwd" else label[r["config"]].split()[0],
                "fwd" if mode == "fwd" else "fwd+adj", cores, f(m["cpu_samples_per_s"]), m["cpu_sample"],
                f(m["cpu_rays_per_s"]), f(m["gpu_samples_per_s"]), f(m["gpu_rays_per_s"]),
                m["gpu_samples_per_s"] / m["cpu_samples_per_s"], roof))
    lv = t.get("learned_vs_plain")
    if lv:
        s = lv["secondary"]
        lines.append("| C3 | time to rel-RMSE ≤1e-2 (secondary-edge term), plain vs learned | — | — | — | "
                     "%.4g s vs %.4g s (%s) | — | %.2f× (target ≥3×) | — |" % (
                         s["time_to_eps_s_plain"], s["time_to_eps_s_learned"], lv["workload"],
                         s["plain_over_learned_time"]))
    with open(base) as fh:
        text = fh.read()
    head = "### Results table (filled by the build's measurement harness, never by hand)"
    a = text.index(head)
    b = text.index("| config | mode |", a)
    e = b
    while e < len(text):
        nl = text.find("\n", e)
        if nl < 0:
            nl = len(text)
        if not text[e:nl].startswith("|"):
            break
        e = nl + 1
    note = ("Written by `python bench.py --fill-baseline <json>` from `python bench.py --table <json>` (%s).  "
            "CPU = the oracle (8x8 tiles from a shared work list) on the %d threads of the host's CPU share; "
            "GPU = one MI355X.  The CPU figures are bounded samples (spp in brackets) of the same scene and "
            "resolution.\n\n" % (os.path.basename(path), t["cpu_threads"]))
    text = text[:a] + head + "\n\n" + note + "\n".join(lines) + "\n" + text[e:]
    with open(base, "w") as fh:
        fh.write(text)


def main():
    args = parse()
    if args.fill_baseline:
        fill_baseline(*args.fill_baseline)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: multi-GPU runs are not part of this build "
              "(the path would shard pixel samples per GPU and all-reduce the gradient buffer)")
        sys.exit(2)
    if args.table:
        bench_table(args)
        return
    if args.workload == "batched":
        bench_batched(args)
        return
    if args.workload == "learned":
        bench_learned(args)
        return
    if args.workload == "torch_step":
        bench_torch_step(args)
        return
    if args.scene:
        from delta_ray_amd import workloads
        w = workloads.mitsuba(args.scene, num_samples=args.scene_spp, max_bounces=args.scene_bounces)
    else:
        w = make_workload(args.workload)
    g = gpu_leg(w, args.steps, args.warmup, args.calibrate)
    value = g["value"]
    out = {"metric": METRIC, "value": round(value, 1), "unit": "samples/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(g["ms_per_step"], 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded procedural scene)",
           "config": {"workload": g["workload"], "samples_per_step": g["samples"],
                      "primary_edge_samples_per_step": g["samples"], "parallelism": "single GPU"},
           "roofline": g["roofline"], "extra": g["extra"]}
    if not args.no_cpu:
        threads, host = host_cpus()
        c = cpu_leg(w, g["d_img"].cpu().numpy(), threads, args.cpu_seconds)
        W, H = w["resolution"]
        out["cpu_baseline"] = {"value": round(c["samples_per_s"], 1), "unit": "samples/s", "cores": threads,
                               "kind": "port", "rays_per_s": round(c["rays_per_s"], 1), "host": host,
                               "sample": "same scene and resolution at %d spp (%d pixel samples + as many "
                                         "primary-edge samples), fwd+adjoint, %.1f s, 8x8 pixel tiles from a "
                                         "shared work list" % (c["spp"], c["spp"] * W * H, c["seconds"])}
        out["extra"]["gpu_over_cpu"] = round(value / c["samples_per_s"], 1)
    if not args.no_learned and args.workload == "heightfield" and not args.scene:
        print("bench: learned-vs-plain leg", file=sys.stderr, flush=True)
        out["extra"]["learned_vs_plain"] = learned_leg(w, args.lv_bench_renders, args.lv_train_renders,
                                                       args.lv_train_steps, args.lv_rounds, args.lv_eps,
                                                       terms=("secondary",), blocks=args.lv_blocks)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

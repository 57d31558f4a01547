"""Learned secondary-edge proposal for a DeviceScene (no reference counterpart).

The edge tree chooses silhouette edges with node importances built from edge
lengths and a BSDF bound (edge_tree.cpp:499-509).  Here a small network
(dr_mlp_*: 24 per-edge features -> 64 -> 64 -> 1, bf16 weights on the matrix
cores) predicts a per-edge weight; node lengths become sums of weight x edge
length (dr_scene_set_edge_weights), and DR_LEARNED_EDGES renders mix that
proposal with the plain one per shading point (probabilities 0.8 / 0.2): the
two tree passes of the chosen proposal are weighted with the reference's
power heuristic, and the proposals with the balance heuristic (a defensive
mixture; a proposal also picks between the silhouette and the normal tree by
its own root lengths), so the estimator stays unbiased whatever the network predicts, the
plain proposal remains part of every estimate, and uninformative weights
give the plain estimator back.

Training data come from the plain estimator itself: DR_EDGE_STATS renders
accumulate, per edge, the number of secondary-edge samples and the sum and
sum of squares of their |gradient contribution| (already divided by the
selection probability): the edge's vertex adjoints and the Dirac derivative
at the shading point, counted for the shapes being optimised.  For a proposal whose per-edge scale factors w_e
multiply the plain one, the second moment of the estimate is (to first
order, when the points' totals vary little) smallest for w_e proportional to
the root mean square of the edge's weighted contributions (sqrt(sum f^2/p / sum p), over the points that select it); the
log of that RMS (centred, with a floor) is the network's regression target,
with a loss weight growing with the sample count.  The applied weights
combine the network's prediction with the edge's own statistic, in
proportion to its sample count (the network is the prior of edges with few
or no samples), and are normalised by their geometric mean.
"""
import ctypes

import numpy as np

from . import _abi


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class EdgeProposal:
    """use_network=False: the ablation without the network -- the applied
    weights come from each edge's own statistics alone (edges without samples
    keep the plain proposal's weight).  stat_shapes: indices of the shapes
    whose gradients the proposal should serve (the optimised ones; None: all),
    which decides what the statistics measure (dr_scene_set_stat_shapes)."""

    def __init__(self, scene, view=None, seed=0, use_network=True, stat_shapes=None):
        import torch
        self.torch = torch
        self.lib = scene.lib
        self.scene = scene
        self.n = int(self.lib.dr_scene_edge_count(scene.handle))
        dev = torch.device("cuda")
        self.features = torch.zeros((max(self.n, 1), _abi.DR_MLP_FEATURES), dtype=torch.float32, device=dev)
        vh = view.handle if view is not None else None
        _abi.check(self.lib, self.lib.dr_scene_edge_features(scene.handle, vh, _ptr(self.features), None),
                   "dr_scene_edge_features")
        ids = np.ascontiguousarray(stat_shapes if stat_shapes is not None else [], np.int32)
        _abi.check(self.lib, self.lib.dr_scene_set_stat_shapes(scene.handle, ids.ctypes.data if len(ids) else None,
                                                               len(ids)), "dr_scene_set_stat_shapes")
        self.stats = torch.zeros((max(self.n, 1), 3), dtype=torch.float32, device=dev)
        self.acc = torch.zeros((max(self.n, 1), 3), dtype=torch.float64, device=dev)
        h = ctypes.c_void_p()
        _abi.check(self.lib, self.lib.dr_mlp_create(seed, ctypes.byref(h)), "dr_mlp_create")
        self.mlp = h
        self.logits = torch.zeros(max(self.n, 1), dtype=torch.float32, device=dev)
        self.seed = seed
        self.use_network = use_network
        self.trained_steps = 0
        self.log_w = np.zeros(self.n)  # current log-weights (all 1: the plain proposal)

    def __del__(self):
        h = getattr(self, "mlp", None)
        if h is not None and h.value:
            self.lib.dr_mlp_destroy(h)
            self.mlp = None

    def reset_stats(self):
        """Start a statistics window: clear the scene's counters and the
        statistics gathered so far."""
        _abi.check(self.lib, self.lib.dr_scene_edge_stats(self.scene.handle, None, 1, None), "dr_scene_edge_stats")
        self.acc.zero_()

    def collect(self):
        """Add the statistics of the DR_EDGE_STATS renders since the last
        call to the training set (and clear them in the scene)."""
        _abi.check(self.lib, self.lib.dr_scene_edge_stats(self.scene.handle, _ptr(self.stats), 1, None),
                   "dr_scene_edge_stats")
        self.torch.cuda.synchronize()
        self.acc += self.stats.double()

    def targets(self, full_weight_at=64.0, floor=1e-3):
        """Per-edge regression targets (log RMS of the weighted contributions,
        centred) and loss weights.  The RMS gets `floor` times the RMS over
        all samples added, so edges that were selected but never contributed
        get a finite target (about log(floor) below the typical edge) instead
        of dominating the loss."""
        t = self.torch
        cnt = self.acc[:, 1]
        seen = cnt > 0
        rms = t.sqrt(self.acc[:, 2] / t.clamp(cnt, min=1.0))
        typical = t.sqrt(self.acc[:, 2].sum() / t.clamp(cnt.sum(), min=1.0)).clamp(min=1e-30)
        y = t.log(rms + floor * typical)
        w = t.clamp(cnt / full_weight_at, max=1.0)
        w = t.where(seen, w, t.zeros_like(w))
        if bool(seen.any()):
            centre = (y * w).sum() / w.sum()
            y = t.where(seen, y - centre, t.zeros_like(y))
        return y.float().contiguous(), w.float().contiguous()

    def train(self, steps=300, batch=1024, lr=3e-3):
        if not self.use_network:
            return 0.0
        y, w = self.targets()
        loss = ctypes.c_float(0.0)
        _abi.check(self.lib, self.lib.dr_mlp_train(self.mlp, _ptr(self.features), _ptr(y), _ptr(w), self.n, steps,
                                                   batch, lr, self.seed + self.trained_steps, ctypes.byref(loss),
                                                   None), "dr_mlp_train")
        self.trained_steps += steps
        return loss.value

    def apply(self, clip=(1e-3, 1e3), prior_count=16.0, temper=0.5):
        """Network + statistics -> the scene's learned node lengths.  Per edge
        the step is (n * y + k * prediction) / (n + k): its own target y from
        n samples, shrunk toward the network's prediction with k =
        `prior_count` pseudo-samples; centred (weights normalised by their
        geometric mean) and multiplied by `temper` (the statistic is noisy and
        the best scale factor differs from point to point, so the step toward
        it is damped).  The step is added to the current log-weights, so that
        rounds of statistics gathered under the learned proposal itself
        (DR_LEARNED_EDGES | DR_EDGE_STATS renders) refine the weights: each
        round's RMS measures what the current proposal still under- or
        over-samples.  Weights are clipped to `clip`, a wide range: the
        second moment of the secondary-edge estimate can concentrate on few
        edges, whose best weights are then hundreds of times the typical
        one."""
        if self.use_network:
            _abi.check(self.lib, self.lib.dr_mlp_infer(self.mlp, _ptr(self.features), self.n, _ptr(self.logits),
                                                       None), "dr_mlp_infer")
            self.torch.cuda.synchronize()
            pred = self.logits[:self.n].double().cpu().numpy()
        else:
            pred = np.zeros(self.n)
        y, _ = self.targets()
        y = y[:self.n].double().cpu().numpy()
        cnt = self.acc[:self.n, 1].cpu().numpy()
        lg = (cnt * y + prior_count * pred) / (cnt + prior_count)
        self.log_w = np.clip(self.log_w + temper * (lg - lg.mean()), np.log(clip[0]), np.log(clip[1]))
        wts = np.exp(self.log_w).astype(np.float32)
        _abi.check(self.lib, self.lib.dr_scene_set_edge_weights(self.scene.handle, wts.ctypes.data),
                   "dr_scene_set_edge_weights")
        return wts

    def set_weights(self, wts):
        w = np.ascontiguousarray(wts, np.float32)
        self.log_w = np.log(np.maximum(w.astype(np.float64), 1e-30))
        _abi.check(self.lib, self.lib.dr_scene_set_edge_weights(self.scene.handle, w.ctypes.data),
                   "dr_scene_set_edge_weights")


def main():
    """Time the proposal's network on the edges of the glossy heightfield
    (128 x 128 grid): `--reps` inferences of every edge (dr_mlp_infer, bf16
    on the matrix cores) and one training call of `--steps` Adam steps.

        python -m delta_ray_amd.learned --reps 200 --steps 300
    """
    import argparse
    import json
    import time
    import torch
    from . import delta_ray as dr
    from . import workloads
    ap = argparse.ArgumentParser(description=main.__doc__.split("\n\n")[0])
    ap.add_argument("--reps", type=int, default=200)
    ap.add_argument("--steps", type=int, default=300)
    a = ap.parse_args()
    torch.cuda.init()
    w = workloads.heightfield()
    scene = dr.DeviceScene(w["camera"], w["shapes"], w["materials"], w["lights"])
    prop = EdgeProposal(scene, seed=1)
    y = torch.randn(prop.n, device="cuda")
    wt = torch.ones(prop.n, device="cuda")
    loss = ctypes.c_float(0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _abi.check(prop.lib, prop.lib.dr_mlp_train(prop.mlp, _ptr(prop.features), _ptr(y), _ptr(wt), prop.n, a.steps,
                                               1024, 3e-3, 1, ctypes.byref(loss), None), "dr_mlp_train")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(a.reps):
        _abi.check(prop.lib, prop.lib.dr_mlp_infer(prop.mlp, _ptr(prop.features), prop.n, _ptr(prop.logits), None),
                   "dr_mlp_infer")
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    flops = 2.0 * (_abi.DR_MLP_FEATURES * 64 + 64 * 64 + 64) * prop.n
    ms = (t2 - t1) / a.reps * 1e3
    print(json.dumps({"edges": prop.n, "infer_ms": round(ms, 4), "infer_tflops": round(flops / (ms * 1e-3) / 1e12, 4),
                      "train_ms": round((t1 - t0) * 1e3, 2), "train_steps": a.steps, "loss": loss.value}))


if __name__ == "__main__":
    main()

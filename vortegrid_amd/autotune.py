"""Autotuner of the batched-effects path.

Chooses, on the live batch, the near-field walk's traversal parameters (near
radius near_k, first layer applied through brick-centred local expansions)
as the fastest candidate whose velocity grids meet an accuracy target
against the FP64 direct sum, then the launch shapes (vg_batch_autotune:
walk waves per workgroup, vorton-advection workgroup and items per thread).
The choice and every trial are returned as a dict and can be written as
JSON; load_choice() applies a saved choice to another batch.

Accuracy is judged on a fixed sample of effects (spread over the vorton
counts) and of grid points per effect: rel-L2 = sqrt(sum |u - u64|^2 /
sum |u64|^2) over the sample, u64 from vg_direct_velocity_f64 with each
effect's own mollification.  Timing is the wall time of vg_batch_solve
(bounds, geometry, trees and walk; the advection passes do not depend on
the method) with the stream synchronised around it.
"""
from __future__ import annotations

import json
import time

import numpy as np

from . import vg

F32 = np.float32
CANDIDATES = ((1, -1), (1, 2), (2, -1), (2, 3), (2, 2), (3, 3), (3, 2), (4, 2))


def _sample_points(g, idx):
    nudge = F32(1.0) - F32(2.0) * F32(np.finfo(F32).eps)
    ix = idx % g.np[0]
    iy = (idx // g.np[0]) % g.np[1]
    iz = idx // (g.np[0] * g.np[1])
    return np.stack([F32(g.min[a]) + np.asarray(i, F32) * (F32(g.cell[a]) * nudge)
                     for a, i in enumerate((ix, iy, iz))], 1).astype(F32)


class _Judge:
    """FP64 references for a fixed sample of effects and points, taken once
    (the grids' geometry does not depend on the method)."""

    def __init__(self, batch, vom_host, radii, n_effects, n_points, seed, stream):
        counts = np.asarray(batch.vorton_counts, np.int64)
        order = np.argsort(counts, kind="stable")
        nonempty = order[counts[order] > 0]
        pick = nonempty[np.linspace(0, len(nonempty) - 1, min(n_effects, len(nonempty))).astype(int)]
        off = np.concatenate([[0], np.cumsum(counts)])
        vpos = batch.download("vpos")
        rng = np.random.default_rng(seed)
        batch.solve(stream)
        vg.synchronize(stream)
        self.items = []
        for e in sorted(set(int(x) for x in pick)):
            g, _ = batch.grid(e)
            spread, circ = batch.mollify(e)
            idx = np.sort(rng.choice(g.capacity, min(n_points, g.capacity), replace=False))
            q = _sample_points(g, idx)
            src = vpos[off[e]:off[e + 1]]
            dp, dw = vg.DeviceArray.from_numpy(src), vg.DeviceArray.from_numpy(vom_host[off[e]:off[e + 1]])
            v = vg.vortons_view(dp, dw, len(src), uniform_size=2 * radii[e])
            dq = vg.DeviceArray.from_numpy(q)
            ref = vg.DeviceArray.zeros((len(q), 3), np.float64)
            vg.direct_velocity_f64(v, dq, len(q), ref, spread, circ, stream=stream)
            vg.synchronize(stream)
            self.items.append((e, idx, ref.numpy()))

    def rel_l2(self, batch):
        num = den = 0.0
        for e, idx, ref in self.items:
            _, field = batch.grid(e)
            d = field[idx].astype(np.float64) - ref
            num += float(np.sum(d * d))
            den += float(np.sum(ref * ref))
        return (num / den) ** 0.5 if den > 0 else 0.0


def tune_batch(batch, vom_host, radii, target=2.5e-2, candidates=CANDIDATES, order=1,
               sample_effects=16, sample_points=512, reps=3, seed=0, stream=None, path=None):
    """Pick (near_k, expand_from) and the launch shapes for `batch` (uploaded
    and stepped at least once).  vom_host: all effects' angular velocities
    (the batch's own order), radii: vorton radius per effect.  Leaves the
    batch set to the choice and returns the record (also written to path)."""
    if stream is None:
        stream = vg.Stream()
    judge = _Judge(batch, np.ascontiguousarray(vom_host, F32).reshape(-1, 3), radii,
                   sample_effects, sample_points, seed, stream)
    trials = []
    for near_k, expand_from in candidates:
        batch.set_method(near_k, order, expand_from)
        batch.solve(stream)  # warm
        vg.synchronize(stream)
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            batch.solve(stream)
            vg.synchronize(stream)
            times.append((time.perf_counter() - t0) * 1e3)
        err = judge.rel_l2(batch)
        trials.append({"near_k": near_k, "expand_from": expand_from, "order": order,
                       "solve_ms": min(times), "rel_l2": err, "meets_target": err <= target})
    ok = [t for t in trials if t["meets_target"]]
    best = min(ok, key=lambda t: t["solve_ms"]) if ok else min(trials, key=lambda t: t["rel_l2"])
    batch.set_method(best["near_k"], order, best["expand_from"])
    batch.solve(stream)
    vg.synchronize(stream)
    shapes = batch.autotune(stream)
    walk_waves, adv_block, adv_items = batch.get_launch()
    rec = {"target_rel_l2": target, "method": {"near_k": best["near_k"], "order": order,
                                               "expand_from": best["expand_from"]},
           "launch": {"walk_waves": walk_waves, "advect_block": adv_block,
                      "advect_items": adv_items},
           "chosen": best, "met_target": bool(ok), "trials": trials,
           "shape_report": shapes.strip().splitlines(),
           "sample": {"effects": [int(e) for e, _, _ in judge.items],
                      "points_per_effect": sample_points}}
    if path:
        with open(path, "w") as fh:
            json.dump(rec, fh, indent=1)
    return rec


def load_choice(batch, path):
    """Apply a choice written by tune_batch to a batch."""
    with open(path) as fh:
        rec = json.load(fh)
    m, l = rec["method"], rec["launch"]
    batch.set_method(m["near_k"], m["order"], m["expand_from"])
    batch.set_launch(l["walk_waves"], l["advect_block"], l["advect_items"])
    return rec

"""Training of the learned far-field correction on the GPU.

The network (vg_mlp) is trained on pairs the direct solver generates: for
random (grid point, applied cluster) pairs of the near-field treecode, the
target is the direct sum over the cluster's member vortons minus the
cluster's monopole (vg_tree_mlp_samples).  The training fields are drawn
from the vortex-ring, colliding-ring and buoyant-plume distributions;
nothing here reads a file.
"""
from __future__ import annotations

import time

import numpy as np

from . import scenarios, vg

# (generator, vorton count, seed) of the training fields
TRAINING_FIELDS = (
    (scenarios.vortex_ring, 16384, 51),
    (scenarios.colliding_rings, 32768, 52),
    (scenarios.jet_plume, 32768, 53),
)


def training_pairs(near_k: int, pairs_per_field: int = 65536, layer_mask: int = 0xFFFF,
                   fields=TRAINING_FIELDS, seed: int = 60):
    """Host arrays (x [n,16], target [n,3]) drawn from the training fields."""
    xs, ts = [], []
    for k, (gen, n, s) in enumerate(fields):
        pos, om, radius = gen(n, seed=s)
        dp, dw = vg.DeviceArray.from_numpy(pos), vg.DeviceArray.from_numpy(om)
        v = vg.vortons_view(dp, dw, len(pos), uniform_size=2 * radius)
        mn, mx = vg.find_bounding_box(dp, len(pos))
        g = vg.finalize_bounding_box(mn, mx, radius, len(pos))
        t = vg.Tree()
        t.build(g, v)
        spread, circ = vg.auto_mollify(g, radius)
        x, tt, _, failed = t.mlp_samples(g, pairs_per_field, near_k=near_k, layer_mask=layer_mask,
                                         spread=spread, circ=circ, seed=seed + k)
        if failed:
            raise RuntimeError(f"{failed} training pairs could not be drawn")
        xs.append(x.numpy())
        ts.append(tt.numpy())
    X = np.concatenate(xs)
    T = np.concatenate(ts)
    perm = np.random.default_rng(seed).permutation(len(X))
    return X[perm], T[perm]


def train_far_field(near_k: int = 1, pairs_per_field: int = 65536, steps=(1500, 1500),
                    lrs=(2e-3, 5e-4), batch: int = 4096, seed: int = 7, layer_mask: int = 0xFFFF):
    """Returns (trained vg.Mlp, info dict with the final loss and timings)."""
    t0 = time.perf_counter()
    X, T = training_pairs(near_k, pairs_per_field, layer_mask)
    t1 = time.perf_counter()
    m = vg.Mlp(seed=seed)
    dx, dt = vg.DeviceArray.from_numpy(X), vg.DeviceArray.from_numpy(T)
    loss = None
    for n_steps, lr in zip(steps, lrs):
        loss = m.train(dx, dt, len(X), batch=batch, steps=n_steps, lr=lr)
    vg.synchronize()
    t2 = time.perf_counter()
    return m, {"pairs": len(X), "final_minibatch_loss": loss, "sample_s": t1 - t0,
               "train_s": t2 - t1, "steps": int(sum(steps)), "batch": batch}

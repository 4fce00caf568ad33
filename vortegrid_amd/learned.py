"""Training of the learned far-field correction on the GPU.

The network (vg_mlp) is trained on pairs the direct solver generates: for
random (grid point, applied cluster) pairs of the near-field treecode, the
target is the direct sum over the cluster's member vortons minus the
cluster's monopole and analytic first-moment term (vg_tree_mlp_samples);
the loss is on the correction vector B c that the network's outputs c
weight.  The training fields are drawn from the vortex-ring, colliding-ring
and buoyant-plume distributions, each
under several random rotations (the Biot-Savart law is rotation-equivariant
and these generators are axis-aligned); nothing here reads a file.

Inputs and targets are standardised for training and the scalings are
folded back into the first and last layers afterwards, so the network the
kernels evaluate takes the raw features and returns the raw correction.
"""
from __future__ import annotations

import time

import numpy as np

from . import scenarios, vg

# (generator, vorton count, seed, keyword arguments) of the training fields:
# the ring, colliding-ring and plume families over a range of shapes
TRAINING_FIELDS = (
    (scenarios.vortex_ring, 16384, 51, {}),
    (scenarios.vortex_ring, 16384, 54, {"R": 0.5, "r": 0.15}),
    (scenarios.vortex_ring, 16384, 55, {"R": 1.2, "r": 0.35}),
    (scenarios.colliding_rings, 32768, 52, {}),
    (scenarios.colliding_rings, 32768, 56, {"separation": 0.6, "r": 0.3}),
    (scenarios.jet_plume, 32768, 53, {}),
    (scenarios.jet_plume, 32768, 57, {"sigma": 0.4, "length": 2.0}),
)
N_IN = vg.MLP_INPUTS


def random_rotation(rng) -> np.ndarray:
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def training_pairs(near_k: int, pairs_per_field: int = 24576, rotations: int = 4,
                   layer_mask: int = 0xFFFF, fields=TRAINING_FIELDS, seed: int = 60):
    """Host arrays (x [n,16], basis [n,3,3], target [n,3]) from rotated
    copies of the training fields."""
    rng = np.random.default_rng(seed)
    xs, bs, ts = [], [], []
    for k, (gen, n, s, kw) in enumerate(fields):
        pos0, om0, radius = gen(n, seed=s, **kw)
        centre = pos0.mean(0)
        for r in range(rotations):
            R = random_rotation(rng).astype(np.float32)
            pos = ((pos0 - centre) @ R.T + centre).astype(np.float32)
            om = (om0 @ R.T).astype(np.float32)
            dp, dw = vg.DeviceArray.from_numpy(pos), vg.DeviceArray.from_numpy(om)
            v = vg.vortons_view(dp, dw, len(pos), uniform_size=2 * radius)
            mn, mx = vg.find_bounding_box(dp, len(pos))
            g = vg.finalize_bounding_box(mn, mx, radius, len(pos))
            t = vg.Tree()
            t.build(g, v)
            spread, circ = vg.auto_mollify(g, radius)
            x, b, tt, _, failed = t.mlp_samples(g, pairs_per_field, near_k=near_k,
                                                layer_mask=layer_mask, spread=spread, circ=circ,
                                                seed=seed + 16 * k + r)
            if failed:
                raise RuntimeError(f"{failed} training pairs could not be drawn")
            xs.append(x.numpy())
            bs.append(b.numpy())
            ts.append(tt.numpy())
    X = np.concatenate(xs)
    B = np.concatenate(bs)
    T = np.concatenate(ts)
    perm = rng.permutation(len(X))
    return X[perm], B[perm], T[perm]


def fold_scalings(params: np.ndarray, mu: np.ndarray, sigma: np.ndarray, t_scale: float):
    """Network trained on (x - mu)/sigma -> t/t_scale  ==>  the same network
    on raw x returning raw t: W1 rows / sigma, b1 - (mu/sigma) W1,
    W4 and b4 times t_scale."""
    p = np.asarray(params, np.float64).copy()
    h = vg.MLP_HIDDEN
    W1 = p[:N_IN * h].reshape(N_IN, h)
    b1 = p[N_IN * h:N_IN * h + h]
    b1 -= (mu / sigma) @ W1
    W1 /= sigma[:, None]
    o4 = N_IN * h + h + 2 * (h * h + h)
    p[o4:o4 + h * 3 + 3] *= t_scale
    return p.astype(np.float32)


def train_far_field(near_k: int = 1, pairs_per_field: int = 24576, rotations: int = 4,
                    steps=(3000, 2000, 1000), lrs=(2e-3, 5e-4, 1e-4), batch: int = 4096,
                    seed: int = 7, layer_mask: int = 0xFFFF):
    """Returns (trained vg.Mlp, info dict with the final loss and timings)."""
    t0 = time.perf_counter()
    X, B, T = training_pairs(near_k, pairs_per_field, rotations, layer_mask)
    t1 = time.perf_counter()
    mu = X.mean(0).astype(np.float64)
    sigma = X.std(0).astype(np.float64)
    sigma[sigma < 1e-12] = 1.0
    t_scale = float(np.sqrt(np.mean(T.astype(np.float64) ** 2))) or 1.0
    Xn = ((X - mu) / sigma).astype(np.float32)
    Tn = (T / t_scale).astype(np.float32)
    m = vg.Mlp(seed=seed)
    dx, dt = vg.DeviceArray.from_numpy(Xn), vg.DeviceArray.from_numpy(Tn)
    db = vg.DeviceArray.from_numpy(B.astype(np.float32))
    loss = None
    for n_steps, lr in zip(steps, lrs):
        loss = m.train(dx, dt, len(X), batch=batch, steps=n_steps, lr=lr, basis=db)
    vg.synchronize()
    t2 = time.perf_counter()
    m.set_params(fold_scalings(m.params(), mu, sigma, t_scale))
    return m, {"pairs": len(X), "final_minibatch_loss_standardised": loss,
               "target_rms": t_scale, "sample_s": t1 - t0, "train_s": t2 - t1,
               "steps": int(sum(steps)), "batch": batch}

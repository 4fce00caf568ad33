"""CPU oracle of the learned far-field correction (test infrastructure only).

The product's learned far-field correction has no reference counterpart:
the reference's only cluster representation is the monopole supervorton
(vortonSim.cpp:1337-1569 AggregateClusters, applied by vorton.h:161-179).
This module restates, in float64 numpy, what the HIP kernels compute so that
the tests can check them:

  features(q, cluster)  16 invariant inputs and the 3 equivariant output
                        vectors per (query, cluster) pair
  forward / backward    the 16 -> 64 -> 64 -> 64 -> 3 GELU(tanh) network and
                        the gradient of its mean-squared loss
  bf16 emulation        optional round-to-nearest-even of every GEMM operand
                        (the MFMA kernels read bf16 operands, accumulate FP32)

Pinned by the tests: the backward pass against central finite differences in
float64 (with and without an output basis), the forward pass against a
per-sample scalar loop, GELU against its closed form at known points, and the
features against their symmetries (rotating the field rotates B c and leaves
the inputs unchanged; reversing every vorticity reverses B c).
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT = 16, 64, 3
SHAPES = [("W1", (N_IN, N_HID)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
          ("W3", (N_HID, N_HID)), ("b3", (N_HID,)), ("W4", (N_HID, N_OUT)), ("b4", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)  # 9603

C_GELU = np.sqrt(2.0 / np.pi)
A_GELU = 0.044715


def unpack(flat):
    """Compact parameter vector (W1 [16][64] row-major [in][out], b1, W2,
    b2, W3, b3, W4 [64][3], b4) -> dict of float64 arrays."""
    flat = np.asarray(flat, np.float64)
    assert flat.size == N_PARAMS
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = flat[o:o + n].reshape(shp)
        o += n
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], np.float64).reshape(-1) for name, _ in SHAPES])


def bf16(a):
    """Round float values to bfloat16 (round to nearest, ties to even), as
    float64.  Inputs are first rounded to float32, as the kernels hold them."""
    a32 = np.ascontiguousarray(np.asarray(a, np.float32))
    u = a32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    t = np.tanh(C_GELU * (z + A_GELU * z ** 3))
    return 0.5 * z * (1.0 + t)


def gelu_grad(z):
    u = C_GELU * (z + A_GELU * z ** 3)
    t = np.tanh(u)
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C_GELU * (1.0 + 3.0 * A_GELU * z * z)


def _op(a, emulate):
    return bf16(a) if emulate else np.asarray(a, np.float64)


def forward(flat, X, emulate_bf16=False):
    """Returns (Y [n,3], cache).  With emulate_bf16, every GEMM operand
    (inputs, weights, activations) is rounded to bf16 first; bias adds,
    GELU and accumulation stay in high precision."""
    p = unpack(flat)
    e = emulate_bf16
    X = np.asarray(X, np.float64)
    Z1 = _op(X, e) @ _op(p["W1"], e) + p["b1"]
    A1 = gelu(Z1)
    Z2 = _op(A1, e) @ _op(p["W2"], e) + p["b2"]
    A2 = gelu(Z2)
    Z3 = _op(A2, e) @ _op(p["W3"], e) + p["b3"]
    A3 = gelu(Z3)
    Y = _op(A3, e) @ _op(p["W4"], e) + p["b4"]
    return Y, (X, Z1, A1, Z2, A2, Z3, A3)


def apply_basis(Y, B):
    """Correction vectors B c for network outputs c (B None: c itself)."""
    return Y if B is None else np.einsum("nik,nk->ni", np.asarray(B, np.float64), Y)


def loss(flat, X, T, emulate_bf16=False, B=None):
    Y, _ = forward(flat, X, emulate_bf16)
    return float(np.mean((apply_basis(Y, B) - np.asarray(T, np.float64)) ** 2))


def backward(flat, X, T, emulate_bf16=False, B=None):
    """Gradient of mean((B c - T)^2) over all n*3 components (B None: the
    identity); returns (loss, compact gradient vector).  With emulate_bf16
    the backward GEMM operands (dY, dZ, activations, weights) are rounded to
    bf16 too."""
    p = unpack(flat)
    e = emulate_bf16
    Y, (X, Z1, A1, Z2, A2, Z3, A3) = forward(flat, X, e)
    T = np.asarray(T, np.float64)
    n = X.shape[0]
    R = apply_basis(Y, B) - T
    dY = 2.0 * R / (3.0 * n)
    if B is not None:
        dY = np.einsum("nik,ni->nk", np.asarray(B, np.float64), dY)
    g = {}
    g["W4"] = _op(A3, e).T @ _op(dY, e)
    g["b4"] = dY.sum(0)
    dZ3 = (_op(dY, e) @ _op(p["W4"], e).T) * gelu_grad(Z3)
    g["W3"] = _op(A2, e).T @ _op(dZ3, e)
    g["b3"] = dZ3.sum(0)
    dZ2 = (_op(dZ3, e) @ _op(p["W3"], e).T) * gelu_grad(Z2)
    g["W2"] = _op(A1, e).T @ _op(dZ2, e)
    g["b2"] = dZ2.sum(0)
    dZ1 = (_op(dZ2, e) @ _op(p["W2"], e).T) * gelu_grad(Z1)
    g["W1"] = _op(X, e).T @ _op(dZ1, e)
    g["b1"] = dZ1.sum(0)
    return float(np.mean(R ** 2)), pack(g)


def numeric_gradient(flat, X, T, idx, h=1e-6, B=None):
    """Central differences of the float64 loss for parameters idx."""
    flat = np.asarray(flat, np.float64).copy()
    out = np.empty(len(idx))
    for j, i in enumerate(idx):
        old = flat[i]
        flat[i] = old + h
        lp = loss(flat, X, T, B=B)
        flat[i] = old - h
        lm = loss(flat, X, T, B=B)
        flat[i] = old
        out[j] = (lp - lm) / (2 * h)
    return out


def init_params(seed):
    """Reference initialiser for tests (uniform +-sqrt(6/(fan_in+fan_out)),
    zero biases); the GPU keeps its own deterministic initialiser."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shp in SHAPES:
        if name[0] == "W":
            lim = np.sqrt(6.0 / (shp[0] + shp[1]))
            p[name] = rng.uniform(-lim, lim, shp)
        else:
            p[name] = np.zeros(shp)
    return pack(p)


# --------------------------------------------------------------- features
def features(q, P, W, M, h):
    """Inputs and output basis of the correction network for query points q
    [n,3] against clusters P (x, y, z, size), W (Omega, S), M (3 rows of the
    first moment D[b][e] = sum Omega_b (y - c)_e, .w = antisymmetric part a)
    on a layer of cell size h (the cube root of the cell volume).

    With x = (q - c)/h, r = |x|, e = x/r, O = Omega/S, Dt = D/(h S),
    at = a/(h S), u = Dt e, v = Dt^T e, the 16 inputs are rotation invariants
    that do not change when every vorticity changes sign:
      1/r, 1/r^2, |O|^2, (O.e)^2, |at|^2, (at.e)^2, |u|^2, |v|^2, (e.u)^2,
      |Dt|_F^2, (tr Dt)^2, O.at, (O.e)(e.u), O.u, O.v, at.(u x e)
    and the outputs c_k weight three vectors that rotate with the field and
    change sign with the vorticity, times r^-3:
      B = r^-3 [O x e, at, u x e]          (columns)
    The correction is (2/3)(size/2)^3 S / h^2 * B c.  Returns (X [n,16],
    B [n,3,3])."""
    q = np.asarray(q, np.float64)
    P = np.asarray(P, np.float64)
    W = np.asarray(W, np.float64)
    M = np.asarray(M, np.float64).reshape(-1, 3, 4)
    x = (q - P[:, :3]) / h
    ir = 1.0 / np.sqrt((x * x).sum(1))
    e = x * ir[:, None]
    S = W[:, 3]
    O = W[:, :3] / S[:, None]
    Dt = M[:, :, :3] / (h * S)[:, None, None]
    at = M[:, :, 3] / (h * S)[:, None]
    u = np.einsum("nbe,ne->nb", Dt, e)
    v = np.einsum("nbe,nb->ne", Dt, e)
    dot = lambda a, b: (a * b).sum(1)
    eu = dot(e, u)
    X = np.stack([ir, ir * ir, dot(O, O), dot(O, e) ** 2, dot(at, at), dot(at, e) ** 2,
                  dot(u, u), dot(v, v), eu ** 2, (Dt * Dt).sum((1, 2)),
                  np.trace(Dt, axis1=1, axis2=2) ** 2, dot(O, at), dot(O, e) * eu, dot(O, u),
                  dot(O, v), dot(at, np.cross(u, e))], 1)
    B = np.stack([np.cross(O, e), at, np.cross(u, e)], 2) * (ir ** 3)[:, None, None]
    return X, B


def correction_scale(P, W, h):
    P = np.asarray(P, np.float64)
    W = np.asarray(W, np.float64)
    r = 0.5 * P[:, 3]
    return (2.0 / 3.0) * r ** 3 * W[:, 3] / (h * h)


# --------------------------------------------------------------- tree walk
FLT_MIN = np.float32(1.1754943508222875e-38)


def weight_sums(ot, pos, om):
    """Per-layer cluster weight sums S (the tree's W.w): leaf cells sum
    |2 Omega| of their vortons, parent cells |2 Omega| of their children's
    summed Omega, both starting from FLT_MIN (vortonSim.cpp:1192-1321,
    1337-1569 weights).  float64, so equal to the kernels' float32 sums to
    rounding."""
    g0, _, _, _ = ot.layer(0)
    start, count, order = cell_lists(ot)
    cell_of = np.repeat(np.arange(g0.capacity), count)
    S = [np.full(g0.capacity, float(FLT_MIN))]
    mag = 2.0 * np.linalg.norm(np.asarray(om, np.float64), axis=1)
    np.add.at(S[0], cell_of, mag[order])
    for l in range(1, ot.depth):
        gp, _, _, _ = ot.layer(l)
        gc, _, omc, _ = ot.layer(l - 1)
        d = ot.decimations(l)
        s = np.zeros(gp.capacity)
        mag = 2.0 * np.linalg.norm(omc.astype(np.float64), axis=1).reshape(gc.np[2], gc.np[1], gc.np[0])
        nc = [gp.np[a] - 1 for a in range(3)]
        acc = np.full((nc[2], nc[1], nc[0]), float(FLT_MIN))
        for cz in range(d[2]):
            for cy in range(d[1]):
                for cx in range(d[0]):
                    acc += mag[cz:cz + nc[2] * d[2]:d[2], cy:cy + nc[1] * d[1]:d[1],
                               cx:cx + nc[0] * d[0]:d[0]]
        s3 = s.reshape(gp.np[2], gp.np[1], gp.np[0])
        s3[:nc[2], :nc[1], :nc[0]] = acc
        S.append(s)
    return S


def layer_h(ot):
    """Cube root of each layer's cell volume (the feature length scale)."""
    out = []
    for l in range(ot.depth):
        g, _, _, _ = ot.layer(l)
        out.append(float(np.cbrt(np.float64(np.float32(g.cell[0] * g.cell[1]) * g.cell[2]))))
    return out


def accepted_clusters(ot, g, point, near_k):
    """Clusters the near-field treecode applies at grid point `point`
    (ix, iy, iz), restating the brick walk of vg_oracle.c mac_brick: the
    point's 4x4x4 brick, per layer the children of the parent near box minus
    the layer's near box (brick cells +- near_k).  Returns ([(layer,
    cluster offset)], leaf near box (lo, hi))."""
    depth = ot.depth
    bb = [p // 4 for p in point]
    cl0, cl1 = [], []
    for a in range(3):
        nc = g.np[a] - 1
        first, last = 4 * bb[a], min(4 * bb[a] + 3, g.np[a] - 1)
        cl0.append(min(first, nc - 1))
        cl1.append(min(last, nc - 1))
    scale = [[1, 1, 1]]
    for l in range(1, depth):
        d = ot.decimations(l)
        scale.append([scale[-1][a] * d[a] for a in range(3)])
    lo, hi = [0, 0, 0], [0, 0, 0]
    out = []
    for l in range(depth - 2, -1, -1):
        gl, _, _, size = ot.layer(l)
        d = ot.decimations(l + 1)
        c0 = [lo[a] * d[a] for a in range(3)]
        c1 = [hi[a] * d[a] + d[a] - 1 for a in range(3)]
        n0, n1 = [], []
        for a in range(3):
            ncl = gl.np[a] - 1
            blo = max(cl0[a] // scale[l][a] - near_k, 0)
            bhi = min(cl1[a] // scale[l][a] + near_k, ncl - 1)
            n0.append(max(blo, c0[a]))
            n1.append(min(bhi, c1[a]))
        for cz in range(c0[2], c1[2] + 1):
            for cy in range(c0[1], c1[1] + 1):
                for cx in range(c0[0], c1[0] + 1):
                    if n0[0] <= cx <= n1[0] and n0[1] <= cy <= n1[1] and n0[2] <= cz <= n1[2]:
                        continue
                    off = cx + gl.np[0] * (cy + gl.np[1] * cz)
                    if size[off] != 0.0:
                        out.append((l, off))
        lo, hi = n0, n1
    if depth == 1:
        g0 = ot.layer(0)[0]
        lo, hi = [0, 0, 0], [g0.np[a] - 2 for a in range(3)]
    return out, (lo, hi)


def cell_lists(ot):
    """(cell_start, cell_count, sorted vorton indices) of the oracle tree's
    leaf layer."""
    t = ot._t
    cap = ot.layer(0)[0].capacity
    start = np.ctypeslib.as_array(t.cell_start, shape=(cap,)).copy()
    count = np.ctypeslib.as_array(t.cell_count, shape=(cap,)).copy()
    order = np.ctypeslib.as_array(t.sorted_idx, shape=(max(int(t.n), 1),)).copy()[:int(t.n)]
    return start, count, order


def cluster_members(ot, layer, off):
    """Vorton indices of cluster `off` on `layer`: the vortons of the leaf
    cells it covers."""
    gl = ot.layer(layer)[0]
    g0 = ot.layer(0)[0]
    start, count, order = cell_lists(ot)
    c = [off % gl.np[0], (off // gl.np[0]) % gl.np[1], off // (gl.np[0] * gl.np[1])]
    sc = [1, 1, 1]
    for l in range(1, layer + 1):
        d = ot.decimations(l)
        sc = [sc[a] * d[a] for a in range(3)]
    idx = []
    for z in range(c[2] * sc[2], min((c[2] + 1) * sc[2], g0.np[2] - 1)):
        for y in range(c[1] * sc[1], min((c[1] + 1) * sc[1], g0.np[1] - 1)):
            for x in range(c[0] * sc[0], min((c[0] + 1) * sc[0], g0.np[0] - 1)):
                k = x + g0.np[0] * (y + g0.np[1] * z)
                idx.extend(order[start[k]:start[k] + count[k]].tolist())
    return np.asarray(idx, np.int64)

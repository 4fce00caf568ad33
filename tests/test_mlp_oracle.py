"""The learned far-field correction's CPU oracle, pinned against what the
mathematics fixes: finite differences, a scalar restatement, the erf GELU,
torch's bfloat16 rounding, the monopole law and the C oracle's tree walk."""
import math

import numpy as np
import pytest

from oracle import mlp_oracle as M
from vortegrid_amd import scenarios


def test_param_layout():
    assert M.N_PARAMS == 9603
    p = M.init_params(3)
    assert np.array_equal(M.pack(M.unpack(p)), p)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = M.init_params(1)
    p += 0.05 * rng.standard_normal(p.size)  # non-zero biases too
    X = rng.standard_normal((40, M.N_IN))
    T = rng.standard_normal((40, M.N_OUT))
    _, g = M.backward(p, X, T)
    idx = rng.choice(p.size, 120, replace=False)
    num = M.numeric_gradient(p, X, T, idx)
    assert np.allclose(g[idx], num, rtol=1e-5, atol=1e-9)


def test_forward_matches_scalar_loop():
    rng = np.random.default_rng(2)
    p = M.init_params(4) + 0.02 * rng.standard_normal(M.N_PARAMS)
    X = rng.standard_normal((3, M.N_IN))
    Y, _ = M.forward(p, X)
    P = M.unpack(p)

    def g(z):  # erf-free tanh GELU, scalar
        return 0.5 * z * (1 + math.tanh(math.sqrt(2 / math.pi) * (z + 0.044715 * z ** 3)))

    for r in range(3):
        h = list(X[r])
        for W, b in ((P["W1"], P["b1"]), (P["W2"], P["b2"]), (P["W3"], P["b3"])):
            h = [g(sum(h[i] * W[i][j] for i in range(len(h))) + b[j]) for j in range(W.shape[1])]
        y = [sum(h[i] * P["W4"][i][j] for i in range(64)) + P["b4"][j] for j in range(3)]
        assert np.allclose(Y[r], y, rtol=1e-12, atol=1e-12)


def test_gelu_against_erf_form():
    z = np.linspace(-6, 6, 241)
    exact = np.array([0.5 * v * (1 + math.erf(v / math.sqrt(2))) for v in z])
    assert np.max(np.abs(M.gelu(z) - exact)) < 1e-3
    assert M.gelu(np.array([0.0]))[0] == 0.0
    # derivative by central differences
    h = 1e-6
    assert np.allclose(M.gelu_grad(z), (M.gelu(z + h) - M.gelu(z - h)) / (2 * h), atol=1e-7)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    a = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 10000),
                        np.float32([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -1.0 - 2 ** -8, 0.0])])
    want = torch.tensor(a).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(M.bf16(a), want)


def _random_cluster(rng):
    c = rng.uniform(-1, 1, 3)
    om = rng.standard_normal(3)
    S = 2 * np.linalg.norm(om) * rng.uniform(1.0, 3.0)
    D = 0.05 * rng.standard_normal((3, 3))
    a = np.array([D[1, 2] - D[2, 1], D[2, 0] - D[0, 2], D[0, 1] - D[1, 0]])
    P = np.array([[*c, 0.1]])
    W = np.array([[*om, S]])
    Mm = np.concatenate([D, a[:, None]], 1).reshape(1, 12)
    return c, om, S, D, P, W, Mm


def test_features_are_invariant_and_basis_equivariant():
    """Rotating the field leaves the inputs unchanged and rotates B c;
    reversing every vorticity leaves the inputs unchanged and reverses B c."""
    rng = np.random.default_rng(7)
    from vortegrid_amd import learned

    for _ in range(20):
        c, om, S, D, P, W, Mm = _random_cluster(rng)
        q = c + rng.uniform(2, 4) * rng.standard_normal(3)
        h = 0.3
        X, B = M.features(q[None], P, W, Mm, h)
        coef = rng.standard_normal(3)
        R = learned.random_rotation(rng)
        Dr = R @ D @ R.T
        ar = np.array([Dr[1, 2] - Dr[2, 1], Dr[2, 0] - Dr[0, 2], Dr[0, 1] - Dr[1, 0]])
        Pr = np.array([[*(R @ c), 0.1]])
        Wr = np.array([[*(R @ om), S]])
        Mr = np.concatenate([Dr, ar[:, None]], 1).reshape(1, 12)
        Xr, Br = M.features((R @ q)[None], Pr, Wr, Mr, h)
        assert np.allclose(Xr, X, rtol=1e-10, atol=1e-14)
        assert np.allclose(Br[0] @ coef, R @ (B[0] @ coef), rtol=1e-10, atol=1e-14)
        Xn, Bn = M.features(q[None], P, np.array([[*(-om), S]]), -Mm, h)
        assert np.allclose(Xn, X, rtol=1e-12, atol=1e-15)
        assert np.allclose(Bn, -B, rtol=1e-12, atol=1e-15)


def test_basis_spans_monopole_and_first_moment():
    """c = (r, 0, 0) gives the monopole law and c = (0, -1, 3) the analytic
    first-moment term, both times the output scale."""
    rng = np.random.default_rng(8)
    for _ in range(20):
        c, om, S, D, P, W, Mm = _random_cluster(rng)
        q = c + rng.uniform(2, 4) * rng.standard_normal(3)
        h = 0.3
        X, B = M.features(q[None], P, W, Mm, h)
        sc = M.correction_scale(P, W, h)[0]
        d = q - c
        s = (2.0 / 3.0) * 0.05 ** 3
        mono = s * np.cross(om, d) / np.linalg.norm(d) ** 3
        a = Mm.reshape(3, 4)[:, 3]
        dip = -s * (a - 3.0 / (d @ d) * np.cross(D @ d, d)) / np.linalg.norm(d) ** 3
        r = 1.0 / X[0, 0]
        assert np.allclose(sc * B[0] @ np.array([r, 0, 0]), mono, rtol=1e-10)
        assert np.allclose(sc * B[0] @ np.array([0, -1, 3]), dip, rtol=1e-10)


def test_basis_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = M.init_params(5) + 0.05 * rng.standard_normal(M.N_PARAMS)
    X = rng.standard_normal((30, M.N_IN))
    Bm = rng.standard_normal((30, 3, 3))
    T = rng.standard_normal((30, 3))
    _, g = M.backward(p, X, T, B=Bm)
    idx = rng.choice(p.size, 80, replace=False)
    assert np.allclose(g[idx], M.numeric_gradient(p, X, T, idx, B=Bm), rtol=1e-5, atol=1e-9)


@pytest.fixture(scope="module")
def small_tree(oracle):
    pos, om, radius = scenarios.colliding_rings(4096, seed=31)
    g = oracle.grid_for(pos, radius)
    ot = oracle.InfluenceTree(g, pos, om, 2 * radius)
    return pos, om, radius, g, ot


def test_cluster_members_partition_the_vortons(small_tree):
    pos, om, radius, g, ot = small_tree
    for l in range(ot.depth):
        gl = ot.layer(l)[0]
        seen = []
        for off in range(gl.capacity):
            seen.extend(M.cluster_members(ot, l, off).tolist())
        assert sorted(seen) == list(range(len(pos)))


@pytest.mark.parametrize("near_k", [0, 1, 2])
def test_accepted_clusters_restate_the_walk(oracle, small_tree, near_k):
    """Monopoles of the enumerated clusters plus the direct sum over the
    leaf near box reproduce the C oracle's near-field treecode."""
    pos, om, radius, g, ot = small_tree
    spread, circ = oracle.auto_mollify(g, radius)
    want = oracle.mac_velocity_grid(ot, g, pos, om, 2 * radius, near_k, 0, spread, circ, 1,
                                    threads=8).reshape(-1, 3)
    start, count, order = M.cell_lists(ot)
    rng = np.random.default_rng(near_k)
    g0 = ot.layer(0)[0]
    for _ in range(12):
        pt = [int(rng.integers(0, g.np[a])) for a in range(3)]
        q = oracle.query_points(g).reshape(-1, 3)[pt[0] + g.np[0] * (pt[1] + g.np[1] * pt[2])]
        acc, (lo, hi) = M.accepted_clusters(ot, g, pt, near_k)
        srcs_p, srcs_w = [], []
        for l, off in acc:
            _, P, W, _ = ot.layer(l)
            srcs_p.append(P[off])
            srcs_w.append(W[off])
        near = []
        for z in range(lo[2], hi[2] + 1):
            for y in range(lo[1], hi[1] + 1):
                for x in range(lo[0], hi[0] + 1):
                    k = x + g0.np[0] * (y + g0.np[1] * z)
                    near.extend(order[start[k]:start[k] + count[k]].tolist())
        u = oracle.direct_velocity(np.array(srcs_p), np.array(srcs_w), 2 * radius, q[None],
                                   spread, circ, policy=1)[0] if srcs_p else 0.0
        if near:
            u = u + oracle.direct_velocity(pos[near], om[near], 2 * radius, q[None], spread, circ,
                                           policy=1)[0]
        o = pt[0] + g.np[0] * (pt[1] + g.np[1] * pt[2])
        assert np.allclose(u, want[o], rtol=1e-4, atol=1e-6 * np.abs(want).max())


def test_folded_standardisation_is_the_same_network():
    """Training on standardised inputs / targets and folding the scalings
    into the first and last layers gives the network on raw values."""
    from vortegrid_amd import learned

    rng = np.random.default_rng(9)
    p = M.init_params(2) + 0.05 * rng.standard_normal(M.N_PARAMS)
    X = rng.standard_normal((50, M.N_IN)) * 10.0 ** rng.uniform(-3, 1, M.N_IN) + rng.standard_normal(M.N_IN)
    mu, sigma, ts = X.mean(0), X.std(0), 3.7e-3
    want = M.forward(p, (X - mu) / sigma)[0] * ts
    got = M.forward(learned.fold_scalings(p, mu, sigma, ts).astype(np.float64), X)[0]
    assert np.allclose(got, want, rtol=2e-4, atol=1e-6 * np.abs(want).max())

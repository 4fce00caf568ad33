"""Pins the CPU oracle (oracle/vg_oracle.c) without running the reference.

Each check is chosen so a plausible restatement mistake (dropped term,
wrong sign or index, transposed operand, wrong rounding order) fails:
closed-form single-vorton profiles, superposition against an independent
float64 Biot-Savart, the reference's own NestedGrid known answer
(nestedGridDiagnostics.cpp:16-96), hand-derived DefineShape results, a
manufactured Poisson solution (testFunc3D, uniformGridDiagnostics.cpp:
319-323), and exactness properties of the grid transfer operators.
"""
import numpy as np
import pytest

F32 = np.float32
EPS = np.finfo(np.float32).eps


# ------------------------------------------------------------ finvsqrtf
def test_finvsqrtf_known_value_and_bound(oracle):
    # 0x5f3759df - (0x3f800000 >> 1) = 0x3f7759df -> 0.96697..., one Newton
    # step gives 0.998307 (Lomont 2003).
    assert abs(oracle.finvsqrtf(1.0) - 0.998307) < 1e-6
    rng = np.random.default_rng(0)
    xs = np.concatenate([10.0 ** rng.uniform(-30, 30, 20000), np.linspace(1, 4, 2001)])
    worst = 0.0
    for x in xs.astype(F32):
        y = oracle.finvsqrtf(float(x))
        exact = 1.0 / np.sqrt(float(x))
        rel = (exact - y) / exact
        assert rel >= -1e-7, "one Newton step from this seed always under-estimates"
        worst = max(worst, rel)
    assert worst <= 1.7523e-3  # Lomont's bound for 0x5f3759df + 1 Newton step
    assert worst > 1.70e-3     # and the bound is attained (no extra iteration)


def test_nearest_pow2(oracle):
    for v, want in [(0, 1), (1, 1), (2, 2), (3, 4), (5, 8), (8, 8), (9, 16), (1000, 1024)]:
        assert oracle.nearest_pow2(v) == want


# ------------------------------------------------------------ pair kernels
def _acc(oracle, fn, q, p, om, size, spread=1.0, circ=1.0, policy=1):
    acc = np.zeros(3, F32)
    fn(acc, np.asarray(q, F32), np.asarray(p, F32), np.asarray(om, F32), size, spread, circ, policy)
    return acc


def test_single_vorton_velocity_profile(oracle):
    L = oracle.lib()
    r = 0.1
    om = [0.0, 0.0, 1.0]  # Omega = w/2 along z
    # outside the core: u = (2/3 r^3 Omega) x d / |d|^3 ; Omega x x-hat = y-hat
    for d in (0.2, 0.5, 1.0, 3.0):
        u = _acc(oracle, L.vgo_accumulate_velocity, [d, 0, 0], [0, 0, 0], om, 2 * r)
        want = (2.0 / 3.0) * r ** 3 / d ** 2
        assert abs(u[0]) < 1e-12 and abs(u[2]) < 1e-12
        assert u[1] == pytest.approx(want, rel=2e-6)
    # inside: linear law u = 2/3 Omega x d
    for d in (0.01, 0.05, 0.09):
        u = _acc(oracle, L.vgo_accumulate_velocity, [d, 0, 0], [0, 0, 0], om, 2 * r)
        assert u[1] == pytest.approx((2.0 / 3.0) * d, rel=2e-6)
    # continuous at the core radius
    a = _acc(oracle, L.vgo_accumulate_velocity, [r * (1 - 1e-6), 0, 0], [0, 0, 0], om, 2 * r)
    b = _acc(oracle, L.vgo_accumulate_velocity, [r * (1 + 1e-6), 0, 0], [0, 0, 0], om, 2 * r)
    assert a[1] == pytest.approx(b[1], rel=1e-5)
    # spreading: radius scaled by spread, strength by circ
    u1 = _acc(oracle, L.vgo_accumulate_velocity, [2.0, 0, 0], [0, 0, 0], om, 2 * r, 2.0, 0.125)
    u0 = _acc(oracle, L.vgo_accumulate_velocity, [2.0, 0, 0], [0, 0, 0], om, 2 * r)
    assert u1[1] == pytest.approx(u0[1], rel=2e-6)  # far field: (2r)^3 / 8 == r^3


def test_pair_velocity_symmetries(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(1)
    for _ in range(200):
        p = rng.normal(size=3)
        om = rng.normal(size=3)
        d = rng.normal(size=3)
        size = float(rng.uniform(0.05, 0.5))
        u_plus = _acc(oracle, L.vgo_accumulate_velocity, p + d, p, om, size)
        u_minus = _acc(oracle, L.vgo_accumulate_velocity, p - d, p, om, size)
        scale = np.abs(u_plus).max() + 1e-30
        assert np.allclose(u_plus, -u_minus, atol=3e-6 * scale)
        assert abs(np.dot(u_plus, om)) <= 1e-5 * scale * np.linalg.norm(om)
        assert abs(np.dot(u_plus, d)) <= 1e-5 * scale * np.linalg.norm(d)


def test_single_vorton_vecpot_profile(oracle):
    L = oracle.lib()
    r = 0.1
    om = np.array([0.3, -0.2, 1.0])
    for d in (0.2, 1.0, 4.0):
        a = _acc(oracle, L.vgo_accumulate_vecpot, [0, d, 0], [0, 0, 0], om, 2 * r)
        assert np.allclose(a, (2.0 / 3.0) * r ** 3 * om / d, rtol=2e-6)
    for d in (0.02, 0.08):
        a = _acc(oracle, L.vgo_accumulate_vecpot, [0, 0, d], [0, 0, 0], om, 2 * r)
        assert np.allclose(a, (2.0 / 3.0) * om * d * d, rtol=2e-6)


def _biot_savart64(pos, om, radius, q):
    """Independent float64 Biot-Savart of uniform-core vortons: u = sum
    (1/4pi) V w x d / max(|d|, r)^3 with V = 4/3 pi r^3, w = 2 Omega."""
    pos = pos.astype(np.float64)
    w = 2.0 * om.astype(np.float64)
    out = np.zeros((len(q), 3))
    V = 4.0 / 3.0 * np.pi * radius ** 3
    for i, x in enumerate(q.astype(np.float64)):
        d = x - pos
        dist = np.linalg.norm(d, axis=1)
        denom = np.maximum(dist, radius) ** 3
        out[i] = (V / (4 * np.pi) * np.cross(w, d) / denom[:, None]).sum(0)
    return out


def test_direct_sum_superposition(oracle):
    rng = np.random.default_rng(2)
    n = 300
    pos = rng.normal(size=(n, 3)).astype(F32)
    om = rng.normal(size=(n, 3)).astype(F32)
    radius = 0.15
    q = rng.normal(size=(64, 3)).astype(F32)
    # the oracle receives the diameter as float32
    want = _biot_savart64(pos, om, float(F32(2 * radius)) / 2, q)
    got = oracle.direct_velocity(pos, om, 2 * radius, q, policy=1)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 2e-6
    got64 = oracle.direct_velocity_f64(pos, om, 2 * radius, q)
    assert np.linalg.norm(got64 - want) / np.linalg.norm(want) < 1e-13
    # reference fast rsqrt: bounded by the finvsqrtf error
    fast = oracle.direct_velocity(pos, om, 2 * radius, q, policy=0)
    assert np.linalg.norm(fast - want) / np.linalg.norm(want) < 1.8e-3
    # threads do not change results
    assert np.array_equal(fast, oracle.direct_velocity(pos, om, 2 * radius, q, policy=0, threads=4))


# ------------------------------------------------------------ geometry
def test_define_shape_known_answers(oracle):
    # NestedGrid<unsigned>::UnitTest box: 1024 elements in [-1,1]x[-1.5,1.5]
    # x[-2.5,2.5]: volume 30, (30/1024)^(-1/3) = 3.249, cells = floor(2*3.249
    # +.5)=6 -> 8, floor(3*3.249+.5)=10 -> 16, floor(5*3.249+.5)=16 -> 16.
    g = oracle.define_shape(1024, [-1, -1.5, -2.5], [1, 1.5, 2.5])
    assert g.shape == (9, 17, 17)
    # ring box: 4096 in 2.6 x 2.6 x 0.6 -> (4.056/4096)^(-1/3) = 10.03:
    # 26.6 -> 26 -> 32 and 6.5 -> 6 -> 8.
    assert oracle.define_shape(4096, [-1.3, -1.3, -0.3], [1.3, 1.3, 0.3]).shape == (33, 33, 9)
    # halving loop: 1 element in 10x1x1 -> x: 10*10^(-1/3)+.5 = 5.1 -> 5 -> 8;
    # 8*1*1 >= 8*1 so halve to 4 -> 5 points.
    assert oracle.define_shape(1, [0, 0, 0], [10, 1, 1]).shape == (5, 2, 2)
    assert oracle.define_shape(1, [0, 0, 0], [1, 1, 1]).shape == (2, 2, 2)
    # flat z: 2-D grid, z keeps one cell and cellsPerExtent.z = 1/FLT_MIN
    g2 = oracle.define_shape(256, [0, 0, 0], [1, 1, 0])
    assert g2.np[2] == 2 and g2.extent[2] == 0.0 and g2.cpe[2] == pytest.approx(1.0 / np.finfo(F32).tiny)
    assert oracle.define_shape(0, [0, 0, 0], [1, 1, 1]).capacity == 0
    # spacing and its reciprocal follow extent / cells with extent nudged by 1+eps
    assert g.extent[0] == F32(F32(2.0) * F32(1.0 + EPS))
    assert g.cell[1] == F32(g.extent[1] / F32(16))
    assert g.cpe[2] == F32(F32(16) / g.extent[2])


def test_nested_grid_reference_unit_test(oracle):
    """nestedGridDiagnostics.cpp:16-96: depth 5, decimations in {1,2},
    child cells = parent cells * decimation, parent spacing = child spacing
    * decimation (exact), and cluster corners agree."""
    g = oracle.define_shape(1024, [-1, -1.5, -2.5], [1, 1.5, 2.5])
    ng = oracle.nested(g)
    assert ng.depth == 5
    assert tuple(ng.layer[0].np) == tuple(g.np)
    for l in range(ng.depth - 1, 0, -1):
        par, ch = ng.layer[l], ng.layer[l - 1]
        for a in range(3):
            d = ng.dec[l][a]
            assert d in (1, 2)
            assert ch.np[a] - 1 == (par.np[a] - 1) * d
            assert F32(par.cell[a]) == F32(F32(ch.cell[a]) * F32(d))
    assert tuple(ng.layer[ng.depth - 1].np) == (2, 2, 2)


def test_indices_round_trip(oracle):
    """uniformGridDiagnostics.cpp:328-420: positions just inside each cell's
    min and max corners map back to that cell."""
    rng = np.random.default_rng(3)
    pts = (np.array([2.0, 3.0, 5.0]) * (rng.uniform(size=(1024, 3)) - 0.5)).astype(F32)
    mn, mx = pts.min(0), pts.max(0)
    g = oracle.define_shape(1024, mn, mx)
    nudge = F32(2.0) * EPS * np.array(g.extent, F32)
    for iz in range(g.np[2] - 1):
        for iy in range(0, g.np[1] - 1, 3):
            for ix in range(0, g.np[0] - 1, 2):
                idx = np.array([ix, iy, iz])
                lo = (np.array(g.min, F32) + idx.astype(F32) * np.array(g.cell, F32) + nudge).astype(F32)
                hi = (lo + np.array(g.cell, F32) - F32(2) * nudge).astype(F32)
                assert tuple(oracle.indices_of_position(g, lo)) == tuple(idx)
                assert tuple(oracle.indices_of_position(g, hi)) == tuple(idx)


# ------------------------------------------------------------ tree
@pytest.fixture(scope="module")
def ring_case(oracle):
    from vortegrid_amd import scenarios

    pos, om, radius = scenarios.vortex_ring(2048, seed=11)
    g = oracle.grid_for(pos, radius)
    return pos, om, radius, g


def test_tree_conserves_circulation(oracle, ring_case):
    pos, om, radius, g = ring_case
    t = oracle.InfluenceTree(g, pos, om, 2 * radius)
    total = om.astype(np.float64).sum(0)
    for l in range(t.depth):
        gl, cp, cw, cs = t.layer(l)
        assert np.allclose(cw.astype(np.float64).sum(0), total, rtol=1e-4, atol=1e-3)
        nz = cs > 0
        # centroids of occupied cells lie in their cells (|w|-weighted mean)
        cells = np.array(np.unravel_index(np.nonzero(nz)[0], gl.shape[::-1])).T[:, ::-1]
        lo = np.array(gl.min) + cells * np.array(gl.cell)
        hi = lo + np.array(gl.cell)
        tol = 1e-5 * np.abs(np.array(gl.extent)).max()
        assert np.all(cp[nz] >= lo - tol) and np.all(cp[nz] <= hi + tol)
        assert np.all(cs[nz] == F32(2 * radius))
    assert t.depth > 2


def test_tree_single_vorton_equals_direct(oracle):
    pos = np.array([[0.31, -0.2, 0.7]], F32)
    om = np.array([[0.2, 0.5, -0.4]], F32)
    g = oracle.define_shape(64, [-1, -1, -1], [1, 1, 1])
    t = oracle.InfluenceTree(g, pos, om, 0.1)
    rng = np.random.default_rng(4)
    for q in rng.uniform(-1, 1, (50, 3)).astype(F32):
        a = t.velocity(q, policy=1)
        b = oracle.direct_velocity(pos, om, 0.1, q[None], policy=1)[0]
        assert np.allclose(a, b, rtol=1e-6, atol=1e-12)


def test_tree_counts_every_vorton_once(oracle, ring_case):
    """With one common Omega every cluster's first moment about its
    |w|-weighted centroid vanishes, so far away the treecode must equal the
    direct sum to second order in (cluster size / distance); a vorton
    dropped or counted twice would show as an O(1/N) error."""
    pos, _, radius, g = ring_case
    om = np.tile(np.array([0.1, -0.3, 0.2], F32), (len(pos), 1))
    t = oracle.InfluenceTree(g, pos, om, 2 * radius)
    q = np.array([[60.0, 10.0, 20.0], [-50.0, 40.0, -30.0], [5.0, 2.0, 80.0]], F32)
    tv = np.array([t.velocity(x, policy=1) for x in q])
    dv = oracle.direct_velocity(pos, om, 2 * radius, q, policy=1)
    assert np.linalg.norm(tv - dv) / np.linalg.norm(dv) < 1e-3


def test_tree_grid_uses_direct_for_single_layer(oracle):
    pos = np.array([[0.5, 0.5, 0.5]], F32)
    om = np.array([[0.0, 0.0, 1.0]], F32)
    g = oracle.define_shape(1, [0, 0, 0], [1, 1, 1])
    t = oracle.InfluenceTree(g, pos, om, 0.2)
    assert t.depth == 1
    a = oracle.velocity_grid(t, g, pos, om, 0.2, technique=1, policy=1)
    b = oracle.velocity_grid(None, g, pos, om, 0.2, technique=0, policy=1)
    assert np.array_equal(a, b)


# ------------------------------------------------------------ grids
def test_splat_partition_of_unity_and_weights(oracle, ring_case):
    pos, om, radius, g = ring_case
    out = oracle.splat_vorticity(g, pos, om, radius, scale=1.0)
    vc = (4.0 / 3.0) * np.pi * radius ** 3 / (g.cell[0] * g.cell[1] * g.cell[2])
    want = (2.0 * om.astype(np.float64)).sum(0) * vc
    assert np.allclose(out.astype(np.float64).sum(0), want, rtol=1e-4)
    # one vorton at fractional cell position (0.25, 0.5, 0.75) of cell (1,2,3)
    frac = np.array([0.25, 0.5, 0.75])
    p = (np.array(g.min) + (np.array([1, 2, 3]) + frac) * np.array(g.cell)).astype(F32)
    w = np.array([[0.0, 0.0, 0.5]], F32)
    one = oracle.splat_vorticity(g, p[None], w, radius, scale=1.0).reshape(g.shape[::-1] + (3,))
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                wt = ((frac[0] if dx else 1 - frac[0]) * (frac[1] if dy else 1 - frac[1])
                      * (frac[2] if dz else 1 - frac[2]))
                assert one[3 + dz, 2 + dy, 1 + dx, 2] == pytest.approx(wt * vc, rel=1e-4)
    assert np.count_nonzero(one[..., 2]) == 8


def _trilinear_field(g, coef):
    pts = oracle_query_points(g)
    x, y, z = pts[:, 0].astype(np.float64), pts[:, 1].astype(np.float64), pts[:, 2].astype(np.float64)
    cols = []
    for c in coef:
        cols.append(c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * y + c[5] * y * z
                    + c[6] * x * z + c[7] * x * y * z)
    return np.stack(cols, 1).astype(F32)


def oracle_query_points(g):
    axes = [np.float32(g.min[a]) + np.arange(g.np[a], dtype=np.float32) * np.float32(g.cell[a])
            for a in range(3)]
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    return np.stack([x, y, z], -1).reshape(-1, 3)


def test_interpolate_exact_on_trilinear(oracle):
    g = oracle.define_shape(4000, [-1, -0.5, 0], [1, 1, 2])
    rng = np.random.default_rng(5)
    coef = rng.normal(size=(3, 8))
    f = _trilinear_field(g, coef)
    for p in rng.uniform([-0.99, -0.49, 0.01], [0.99, 0.99, 1.99], (200, 3)).astype(F32):
        v = oracle.interpolate(g, f, p)
        x, y, z = (float(t) for t in p)
        want = [c[0] + c[1] * x + c[2] * y + c[3] * z + c[4] * x * y + c[5] * y * z + c[6] * x * z
                + c[7] * x * y * z for c in coef]
        assert np.allclose(v, want, atol=2e-5 * (1 + np.abs(want).max()))


def test_restrict_and_prolong(oracle):
    g = oracle.define_shape(4000, [0, 0, 0], [1, 1.5, 2])
    ng = oracle.nested(g)
    lo = ng.layer[1]
    rng = np.random.default_rng(6)
    hi_f = rng.normal(size=(g.capacity, 3)).astype(F32)
    lo_f = oracle.downsample(lo, g, hi_f)
    H = hi_f.reshape(g.shape[::-1] + (3,))
    Lo = lo_f.reshape(lo.shape[::-1] + (3,))
    d = ng.dec[1]
    assert np.array_equal(Lo, H[:: d[2], :: d[1], :: d[0]])
    # accurate (full-weighting) restriction keeps constants on interior points
    const = np.ones((g.capacity, 3), F32)
    acc = oracle.downsample(lo, g, const, accurate=True).reshape(lo.shape[::-1] + (3,))
    assert np.allclose(acc[1:-1, 1:-1, 1:-1], 1.0, atol=1e-6)
    # prolong: linear coarse field reproduced on fine interior, boundary kept
    coef = rng.normal(size=(3, 8))
    coef[:, 4:] = 0.0
    lo_lin = _trilinear_field(lo, coef)
    fine0 = np.full((g.capacity, 3), 7.0, F32)
    up = oracle.upsample(g, fine0, lo, lo_lin).reshape(g.shape[::-1] + (3,))
    want = _trilinear_field(g, coef).reshape(g.shape[::-1] + (3,))
    assert np.allclose(up[1:-1, 1:-1, 1:-1], want[1:-1, 1:-1, 1:-1], atol=1e-4)
    assert np.all(up[0] == 7.0) and np.all(up[:, :, -1] == 7.0)


def _test_func_1d(x):
    return x * (4.0 - x * 4.0)


def test_poisson_manufactured_solution(oracle):
    """testFunc3D = f(x)f(y)f(z), f = x(4-4x) (uniformGridDiagnostics.cpp:
    319-323): f'' = -8, so the 7-point Laplacian is exact and the discrete
    solution with Dirichlet data = testFunc3D is testFunc3D itself."""
    g = oracle.define_shape(2000, [0, 0, 0], [1, 1, 1])
    p = oracle_query_points(g).astype(np.float64)
    fx, fy, fz = (_test_func_1d(p[:, a]) for a in range(3))
    u = fx * fy * fz
    lap = -8.0 * (fy * fz + fx * fz + fx * fy)
    exact = np.repeat(u[:, None], 3, 1).astype(F32)
    rhs = np.repeat(lap[:, None], 3, 1).astype(F32)
    assert oracle.residual_rel(g, exact, rhs) < 1e-5
    start = exact.copy().reshape(g.shape[::-1] + (3,))
    start[1:-1, 1:-1, 1:-1] = 0.0
    sol = oracle.solve_vector_poisson(g, start.reshape(-1, 3), rhs, 300)
    assert np.abs(sol - exact).max() < 2e-5
    # one red pass only touches rows with (iy + iz) even, interior x
    one = oracle.sor_pass(g, start.reshape(-1, 3), rhs, 0).reshape(g.shape[::-1] + (3,))
    changed = np.any(one != start, axis=(2, 3))
    iz, iy = np.nonzero(changed)
    assert np.all((iy + iz) % 2 == 0)
    # the multigrid cascade improves on the cold start
    casc = oracle.mg_cascade(g, start.reshape(-1, 3), rhs, 16)
    assert np.abs(casc - exact).max() < 0.5 * np.abs(start.reshape(-1, 3) - exact).max()
    assert oracle.residual_rel(g, casc, rhs) < oracle.residual_rel(g, start.reshape(-1, 3), rhs)


def test_curl_exact_on_linear_fields(oracle):
    g = oracle.define_shape(3000, [-1, 0, 0.5], [1, 2, 1.5])
    rng = np.random.default_rng(8)
    M = rng.normal(size=(3, 3))
    c0 = rng.normal(size=3)
    p = oracle_query_points(g).astype(np.float64)
    A = (p @ M.T + c0).astype(F32)
    cu = oracle.curl(g, A)
    # A_b = sum_a M[b,a] x_a: d A_b / d a = M[b, a]
    want = np.array([M[2, 1] - M[1, 2], M[0, 2] - M[2, 0], M[1, 0] - M[0, 1]])
    assert np.allclose(cu, want, atol=5e-4)


def test_advection_closed_form(oracle):
    g = oracle.define_shape(1000, [0, 0, 0], [1, 1, 1])
    v0 = np.array([0.3, -0.2, 0.1], F32)
    field = np.tile(v0, (g.capacity, 1))
    rng = np.random.default_rng(9)
    pos = rng.uniform(0.05, 0.95, (500, 3)).astype(F32)
    vel = rng.normal(size=(500, 3)).astype(F32)
    p1, v1 = oracle.advect(g, field, pos, vel, 0.1, gain=1.0)
    assert np.allclose(v1, v0, atol=1e-6)
    assert np.allclose(p1, pos + 0.1 * v0, atol=1e-6)
    p2, v2 = oracle.advect(g, field, pos, vel, 0.1, gain=0.25)
    assert np.allclose(v2, vel + 0.25 * (v0 - vel), atol=1e-6)


# ------------------------------------------------------------ near-field treecode (accurate mode)
def test_mac_covering_everything_is_the_direct_sum(oracle):
    """With a near box larger than the grid no cluster is applied: the walk
    degenerates to the direct sum over every vorton."""
    from vortegrid_amd import scenarios

    pos, om, radius = scenarios.vortex_ring(3000, seed=31)
    g = oracle.grid_for(pos, radius)
    t = oracle.InfluenceTree(g, pos, om, 2 * radius)
    mac = oracle.mac_velocity_grid(t, g, pos, om, 2 * radius, near_k=64, order=1, policy=1, threads=8)
    direct = oracle.velocity_grid(None, g, pos, om, 2 * radius, technique=0, policy=1, threads=8)
    assert np.linalg.norm(mac - direct) / np.linalg.norm(direct) < 2e-6


def test_mac_first_moments_are_shifted_sums(oracle, ring_case):
    """Leaf first moments equal sum Omega (y - c); a parent's moment about
    its own centroid equals the sum over all its vortons (parallel-axis
    theorem), checked against a float64 recomputation from the vortons."""
    pos, om, radius, g = ring_case
    t = oracle.InfluenceTree(g, pos, om, 2 * radius)
    top = t.depth - 2
    gl, cp, cw, cs = t.layer(top)
    D = t.dipoles(top)
    scale = [1, 1, 1]
    for l in range(1, top + 1):
        scale = [scale[a] * t.decimations(l)[a] for a in range(3)]
    g0 = t.layer(0)[0]
    idx = np.floor((pos.astype(np.float64) - np.array(g0.min)) * np.array(g0.cpe)).astype(int)
    cell = idx // np.array(scale)
    key = cell[:, 0] + gl.np[0] * (cell[:, 1] + gl.np[1] * cell[:, 2])
    for k in np.unique(key):
        sel = key == k
        want = np.einsum("nb,ne->be", om[sel].astype(np.float64),
                         pos[sel].astype(np.float64) - cp[k].astype(np.float64))
        assert np.allclose(D[k], want, atol=2e-4 * (np.abs(want).max() + 1e-3))


def test_mac_accuracy_improves_with_order_and_near_radius(oracle):
    from vortegrid_amd import scenarios

    pos, om, radius = scenarios.random_rings(16384, seed=32, rings=16)
    g = oracle.grid_for(pos, radius)
    t = oracle.InfluenceTree(g, pos, om, 2 * radius)
    s, c = oracle.auto_mollify(g, radius)
    q = oracle.query_points(g)
    sel = np.random.default_rng(0).choice(len(q), 1500, replace=False)
    ref = oracle.direct_velocity_f64(pos, om, 2 * radius, q[sel], s, c, threads=8)
    err = {}
    for k in (1, 2):
        for order in (0, 1):
            v = oracle.mac_velocity_grid(t, g, pos, om, 2 * radius, k, order, s, c, 1, threads=8)
            err[k, order] = np.linalg.norm(v[sel] - ref) / np.linalg.norm(ref)
    ref_rule = oracle.velocity_grid(t, g, pos, om, 2 * radius, 1, s, c, 1, threads=8)[sel]
    e_ref = np.linalg.norm(ref_rule - ref) / np.linalg.norm(ref)
    assert err[1, 0] < 0.5 * e_ref
    assert err[1, 1] < 0.5 * err[1, 0] and err[2, 1] < 0.5 * err[2, 0]
    assert err[2, 1] < err[1, 1] and err[2, 0] < err[1, 0]
    assert err[2, 1] < 5e-3


# ------------------------------------------------------------ local expansions
@pytest.fixture(scope="module")
def rings2(oracle):
    from vortegrid_amd import scenarios

    pos, om, radius = scenarios.colliding_rings(16384, seed=2)
    return pos, om, radius, oracle.grid_for(pos, radius)


def _cluster_field(q, cp, w, size, D, order):
    """Monopole + first-moment far field of one cluster, float64."""
    d = np.asarray(q, np.float64) - np.asarray(cp, np.float64)
    r2 = d @ d
    r = np.sqrt(r2)
    s = (2.0 / 3.0) * (0.5 * size) ** 3
    u = s * np.cross(np.asarray(w, np.float64), d) / r ** 3
    if order:
        D = np.asarray(D, np.float64).reshape(3, 3)
        a = np.array([D[1, 2] - D[2, 1], D[2, 0] - D[0, 2], D[0, 1] - D[1, 0]])
        u -= s * (a - 3.0 / r2 * np.cross(D @ d, d)) / r ** 3
    return u


@pytest.mark.parametrize("order", [0, 1])
def test_local_expansion_matches_derivatives(oracle, order):
    """Value, gradient and (monopole) Hessian of the expansion equal the
    cluster field and its central differences at the expansion centre."""
    rng = np.random.default_rng(order)
    for _ in range(10):
        cp = rng.uniform(-1, 1, 3).astype(np.float32)
        cb = (cp + rng.uniform(3, 6) * rng.standard_normal(3) / np.sqrt(3)).astype(np.float32)
        w = rng.standard_normal(3).astype(np.float32)
        D = (0.3 * rng.standard_normal(9)).astype(np.float32)
        size = 0.2
        c0, c1, c2 = oracle.local_expansion(cb, cp, w, size, D, order)
        f = lambda x, o=order: _cluster_field(x, cp, w, size, D, o)
        x0 = cb.astype(np.float64)
        assert np.allclose(c0, f(x0), rtol=1e-6, atol=1e-12)
        h = 1e-4
        E = np.eye(3)
        for j in range(3):
            g = (f(x0 + h * E[j]) - f(x0 - h * E[j])) / (2 * h)
            assert np.allclose(c1[:, j], g, rtol=1e-5, atol=1e-9 * np.abs(c1).max())
            for k in range(3):
                fm = lambda x: _cluster_field(x, cp, w, size, D, 0)
                hess = (fm(x0 + h * (E[j] + E[k])) - fm(x0 + h * (E[j] - E[k])) -
                        fm(x0 - h * (E[j] - E[k])) + fm(x0 - h * (E[j] + E[k]))) / (4 * h * h)
                assert np.allclose(c2[:, j, k], hess, rtol=2e-4, atol=1e-6 * np.abs(c2).max())


def test_mac_expansion_small_change(oracle, rings2):
    """Expanding the upper layers changes the near-field treecode little,
    and leaves its error against the FP64 direct sum nearly unchanged;
    expand_from out of range is the plain walk bit for bit."""
    pos, om, radius, g = rings2
    t = oracle.InfluenceTree(g, pos, om, 2 * radius)
    spread, circ = oracle.auto_mollify(g, radius)
    base = oracle.mac_velocity_grid(t, g, pos, om, 2 * radius, 1, 1, spread, circ, 1, threads=8)
    off = oracle.mac_velocity_grid_ex(t, g, pos, om, 2 * radius, 1, 1, 99, spread, circ, 1,
                                      threads=8)
    assert np.array_equal(base, off)
    ex = oracle.mac_velocity_grid_ex(t, g, pos, om, 2 * radius, 1, 1, 2, spread, circ, 1,
                                     threads=8)
    rel = np.linalg.norm(ex - base) / np.linalg.norm(base)
    assert 0 < rel < 5e-3
    rng = np.random.default_rng(0)
    sel = rng.choice(g.capacity, 300, replace=False)
    q = oracle.query_points(g).reshape(-1, 3)[sel]
    ref = oracle.direct_velocity_f64(pos, om, 2 * radius, q, spread, circ, threads=8)
    e_base = np.linalg.norm(base[sel] - ref) / np.linalg.norm(ref)
    e_ex = np.linalg.norm(ex[sel] - ref) / np.linalg.norm(ref)
    assert e_ex < 1.25 * e_base + 1e-4


def _dyadic_grid(oracle):
    """A grid whose point coordinates and spacings are exact binary fractions."""
    g = oracle.Grid()
    for a, (n, h) in enumerate([(17, 0.125), (9, 0.25), (33, 0.125)]):
        g.min[a] = -0.5 * a
        g.cell[a] = h
        g.cpe[a] = 1.0 / h
        g.np[a] = n
        g.extent[a] = (n - 1) * h
    return g


def test_jacobian_exact_on_linear_fields(oracle):
    """ComputeJacobian (uniformGridMath.cpp:304-451): v = M x + c has
    d v_b / d a = M[b, a] everywhere, faces (one-sided) included; with
    dyadic M, c and coordinates every difference is exact."""
    g = _dyadic_grid(oracle)
    rng = np.random.default_rng(21)
    M = rng.integers(-8, 9, (3, 3)) / 4.0
    c0 = rng.integers(-8, 9, 3) / 8.0
    p = oracle_query_points(g).astype(np.float64)
    v = (p @ M.T + c0).astype(F32)
    J = oracle.jacobian(g, v)
    assert np.array_equal(J, np.broadcast_to(M.T.astype(F32), J.shape))


def test_jacobian_one_sided_on_faces_central_inside(oracle):
    """v_b = x_a^2: the central difference is exact inside (2 x_a) and the
    one-sided face differences give 2 x_a + h at index 0 and 2 x_a - h at
    n-1, so using 1/(2h) or a central stencil on a face fails this."""
    g = _dyadic_grid(oracle)
    p = oracle_query_points(g).astype(np.float64)
    idx = np.stack(np.meshgrid(np.arange(g.np[2]), np.arange(g.np[1]), np.arange(g.np[0]),
                               indexing="ij")[::-1], -1).reshape(-1, 3)
    for a in range(3):
        v = np.zeros_like(p)
        v[:, (a + 1) % 3] = p[:, a] ** 2
        J = oracle.jacobian(g, v.astype(F32)).astype(np.float64)
        h = float(g.cell[a])
        want = 2 * p[:, a]
        want = np.where(idx[:, a] == 0, want + h, want)
        want = np.where(idx[:, a] == g.np[a] - 1, want - h, want)
        assert np.array_equal(J[:, a, (a + 1) % 3], want)
        others = [(aa, bb) for aa in range(3) for bb in range(3) if (aa, bb) != (a, (a + 1) % 3)]
        for aa, bb in others:
            assert not np.any(J[:, aa, bb])


def test_jacobian_agrees_with_curl(oracle):
    """The curl kernel's internal Jacobian (ComputeCurlFromJacobian,
    uniformGridMath.cpp:207-231) and ComputeJacobian are the same stencil."""
    g = oracle.define_shape(3000, [-1, 0, 0.5], [1, 2, 1.5])
    rng = np.random.default_rng(22)
    v = rng.normal(size=(g.capacity, 3)).astype(F32)
    J = oracle.jacobian(g, v)
    cu = np.stack([J[:, 1, 2] - J[:, 2, 1], J[:, 2, 0] - J[:, 0, 2], J[:, 0, 1] - J[:, 1, 0]], 1)
    assert np.array_equal(cu, oracle.curl(g, v))


def test_stretch_tilt_closed_form(oracle):
    """v = M x: J is the constant M^T, so the transpose formulation
    (vortonSim.cpp:2800-2801, s = Omega * J, mat33.h:141-145) gives
    Omega' = Omega + 0.5 dt M^T Omega.  Dyadic data keep every operation
    exact, and a non-symmetric M makes the classical J Omega fail."""
    g = _dyadic_grid(oracle)
    rng = np.random.default_rng(23)
    M = rng.integers(-8, 9, (3, 3)) / 4.0
    assert not np.array_equal(M, M.T)
    p = oracle_query_points(g).astype(np.float64)
    v = (p @ M.T).astype(F32)
    J = oracle.jacobian(g, v)
    # positions at quarter-cell offsets: trilinear weights are dyadic
    n = 300
    cells = np.stack([rng.integers(0, g.np[a] - 1, n) for a in range(3)], 1)
    frac = rng.integers(0, 4, (n, 3)) / 4.0
    pos = (np.float64(g.min) + (cells + frac) * np.float64(g.cell)).astype(F32)
    om = (rng.integers(-16, 17, (n, 3)) / 16.0).astype(F32)
    dt = 0.125
    w = oracle.stretch_tilt(g, J, pos, om, dt)
    want = om.astype(np.float64) + 0.5 * dt * (om.astype(np.float64) @ M)
    assert np.array_equal(w.astype(np.float64), want)
    classical = om.astype(np.float64) + 0.5 * dt * (om.astype(np.float64) @ M.T)
    assert not np.array_equal(w.astype(np.float64), classical)


def test_stretch_tilt_skips_flat_domains(oracle):
    """StretchAndTiltVortons returns early when any grid extent is zero
    (vortonSim.cpp:2754-2759)."""
    g = _dyadic_grid(oracle)
    g.extent[2] = 0.0
    rng = np.random.default_rng(24)
    J = rng.normal(size=(g.capacity, 3, 3)).astype(F32)
    pos = rng.uniform(0.1, 0.9, (50, 3)).astype(F32)
    om = rng.normal(size=(50, 3)).astype(F32)
    assert np.array_equal(oracle.stretch_tilt(g, J, pos, om, 0.1), om)


# ------------------------------------------------------------ PSE diffusion
def test_pse_exp_against_libm(oracle):
    """The PSE distance law's exp (Cody-Waite + degree-7 Taylor) stays within
    2 ulp of the C library's expf over its whole argument range [-87, 0]."""
    import ctypes

    libm = ctypes.CDLL("libm.so.6")
    libm.expf.restype = ctypes.c_float
    libm.expf.argtypes = [ctypes.c_float]
    xs = np.concatenate([np.linspace(-87.0, 0.0, 20001, dtype=F32),
                         -np.geomspace(1e-8, 1.0, 2001).astype(F32), [F32(0.0)]])
    worst = 0
    for x in xs:
        a = np.float32(oracle.pse_exp(float(x)))
        b = np.float32(libm.expf(float(x)))
        ulp = abs(int(a.view(np.int32)) - int(b.view(np.int32)))
        worst = max(worst, ulp)
    assert worst <= 2
    assert oracle.pse_exp(0.0) == 1.0
    assert oracle.pse_exp(-100.0) == 0.0


def _pse_grid(oracle):
    g = oracle.Grid()
    for a in range(3):
        g.min[a], g.cell[a], g.cpe[a], g.np[a] = 0.0, 0.25, 4.0, 9
        g.extent[a] = 2.0
    return g


def _cell_center(ix, iy, iz, frac=(0.5, 0.5, 0.5)):
    return [0.25 * (ix + frac[0]), 0.25 * (iy + frac[1]), 0.25 * (iz + frac[2])]


def _law(oracle, p, q, size, nu, dt):
    d2 = float(np.sum((np.float64(p) - np.float64(q)) ** 2))
    return min(0.5, 2 * nu * dt * np.exp(-d2 / (4 * size) ** 2))


def test_pse_equal_vorticity_only_dissipates(oracle):
    """Equal Omega within one z slice: every vorton meets only partners not
    yet dissipated, every exchange is exactly zero, so each ends at
    Omega - Omega nu dt (vortonSim.cpp:3800)."""
    g = _pse_grid(oracle)
    rng = np.random.default_rng(40)
    pos = rng.uniform(0.01, 1.99, (200, 3)).astype(F32)
    pos[:, 2] = rng.uniform(0.51, 0.74, 200).astype(F32)   # cell slice z = 2
    om = np.tile(np.float32([0.3, -0.7, 1.1]), (200, 1))
    nu, dt = 0.2, 0.1
    out = oracle.pse_diffuse(g, pos, om, 0.2, nu, dt)
    f = np.float32(nu) * np.float32(dt)
    assert np.array_equal(out, om - om * f)


def test_pse_neighbour_set_is_the_forward_half(oracle):
    """Vortons exchange with their own cell and the 7 forward cells only
    (vortonSim.cpp:3711-3728): a pair across the (-1,+1,0) diagonal or two
    cells apart never exchanges; a pair across (+1,0,0) exchanges once."""
    g = _pse_grid(oracle)
    nu, dt, size = 0.5, 0.2, 0.2
    diss = np.float32(nu) * np.float32(dt)
    a, b = np.float32([1, 0, 0]), np.float32([0, 1, 0])
    for (c1, c2, exchanges) in [((3, 3, 3), (2, 4, 3), False), ((3, 3, 3), (5, 3, 3), False),
                                ((3, 3, 3), (4, 3, 3), True), ((3, 3, 2), (4, 4, 3), True)]:
        pos = np.float32([_cell_center(*c1), _cell_center(*c2)])
        out = oracle.pse_diffuse(g, pos, np.stack([a, b]), size, nu, dt).astype(np.float64)
        if not exchanges:
            want = np.stack([a - a * diss, b - b * diss]).astype(np.float64)
            assert np.array_equal(out, want), (c1, c2)
        else:
            # the vorton in the earlier cell (an even or the same slice)
            # exchanges once, then both dissipate
            first, second = (0, 1)
            f = _law(oracle, pos[first], pos[second], size, nu, dt)
            wa = a - f * (a - b)
            wb = b + f * (a - b)
            want = np.stack([wa * (1 - nu * dt), wb * (1 - nu * dt)])
            assert np.allclose(out, want, rtol=2e-6, atol=1e-7), (c1, c2)
            once_more = np.stack([wa - f * (wa - wb), wb + f * (wa - wb)]) * (1 - nu * dt)
            assert not np.allclose(out, once_more, rtol=2e-6, atol=1e-7)


def test_pse_even_slices_before_odd(oracle):
    """The threaded reference runs all even z slices, then all odd ones
    (vortonSim.cpp:3855-3860).  B in slice 2 is processed (and dissipates)
    before A in slice 1 exchanges with it; the serial order would exchange
    first and dissipate after."""
    g = _pse_grid(oracle)
    nu, dt, size = 0.5, 0.2, 0.2
    d = nu * dt
    A, B = np.float64([1, 0, 0]), np.float64([0, 0, 1])
    pos = np.float32([_cell_center(3, 3, 1), _cell_center(3, 3, 2)])
    out = oracle.pse_diffuse(g, pos, np.float32([A, B]), size, nu, dt).astype(np.float64)
    f = _law(oracle, pos[0], pos[1], size, nu, dt)
    b1 = B * (1 - d)                      # even phase: B alone, dissipates
    a2 = A - f * (A - b1)                 # odd phase: A exchanges with B
    b2 = b1 + f * (A - b1)
    want = np.stack([a2 * (1 - d), b2])
    assert np.allclose(out, want, rtol=2e-6, atol=1e-7)
    serial_b = (B + f * (A - B)) * (1 - d)
    assert not np.allclose(out[1], serial_b, rtol=2e-6, atol=1e-7)


def test_pse_clamp_dead_vortons_and_cell_order(oracle):
    """The exchange fraction clamps at 0.5 (vortonSim.cpp:3651); a dead
    vorton initiates nothing and does not dissipate but still receives
    exchanges; within a cell the lower index goes first."""
    g = _pse_grid(oracle)
    nu, dt, size = 5.0, 0.2, 0.2     # 2 nu dt = 2: clamped to 0.5 at short range
    A, B, C = np.float64([1, 0, 0]), np.float64([0, 1, 0]), np.float64([0, 0, 1])
    pos = np.float32([_cell_center(2, 2, 2), _cell_center(2, 2, 2, (0.6, 0.5, 0.5)),
                      _cell_center(6, 6, 6)])
    alive = np.array([1, 0, 1], bool)
    out = oracle.pse_diffuse(g, pos, np.float32([A, B, C]), size, 0.5 / (2 * dt) * 2, dt,
                             alive=alive).astype(np.float64)
    nu = 0.5 / (2 * dt) * 2
    d = np.float64(np.float32(nu) * np.float32(dt))
    f = _law(oracle, pos[0], pos[1], size, nu, dt)
    assert f == 0.5
    a1, b1 = A - f * (A - B), B + f * (A - B)     # A (alive, lower index) exchanges with B
    want = np.stack([a1 * (1 - d), b1, C * (1 - d)])   # dead B neither initiates nor dissipates
    assert np.allclose(out, want, rtol=2e-6, atol=1e-7)


def test_pse_simultaneous_closed_forms(oracle):
    """Simultaneous PSE: every exchange from the starting values.  Equal
    Omega anywhere is pure dissipation; the cross-slice pair that the
    threaded order treats in two phases gets both exchange and dissipation
    from the starting values; with every vorton alive the exchanges cancel in
    the sum, so sum(Omega' / (1 - nu dt)) = sum(Omega)."""
    g = _pse_grid(oracle)
    nu, dt, size = 0.5, 0.2, 0.2
    d = np.float32(nu) * np.float32(dt)
    rng = np.random.default_rng(41)
    pos = rng.uniform(0.01, 1.99, (300, 3)).astype(F32)
    om = np.tile(np.float32([0.3, -0.7, 1.1]), (300, 1))
    out = oracle.pse_diffuse(g, pos, om, 0.2, nu, dt, simultaneous=True)
    assert np.array_equal(out, om - om * d)
    A, B = np.float64([1, 0, 0]), np.float64([0, 0, 1])
    p2 = np.float32([_cell_center(3, 3, 1), _cell_center(3, 3, 2)])
    out2 = oracle.pse_diffuse(g, p2, np.float32([A, B]), size, nu, dt,
                              simultaneous=True).astype(np.float64)
    f = _law(oracle, p2[0], p2[1], size, nu, dt)
    want = np.stack([A - f * (A - B), B + f * (A - B)]) * (1 - nu * dt)
    assert np.allclose(out2, want, rtol=2e-6, atol=1e-7)
    om3 = rng.normal(size=(300, 3)).astype(F32)
    out3 = oracle.pse_diffuse(g, pos, om3, 0.2, 0.05, 0.1, simultaneous=True).astype(np.float64)
    dd = np.float64(np.float32(0.05) * np.float32(0.1))
    assert np.allclose((out3 / (1 - dd)).sum(0), om3.astype(np.float64).sum(0), atol=2e-5)

"""GPU parity: every HIP operator against the CPU oracle on the same seeded
inputs, through the C-ABI.  Bar: bit-exact for index/grid work (binning,
tree layers, restrict/prolong, SOR, curl, advection, AABB); rel-L2 <= 1e-5
for the FP32 pair sums; <= 1e-12 for the FP64 direct sum."""
import numpy as np
import pytest

from vortegrid_amd import scenarios

pytestmark = pytest.mark.gpu
F32 = np.float32


def case(msg):
    print(f"[case] {msg}", flush=True)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def upload(vg, a):
    return vg.DeviceArray.from_numpy(np.ascontiguousarray(a, F32))


def vortons(vg, pos, om, radius):
    dp, dw = upload(vg, pos), upload(vg, om)
    v = vg.vortons_view(dp, dw, len(pos), uniform_size=2 * radius)
    return v, (dp, dw)


@pytest.fixture(scope="module")
def ring(oracle):
    pos, om, radius = scenarios.vortex_ring(4096, seed=1)
    g = oracle.grid_for(pos, radius)
    return pos, om, radius, g


# ------------------------------------------------------------ direct
@pytest.mark.parametrize("policy", [0, 1])
def test_direct_velocity_ring(vg, oracle, ring, policy):
    pos, om, radius, og = ring
    case(f"direct velocity ring N=4096 policy={policy}")
    v, keep = vortons(vg, pos, om, radius)
    g = vg.GridDesc.from_fields(og)
    out = vg.DeviceArray.zeros((len(pos), 3))
    vg.direct_velocity(v, keep[0], len(pos), out, policy=policy)
    gout = vg.DeviceArray.zeros((g.capacity, 3))
    vg.direct_velocity_grid(v, g, gout, policy=policy)
    vg.synchronize()
    want = oracle.direct_velocity(pos, om, 2 * radius, pos, policy=policy, threads=8)
    assert rel_l2(out.numpy(), want) < 1e-5
    wgrid = oracle.velocity_grid(None, og, pos, om, 2 * radius, technique=0, policy=policy,
                                 threads=8)
    assert rel_l2(gout.numpy(), wgrid) < 1e-5


def test_direct_vecpot_and_f64(vg, oracle, ring):
    pos, om, radius, _ = ring
    case("direct vecpot + fp64 ring N=4096")
    v, keep = vortons(vg, pos, om, radius)
    q = pos[::7].copy()
    dq = upload(vg, q)
    out = vg.DeviceArray.zeros((len(q), 3))
    vg.direct_vecpot(v, dq, len(q), out, policy=0)
    o64 = vg.DeviceArray.zeros((len(q), 3), np.float64)
    vg.direct_velocity_f64(v, dq, len(q), o64, spread=1.3, circ=0.5)
    vg.synchronize()
    assert rel_l2(out.numpy(), oracle.direct_vecpot(pos, om, 2 * radius, q, policy=0)) < 1e-5
    w64 = oracle.direct_velocity_f64(pos, om, 2 * radius, q, spread=1.3, circ=0.5, threads=8)
    assert rel_l2(o64.numpy(), w64) < 1e-12


def test_direct_strided_particle_records(vg, oracle, ring):
    """72-byte Particle records (stride 18 floats): pos at 0, angvel at 9."""
    pos, om, radius, _ = ring
    case("direct velocity on Particle AoS stride 18")
    n = 1000
    rec = np.zeros((n, 18), F32)
    rec[:, 0:3] = pos[:n]
    rec[:, 9:12] = om[:n]
    rec[:, 13] = 2 * radius
    d = upload(vg, rec)
    v = vg.Vortons()
    v.pos, v.angvel, v.size = d.ptr.value, d.offset(9).value, d.offset(13).value
    v.stride, v.n, v.uniform_size = 18, n, 0.0
    out = vg.DeviceArray.zeros((n, 3))
    vg.direct_velocity(v, d, n, out, q_stride=18, policy=1)
    vg.synchronize()
    want = oracle.direct_velocity(pos[:n], om[:n], 2 * radius, pos[:n], policy=1)
    assert rel_l2(out.numpy(), want) < 1e-5


# ------------------------------------------------------------ tree
@pytest.fixture(scope="module")
def rings2(oracle):
    pos, om, radius = scenarios.colliding_rings(16384, seed=2)
    g = oracle.grid_for(pos, radius)
    return pos, om, radius, g


def test_tree_layers_bit_exact(vg, oracle, rings2):
    pos, om, radius, og = rings2
    case(f"tree build colliding rings N=16384 grid={og.shape}")
    v, keep = vortons(vg, pos, om, radius)
    t = vg.Tree()
    t.build(vg.GridDesc.from_fields(og), v)
    vg.synchronize()
    ot = oracle.InfluenceTree(og, pos, om, 2 * radius)
    assert t.depth == ot.depth
    for l in range(ot.depth):
        g1, p1, w1, s1 = t.layer(l)
        g2, p2, w2, s2 = ot.layer(l)
        assert g1.as_dict() == {k: [float(x) for x in getattr(g2, k)] for k in
                                ("min", "extent", "cell", "cpe")} | {"np": list(g2.np)}
        assert np.array_equal(s1, s2), l
        assert np.array_equal(w1, w2), l
        assert np.array_equal(p1, p2), l


@pytest.mark.parametrize("policy", [0, 1])
def test_tree_velocity_grid(vg, oracle, rings2, policy):
    pos, om, radius, og = rings2
    case(f"tree velocity grid N=16384 policy={policy}")
    v, keep = vortons(vg, pos, om, radius)
    g = vg.GridDesc.from_fields(og)
    t = vg.Tree()
    t.build(g, v)
    spread, circ = vg.auto_mollify(g, radius)
    out = vg.DeviceArray.zeros((g.capacity, 3))
    t.velocity_grid(g, out, spread, circ, policy)
    vg.synchronize()
    ot = oracle.InfluenceTree(og, pos, om, 2 * radius)
    want = oracle.velocity_grid(ot, og, pos, om, 2 * radius, technique=1, spread=spread,
                                circ=circ, policy=policy, threads=8)
    assert rel_l2(out.numpy(), want) < 1e-5


def test_tree_velocity_and_vecpot_at_points(vg, oracle, rings2):
    pos, om, radius, og = rings2
    case("tree velocity/vecpot at vorton positions")
    v, keep = vortons(vg, pos, om, radius)
    t = vg.Tree()
    t.build(vg.GridDesc.from_fields(og), v)
    q = pos[::13].copy()
    dq = upload(vg, q)
    vel = vg.DeviceArray.zeros((len(q), 3))
    vp = vg.DeviceArray.zeros((len(q), 3))
    t.velocity(dq, len(q), vel, 1.0, 1.0, 0)
    t.vecpot(dq, len(q), vp, 1.0, 1.0, 0)
    vg.synchronize()
    ot = oracle.InfluenceTree(og, pos, om, 2 * radius)
    wv = np.array([ot.velocity(x, 1.0, 1.0, 0) for x in q])
    wp = np.array([ot.vecpot(x, 1.0, 1.0, 0) for x in q])
    assert rel_l2(vel.numpy(), wv) < 1e-5
    assert rel_l2(vp.numpy(), wp) < 1e-5


@pytest.fixture(scope="module")
def slab(oracle):
    """6000 vortons in a slab 2 x 2 x 0.08: a 64 x 64 x 4-cell tree whose
    layers 1-2 halve every axis and layers 3-6 keep z (decimation 2, 2, 1)."""
    rng = np.random.default_rng(41)
    n = 6000
    pos = np.column_stack([rng.uniform(-1, 1, n), rng.uniform(-1, 1, n),
                           rng.uniform(-0.04, 0.04, n)]).astype(F32)
    om = rng.normal(size=(n, 3)).astype(F32)
    radius = 0.02
    og = oracle.grid_for(pos, radius)
    ot = oracle.InfluenceTree(og, pos, om, 2 * radius)
    assert [tuple(ot.decimations(l)) for l in range(1, ot.depth)] == \
        [(2, 2, 2)] * 2 + [(2, 2, 1)] * (ot.depth - 3)
    return pos, om, radius, og, ot


@pytest.mark.parametrize("policy", [0, 1])
def test_tree_walk_mixed_decimation_and_outside_queries(vg, oracle, slab, policy):
    """The walk visits the non-empty children of a layer through the parents'
    child masks where the layer halves every axis and enumerates the children
    elsewhere: grids, faces and point queries on a tree with both kinds of
    layer, with query grids that cover, undershoot and overreach the tree."""
    pos, om, radius, og, ot = slab
    case(f"tree walk on a slab tree {list(og.np)}, mixed decimation, policy={policy}")
    v, keep = vortons(vg, pos, om, radius)
    t = vg.Tree()
    t.build(vg.GridDesc.from_fields(og), v)
    spread, circ = vg.auto_mollify(vg.GridDesc.from_fields(og), radius)
    mn, mx = np.float32(og.min), np.float32(og.min) + np.float32(og.extent)
    wide = oracle.define_shape(len(pos), mn - (mx - mn), mx + (mx - mn))
    coarse = oracle.define_shape(64, mn, mx)
    for qg in (og, wide, coarse):
        g = vg.GridDesc.from_fields(qg)
        out = vg.DeviceArray.zeros((g.capacity, 3))
        t.velocity_grid(g, out, spread, circ, policy)
        vg.synchronize()
        want = oracle.velocity_grid(ot, qg, pos, om, 2 * radius, technique=1, spread=spread,
                                    circ=circ, policy=policy, threads=8)
        assert rel_l2(out.numpy(), want) < 1e-5, list(qg.np)
    g = vg.GridDesc.from_fields(og)
    for copy_odd in (True, False):
        out = vg.DeviceArray.zeros((g.capacity, 3))
        t.vecpot_boundary(g, out, 1.0, 1.0, policy, copy_odd)
        vg.synchronize()
        want = oracle.vecpot_boundary(ot, og, pos, om, 2 * radius, 1.0, 1.0, policy, copy_odd,
                                      threads=8)
        assert rel_l2(out.numpy(), want) < 1e-5, copy_odd
    rng = np.random.default_rng(43)
    q = np.concatenate([pos[::29], rng.uniform(mn - 0.5, mx + 0.5, (300, 3))]).astype(F32)
    dq = upload(vg, q)
    vel = vg.DeviceArray.zeros((len(q), 3))
    vp = vg.DeviceArray.zeros((len(q), 3))
    t.velocity(dq, len(q), vel, spread, circ, policy)
    t.vecpot(dq, len(q), vp, spread, circ, policy)
    vg.synchronize()
    wv = np.array([ot.velocity(x, spread, circ, policy) for x in q])
    wp = np.array([ot.vecpot(x, spread, circ, policy) for x in q])
    assert rel_l2(vel.numpy(), wv) < 1e-5
    assert rel_l2(vp.numpy(), wp) < 1e-5


@pytest.mark.parametrize("copy_odd", [True, False])
def test_boundary_vecpot(vg, oracle, rings2, copy_odd):
    pos, om, radius, og = rings2
    case(f"boundary vecpot copy_odd={copy_odd}")
    v, keep = vortons(vg, pos, om, radius)
    g = vg.GridDesc.from_fields(og)
    t = vg.Tree()
    t.build(g, v)
    out = vg.DeviceArray.zeros((g.capacity, 3))
    t.vecpot_boundary(g, out, 1.0, 1.0, 0, copy_odd)
    vg.synchronize()
    ot = oracle.InfluenceTree(og, pos, om, 2 * radius)
    want = oracle.vecpot_boundary(ot, og, pos, om, 2 * radius, 1.0, 1.0, 0, copy_odd, threads=8)
    got = out.numpy()
    assert rel_l2(got, want) < 1e-5
    # interior untouched
    G = got.reshape(og.shape[::-1] + (3,))
    assert np.all(G[1:-1, 1:-1, 1:-1] == 0)


def test_splats(vg, oracle, rings2):
    pos, om, radius, og = rings2
    case("vorticity splat + velocity populate")
    v, keep = vortons(vg, pos, om, radius)
    g = vg.GridDesc.from_fields(og)
    t = vg.Tree()
    t.build(g, v)
    out = vg.DeviceArray.zeros((g.capacity, 3))
    t.splat_vorticity(radius, -1.0, out)
    rng = np.random.default_rng(12)
    vel = rng.normal(size=pos.shape).astype(F32)
    dvel = upload(vg, vel)
    pv = vg.DeviceArray.zeros((g.capacity, 3))
    t.populate_velocity_grid(dvel, pv)
    vg.synchronize()
    assert rel_l2(out.numpy(), oracle.splat_vorticity(og, pos, om, radius, -1.0)) < 1e-6
    assert rel_l2(pv.numpy(), oracle.populate_velocity_grid(og, pos, vel)) < 1e-6


# ------------------------------------------------------------ grid operators
@pytest.fixture(scope="module")
def mg_case(oracle):
    pos, om, radius = scenarios.jet_plume(20000, seed=3)
    g = oracle.grid_for(pos, radius)
    negw = oracle.splat_vorticity(g, pos, om, radius, -1.0)
    return pos, om, radius, g, negw


def test_restrict_prolong_sor_curl_bit_exact(vg, oracle, mg_case):
    pos, om, radius, og, negw = mg_case
    case(f"restrict/prolong/sor/curl grid={og.shape}")
    g = vg.GridDesc.from_fields(og)
    ng = oracle.nested(og)
    lo = ng.layer[1]
    rng = np.random.default_rng(13)
    f = rng.normal(size=(og.capacity, 3)).astype(F32)
    df = upload(vg, f)
    dlo = vg.DeviceArray.zeros((lo.capacity, 3))
    glo = vg.GridDesc.from_fields(lo)
    vg.downsample_inject(glo, dlo, g, df)
    up = upload(vg, f)
    vg.upsample_interior(g, up, glo, dlo)
    sol = upload(vg, f)
    dnegw = upload(vg, negw)
    vg.sor_solve(g, sol, dnegw, 3)
    cu = vg.DeviceArray.zeros((og.capacity, 3))
    vg.curl(g, df, cu)
    vg.synchronize()
    wlo = oracle.downsample(lo, og, f)
    assert np.array_equal(dlo.numpy(), wlo)
    assert np.array_equal(up.numpy(), oracle.upsample(og, f, lo, wlo))
    assert np.array_equal(sol.numpy(), oracle.solve_vector_poisson(og, f, negw, 3))
    assert np.array_equal(cu.numpy(), oracle.curl(og, f))


@pytest.mark.parametrize("shape", [(9, 129, 131), (12, 33, 20)])
def test_sor_row_kernels_bit_exact(vg, oracle, shape):
    """The SOR colour pass has two kernels, chosen by the number of rows per
    colour: three components per lane below 8192 rows, one lane per
    component from there (9x129x131 has 65 x 129 = 8385 rows per colour,
    12x33x20 has 17 x 18 = 306).  Both equal the oracle's sequential
    SolveVectorPoisson bit for bit."""
    case(f"sor row kernels, grid {shape}")
    rng = np.random.default_rng(15)
    og = oracle.define_shape(int(np.prod([c - 1 for c in shape])), np.zeros(3),
                             np.array([c - 1 for c in shape], np.float64) * 0.05, pow2=False)
    assert og.shape == shape
    g = vg.GridDesc.from_fields(og)
    f = rng.normal(size=(og.capacity, 3)).astype(F32)
    u0 = rng.normal(size=(og.capacity, 3)).astype(F32)
    sol = upload(vg, u0)
    vg.sor_solve(g, sol, upload(vg, f), 3)
    vg.synchronize()
    assert np.array_equal(sol.numpy(), oracle.solve_vector_poisson(og, u0, f, 3))


def test_mg_cascade(vg, oracle, mg_case):
    pos, om, radius, og, negw = mg_case
    case(f"mg cascade grid={og.shape}")
    g = vg.GridDesc.from_fields(og)
    rng = np.random.default_rng(14)
    a0 = rng.normal(size=(og.capacity, 3)).astype(F32) * 0.01
    da = upload(vg, a0)
    mg = vg.Multigrid()
    mg.cascade(g, da, upload(vg, negw), 16)
    vg.synchronize()
    start = a0.reshape(og.shape[::-1] + (3,)).copy()
    start[1:-1, 1:-1, 1:-1] = 0
    want = oracle.mg_cascade(og, start.reshape(-1, 3), negw, 16)
    assert np.array_equal(da.numpy(), want)


def test_vcycle_manufactured_solution(vg, oracle):
    case("v-cycle testFunc3D 65^3")
    og = oracle.define_shape(262144, [0, 0, 0], [1, 1, 1])
    axes = [np.float32(og.min[a]) + np.arange(og.np[a], dtype=F32) * np.float32(og.cell[a])
            for a in range(3)]
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    f1 = lambda t: t.astype(np.float64) * (4.0 - 4.0 * t.astype(np.float64))
    fx, fy, fz = f1(x), f1(y), f1(z)
    u = (fx * fy * fz).reshape(-1)
    lap = (-8.0 * (fy * fz + fx * fz + fx * fy)).reshape(-1)
    exact = np.repeat(u[:, None], 3, 1).astype(F32)
    rhs = np.repeat(lap[:, None], 3, 1).astype(F32)
    start = exact.reshape(og.shape[::-1] + (3,)).copy()
    start[1:-1, 1:-1, 1:-1] = 0
    g = vg.GridDesc.from_fields(og)
    da = upload(vg, start.reshape(-1, 3))
    mg = vg.Multigrid()
    rel, cycles = mg.vcycle_solve(g, da, upload(vg, rhs), tol=1e-6, max_cycles=40)
    vg.synchronize()
    assert rel <= 1e-6, (rel, cycles)
    assert cycles <= 25
    assert np.abs(da.numpy() - exact).max() < 1e-4


def test_advection_and_bbox_bit_exact(vg, oracle, rings2):
    pos, om, radius, og = rings2
    case("tracer advection + bbox")
    rng = np.random.default_rng(15)
    field = rng.normal(size=(og.capacity, 3)).astype(F32)
    mn, mx = np.array(og.min), np.array(og.min) + np.array(og.extent)
    tr = scenarios.tracers_in_box(50000, mn + 1e-3, mx - 1e-3, seed=16)
    vel = rng.normal(size=tr.shape).astype(F32)
    g = vg.GridDesc.from_fields(og)
    dfield = upload(vg, field)
    for gain in (1.0, 0.3):
        dp, dv = upload(vg, tr), upload(vg, vel)
        vg.advect_tracers(g, dfield, dp, dv, len(tr), 1.0 / 30.0, gain)
        vg.synchronize()
        wp, wv = oracle.advect(og, field, tr, vel, 1.0 / 30.0, gain)
        assert np.array_equal(dp.numpy(), wp)
        assert np.array_equal(dv.numpy(), wv)
    # velocity not stored: same positions
    dp = upload(vg, tr)
    vg.advect_tracers(g, dfield, dp, None, len(tr), 1.0 / 30.0, 1.0)
    vg.synchronize()
    assert np.array_equal(dp.numpy(), oracle.advect(og, field, tr, vel, 1.0 / 30.0, 1.0)[0])
    # bounding box
    dpos = upload(vg, pos)
    a, b = vg.find_bounding_box(dpos, len(pos))
    assert np.array_equal(a, pos.min(0)) and np.array_equal(b, pos.max(0))
    # vorton box and tracer box in one pass: the same as the two calls
    dtr = upload(vg, tr)
    sim = vg.Sim(vg.sim_config(technique=vg.TECH_TREE))
    v = vg.vortons_view(dpos, upload(vg, om), len(pos), uniform_size=2 * radius)
    sim.set_vortons(v)
    tmn, tmx = sim.find_bounding_boxes(dtr, len(tr))
    assert np.array_equal(tmn, tr.min(0)) and np.array_equal(tmx, tr.max(0))
    smn, smx = sim.bounds()
    assert np.array_equal(smn, pos.min(0)) and np.array_equal(smx, pos.max(0))


def test_out_of_domain_tracers_are_clamped(vg, oracle, rings2):
    _, _, _, og = rings2
    case("out-of-domain tracers")
    field = np.ones((og.capacity, 3), F32)
    tr = np.array([[-100.0, 0.0, 0.0], [100.0, 100.0, 100.0], [np.float32(og.min[0]), 0, 0]], F32)
    g = vg.GridDesc.from_fields(og)
    dp = upload(vg, tr)
    dv = vg.DeviceArray.zeros(tr.shape)
    vg.advect_tracers(g, upload(vg, field), dp, dv, len(tr), 0.1, 1.0)
    vg.synchronize()
    assert np.all(np.isfinite(dp.numpy())) and np.all(np.isfinite(dv.numpy()))


# ------------------------------------------------------------ whole frames
def _frame_oracle(oracle, pos, om, radius, technique):
    og = oracle.grid_for(pos, radius)
    ot = oracle.InfluenceTree(og, pos, om, 2 * radius)
    if technique == 1:
        spread, circ = oracle.auto_mollify(og, radius)
        vel = oracle.velocity_grid(ot, og, pos, om, 2 * radius, 1, spread, circ, 0, threads=8)
    else:
        negw = oracle.splat_vorticity(og, pos, om, radius, -1.0)
        a0 = oracle.vecpot_boundary(ot, og, pos, om, 2 * radius, 1.0, 1.0, 0, True, threads=8)
        a = oracle.mg_cascade(og, a0, negw, 16)
        vel = oracle.curl(og, a)
    vv = np.zeros_like(pos)
    _, vv = oracle.advect(og, vel, pos, vv, 0.0, 1.0)
    return og, vel, vv


@pytest.mark.parametrize("technique", [1, 2])
def test_sim_frame(vg, oracle, technique):
    pos, om, radius = scenarios.colliding_rings(8192, seed=21)
    case(f"sim frame technique={technique} N=8192")
    dp, dw = upload(vg, pos), upload(vg, om)
    sim = vg.Sim(vg.sim_config(technique=technique))
    sim.set_vortons(vg.vortons_view(dp, dw, len(pos), uniform_size=2 * radius))
    sim.find_bounding_box()
    sim.update_bounding_box(pos[0], pos[0], True)
    vel = vg.DeviceArray.zeros(pos.shape)
    sim.update(1.0 / 30.0, 0, vel)
    g, p = sim.velocity_grid()
    got = vg.fetch(p, 3 * g.capacity).reshape(-1, 3)
    og, want, wvel = _frame_oracle(oracle, pos, om, radius, 1 if technique == 1 else 2)
    assert g.as_dict() == {k: [float(x) for x in getattr(og, k)] for k in
                           ("min", "extent", "cell", "cpe")} | {"np": list(og.np)}
    assert rel_l2(got, want) < 1e-5
    assert rel_l2(vel.numpy(), wvel) < 1e-5


def test_sim_direct_frame(vg, oracle):
    pos, om, radius = scenarios.vortex_ring(2048, seed=22)
    case("sim frame technique=direct N=2048")
    dp, dw = upload(vg, pos), upload(vg, om)
    sim = vg.Sim(vg.sim_config(technique=0))
    sim.set_vortons(vg.vortons_view(dp, dw, len(pos), uniform_size=2 * radius))
    sim.find_bounding_box()
    sim.update_bounding_box(pos[0], pos[0], True)
    vel = vg.DeviceArray.zeros(pos.shape)
    sim.update(1.0 / 30.0, 0, vel)
    g, p = sim.velocity_grid()
    got = vg.fetch(p, 3 * g.capacity).reshape(-1, 3)
    og = oracle.grid_for(pos, radius)
    wv = oracle.direct_velocity(pos, om, 2 * radius, pos, policy=0, threads=8)
    assert rel_l2(vel.numpy(), wv) < 1e-5
    assert rel_l2(got, oracle.populate_velocity_grid(og, pos, wv)) < 1e-5


def test_empty_inputs(vg):
    case("empty vortons")
    g = vg.define_shape(100, [0, 0, 0], [1, 1, 1])
    pos = np.zeros((1, 3), F32)
    dp = upload(vg, pos)
    v = vg.vortons_view(dp, dp, 0, uniform_size=0.1)
    out = vg.DeviceArray.from_numpy(np.full((g.capacity, 3), 5.0, F32))
    vg.direct_velocity_grid(v, g, out)
    t = vg.Tree()
    t.build(g, v)
    out2 = vg.DeviceArray.from_numpy(np.full((g.capacity, 3), 5.0, F32))
    t.velocity_grid(g, out2)
    vg.synchronize()
    assert np.all(out.numpy() == 0) and np.all(out2.numpy() == 0)


def test_deterministic_reruns(vg, rings2):
    pos, om, radius, og = rings2
    case("bitwise determinism of tree + direct")
    v, keep = vortons(vg, pos, om, radius)
    g = vg.GridDesc.from_fields(og)
    res = []
    for _ in range(2):
        t = vg.Tree()
        t.build(g, v)
        out = vg.DeviceArray.zeros((g.capacity, 3))
        t.velocity_grid(g, out, 1.0, 1.0, 1)
        d = vg.DeviceArray.zeros((len(pos), 3))
        vg.direct_velocity(v, keep[0], len(pos), d, policy=1)
        vg.synchronize()
        res.append((out.numpy(), d.numpy()))
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])


# ------------------------------------------------------------ near-field treecode (accurate mode)
def test_tree_moments_bit_exact(vg, oracle, rings2):
    pos, om, radius, og = rings2
    case("tree first moments vs oracle")
    v, keep = vortons(vg, pos, om, radius)
    t = vg.Tree()
    t.build(vg.GridDesc.from_fields(og), v)
    vg.synchronize()
    ot = oracle.InfluenceTree(og, pos, om, 2 * radius)
    for l in range(ot.depth):
        assert np.array_equal(t.moments(l), ot.dipoles(l)), l


@pytest.mark.parametrize("k,order,policy", [(1, 0, 1), (1, 1, 1), (2, 1, 1), (2, 1, 0), (0, 1, 1)])
def test_mac_velocity_grid(vg, oracle, rings2, k, order, policy):
    pos, om, radius, og = rings2
    case(f"mac velocity grid k={k} order={order} policy={policy}")
    v, keep = vortons(vg, pos, om, radius)
    g = vg.GridDesc.from_fields(og)
    t = vg.Tree()
    t.build(g, v)
    spread, circ = vg.auto_mollify(g, radius)
    out = vg.DeviceArray.zeros((g.capacity, 3))
    t.velocity_grid_mac(g, out, k, order, spread, circ, policy)
    vg.synchronize()
    ot = oracle.InfluenceTree(og, pos, om, 2 * radius)
    want = oracle.mac_velocity_grid(ot, og, pos, om, 2 * radius, k, order, spread, circ, policy,
                                    threads=8)
    assert rel_l2(out.numpy(), want) < 1e-5


@pytest.mark.parametrize("k,order,expand_from", [(1, 1, 2), (2, 1, 3), (1, 0, 2), (2, 1, 99)])
def test_mac_local_expansions(vg, oracle, rings2, k, order, expand_from):
    pos, om, radius, og = rings2
    case(f"mac velocity grid with local expansions k={k} order={order} from={expand_from}")
    v, keep = vortons(vg, pos, om, radius)
    g = vg.GridDesc.from_fields(og)
    t = vg.Tree()
    t.build(g, v)
    spread, circ = vg.auto_mollify(g, radius)
    out = vg.DeviceArray.zeros((g.capacity, 3))
    t.velocity_grid_mac(g, out, k, order, spread, circ, 1, expand_from=expand_from)
    vg.synchronize()
    ot = oracle.InfluenceTree(og, pos, om, 2 * radius)
    want = oracle.mac_velocity_grid_ex(ot, og, pos, om, 2 * radius, k, order, expand_from, spread,
                                       circ, 1, threads=8)
    assert rel_l2(out.numpy(), want) < 1e-5
    c, p, e = t.mac_work(g, k, expand_from)
    c0, p0, e0 = t.mac_work(g, k, -1)
    assert p == p0 and e0 == 0
    if expand_from <= t.depth - 2:
        assert e > 0 and c < c0


def shaped_grid(oracle, pos, radius, cells):
    """A grid of exactly cells[a] cells per axis over the padded vorton box
    (extents proportional to the cell counts, one spacing h)."""
    mn, mx = oracle.find_bbox(pos)
    mn = mn - 3 * radius
    ext = (mx + 3 * radius) - mn
    h = max(float(ext[a]) / cells[a] for a in range(3))
    og = oracle.define_shape(int(np.prod(cells)), mn, mn + h * np.array(cells, np.float64),
                             pow2=False)
    assert og.shape == tuple(c + 1 for c in cells)
    return og


@pytest.mark.parametrize("cells,k,expand_from", [((48, 24, 12), 2, -1), ((48, 24, 12), 2, 2),
                                                 ((24, 24, 24), 3, 2), ((12, 24, 48), 2, 2)])
def test_mac_velocity_grid_odd_shapes(vg, oracle, rings2, cells, k, expand_from):
    """Grids whose cell counts are not powers of two but leave no cell
    without a parent: the NestedGrid decimation is then 3 or 1 on some axes
    (3 cells under one parent, or a 2-point axis kept).  The walk's
    per-parent child masks must give the same clusters as the oracle's
    children box."""
    pos, om, radius, _ = rings2
    og = shaped_grid(oracle, pos, radius, cells)
    ng = oracle.nested(og)
    decs = {int(ng.dec[l][a]) for l in range(1, ng.depth) for a in range(3)}
    case(f"mac velocity grid, grid {og.shape}, decimations {sorted(decs)}, k={k} from={expand_from}")
    assert 3 in decs and not oracle.nested_has_orphans(og)
    v, keep = vortons(vg, pos, om, radius)
    g = vg.GridDesc.from_fields(og)
    t = vg.Tree()
    t.build(g, v)
    spread, circ = vg.auto_mollify(g, radius)
    out = vg.DeviceArray.zeros((g.capacity, 3))
    t.velocity_grid_mac(g, out, k, 1, spread, circ, 1, expand_from=expand_from)
    vg.synchronize()
    ot = oracle.InfluenceTree(og, pos, om, 2 * radius)
    for l in range(ot.depth):
        g1, p1, w1, s1 = t.layer(l)
        g2, p2, w2, s2 = ot.layer(l)
        assert np.array_equal(p1, p2) and np.array_equal(w1, w2), l
    want = oracle.mac_velocity_grid_ex(ot, og, pos, om, 2 * radius, k, 1, expand_from, spread,
                                       circ, 1, threads=8)
    assert np.isfinite(out.numpy()).all()
    assert rel_l2(out.numpy(), want) < 1e-5


@pytest.mark.parametrize("cells", [(20, 20, 20), (36, 36, 36), (40, 20, 10)])
def test_mac_refuses_grids_with_orphan_cells(vg, oracle, rings2, cells):
    """20 -> 10 -> 5 -> 2 leaves the fifth cell of the 5-cell layer without a
    parent cluster, so no walk from the top would reach its vortons: the
    near-field walk, its work count and the batch refuse such shapes.  The
    reference-rule treecode keeps the reference's own behaviour there."""
    pos, om, radius, _ = rings2
    og = shaped_grid(oracle, pos, radius, cells)
    case(f"mac refuses orphan grid {og.shape}")
    assert oracle.nested_has_orphans(og)
    v, keep = vortons(vg, pos, om, radius)
    g = vg.GridDesc.from_fields(og)
    t = vg.Tree()
    t.build(g, v)
    out = vg.DeviceArray.zeros((g.capacity, 3))
    with pytest.raises(vg.VGError, match="parent"):
        t.velocity_grid_mac(g, out, 2, 1, 1.0, 1.0, 1)
    with pytest.raises(vg.VGError, match="parent"):
        t.mac_work(g, 2, -1)
    with pytest.raises(vg.VGError, match="parent"):
        vg.Batch([100], [0], [0.05], cells=cells[0])
    # the reference rule still runs (it reproduces the reference's layers)
    t.velocity_grid(g, out, 1.0, 1.0, 1)
    vg.synchronize()


@pytest.mark.parametrize("k,expand_from", [(2, -1), (2, 2), (3, 2)])
def test_mac_velocity_grid_mixed_sizes(vg, oracle, rings2, k, expand_from):
    """Vortons of two sizes: clusters take the size of their last non-empty
    child, so the walk meets size changes inside chunks (the general loop)
    as well as chunks of one size (the loop without the size test)."""
    pos, om, radius, og = rings2
    case(f"mac velocity grid, two vorton sizes, k={k} from={expand_from}")
    rng = np.random.default_rng(61)
    size = np.where(rng.random(len(pos)) < 0.5, 2 * radius, 3 * radius).astype(F32)
    dp, dw = upload(vg, pos), upload(vg, om)
    ds = upload(vg, np.repeat(size[:, None], 3, axis=1))
    v = vg.Vortons()
    v.pos, v.angvel, v.size = dp.ptr.value, dw.ptr.value, ds.ptr.value
    v.stride, v.n, v.uniform_size = 3, len(pos), 0.0
    g = vg.GridDesc.from_fields(og)
    t = vg.Tree()
    t.build(g, v)
    spread, circ = vg.auto_mollify(g, radius)
    out = vg.DeviceArray.zeros((g.capacity, 3))
    t.velocity_grid_mac(g, out, k, 1, spread, circ, 1, expand_from=expand_from)
    vg.synchronize()
    ot = oracle.InfluenceTree(og, pos, om, size)
    want = oracle.mac_velocity_grid_ex(ot, og, pos, om, size, k, 1, expand_from, spread, circ, 1,
                                       threads=8)
    assert np.isfinite(out.numpy()).all()
    assert rel_l2(out.numpy(), want) < 1e-5


def test_mac_work_and_accuracy(vg, oracle):
    pos, om, radius = scenarios.random_rings(32768, seed=33, rings=16)
    case("mac work counts + accuracy vs fp64 N=32768")
    og = oracle.grid_for(pos, radius)
    v, keep = vortons(vg, pos, om, radius)
    g = vg.GridDesc.from_fields(og)
    t = vg.Tree()
    t.build(g, v)
    spread, circ = vg.auto_mollify(g, radius)
    c1, p1, _ = t.mac_work(g, 1)
    c2, p2, _ = t.mac_work(g, 2)
    assert c2 > c1 > 0 and p2 > p1 > 0
    out = vg.DeviceArray.zeros((g.capacity, 3))
    t.velocity_grid_mac(g, out, 2, 1, spread, circ, 1)
    q = oracle.query_points(og)
    sel = np.random.default_rng(3).choice(len(q), 2000, replace=False)
    dq = upload(vg, q[sel])
    ref = vg.DeviceArray.zeros((len(sel), 3), np.float64)
    vg.direct_velocity_f64(v, dq, len(sel), ref, spread, circ)
    vg.synchronize()
    err = rel_l2(out.numpy()[sel], ref.numpy())
    assert err < 1e-2, err

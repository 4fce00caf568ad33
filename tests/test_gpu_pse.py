"""Viscous diffusion by particle strength exchange
(DiffuseAndDissipateVorticityPSE, vortonSim.cpp:3616-3922) on the GPU: the
partition (PartitionVortons) and the order-dependent in-place exchange, bit
for bit against the CPU oracle, including dead vortons, the exchange clamp,
72-byte records, 1M vortons, and the simulation entry point."""
import numpy as np
import pytest

from vortegrid_amd import scenarios

pytestmark = pytest.mark.gpu
F32 = np.float32


def case(msg):
    print(f"[case] {msg}", flush=True)


def upload(vg, a):
    return vg.DeviceArray.from_numpy(np.ascontiguousarray(a))


def build_tree(vg, oracle, pos, om, radius):
    og = oracle.grid_for(pos, radius)
    g = vg.GridDesc.from_fields(og)
    dp, dw = upload(vg, pos), upload(vg, om)
    v = vg.vortons_view(dp, dw, len(pos), uniform_size=2 * radius)
    t = vg.Tree()
    t.build(g, v)
    return og, g, t, (dp, dw)


def test_partition_matches_reference_rule(vg, oracle):
    """PartitionParticles: vorton i goes to the cell of its position
    (truncation), cells list vortons in increasing index order."""
    pos, om, radius = scenarios.colliding_rings(8192, seed=51)
    case(f"partition N={len(pos)}")
    og, g, t, keep = build_tree(vg, oracle, pos, om, radius)
    start, count, index = t.partition()
    vg.synchronize()
    idx = [((pos[:, a] - F32(og.min[a])) * F32(og.cpe[a])).astype(np.int64) for a in range(3)]
    key = idx[0] + og.np[0] * (idx[1] + og.np[1] * idx[2])
    want_count = np.bincount(key, minlength=og.capacity).astype(np.uint32)
    assert np.array_equal(count, want_count)
    assert np.array_equal(start[1:], np.cumsum(want_count)[:-1].astype(np.uint32))
    assert np.array_equal(index, np.argsort(key, kind="stable").astype(np.uint32))


@pytest.mark.parametrize("viscosity,dt", [(0.01, 1.0 / 30.0), (0.5, 0.2), (2.0, 0.4)])
def test_pse_bit_exact(vg, oracle, viscosity, dt):
    pos, om, radius = scenarios.colliding_rings(8192, seed=52)
    case(f"PSE N={len(pos)} nu={viscosity} dt={dt}")
    og, g, t, (dp, dw) = build_tree(vg, oracle, pos, om, radius)
    t.pse_diffuse(dw, viscosity, dt)
    vg.synchronize()
    want = oracle.pse_diffuse(og, pos, om, 2 * radius, viscosity, dt)
    assert not np.array_equal(want, om)
    assert np.array_equal(dw.numpy(), want)


@pytest.mark.parametrize("dense", [120, 900])
def test_pse_dense_cells(vg, oracle, dense):
    """Cells holding more than a wave's worth of vortons, and neighbourhoods
    beyond the register-resident partner list (900 in one cell), next to a
    sparse ring: the same bits as the oracle."""
    pos, om, radius = scenarios.vortex_ring(2048, seed=56)
    rng = np.random.default_rng(56)
    centre = pos[0]
    blob = (centre + rng.uniform(-0.01, 0.01, (dense, 3))).astype(F32)
    pos = np.concatenate([pos, blob]).astype(F32)
    om = np.concatenate([om, rng.normal(size=(dense, 3)).astype(F32)])
    case(f"PSE with a dense blob of {dense} vortons")
    og, g, t, (dp, dw) = build_tree(vg, oracle, pos, om, radius)
    start, count, index = t.partition()
    assert count.max() >= dense
    t.pse_diffuse(dw, 0.5, 0.2)
    vg.synchronize()
    want = oracle.pse_diffuse(og, pos, om, 2 * radius, 0.5, 0.2)
    assert np.array_equal(dw.numpy(), want)


def test_pse_in_chain_fractions(vg, oracle, monkeypatch):
    """VG_PSE_FRACTIONS=0 keeps the exchange fractions in the serial chain
    (the path for fraction blocks of 2^31 floats or more): the same bits."""
    pos, om, radius = scenarios.vortex_ring(2048, seed=58)
    rng = np.random.default_rng(58)
    blob = (pos[0] + rng.uniform(-0.01, 0.01, (120, 3))).astype(F32)
    pos = np.concatenate([pos, blob]).astype(F32)
    om = np.concatenate([om, rng.normal(size=(120, 3)).astype(F32)])
    case("PSE with in-chain fractions")
    og, g, t, (dp, dw) = build_tree(vg, oracle, pos, om, radius)
    monkeypatch.setenv("VG_PSE_FRACTIONS", "0")
    t.pse_diffuse(dw, 0.5, 0.2)
    vg.synchronize()
    want = oracle.pse_diffuse(og, pos, om, 2 * radius, 0.5, 0.2)
    assert np.array_equal(dw.numpy(), want)


@pytest.mark.parametrize("viscosity,dt", [(0.01, 1.0 / 30.0), (0.5, 0.2)])
def test_pse_simultaneous(vg, oracle, viscosity, dt):
    """VG_PSE_SIMULTANEOUS: bit-exact against the oracle's simultaneous form;
    at the demo viscosity within 1e-5 rel-L2 of the reference order."""
    pos, om, radius = scenarios.colliding_rings(8192, seed=57)
    case(f"simultaneous PSE N={len(pos)} nu={viscosity} dt={dt}")
    og, g, t, (dp, dw) = build_tree(vg, oracle, pos, om, radius)
    t.pse_diffuse(dw, viscosity, dt, order=1)
    vg.synchronize()
    got = dw.numpy()
    assert np.array_equal(got, oracle.pse_diffuse(og, pos, om, 2 * radius, viscosity, dt,
                                                  simultaneous=True))
    if viscosity == 0.01:
        ref = oracle.pse_diffuse(og, pos, om, 2 * radius, viscosity, dt)
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-5


def test_pse_records_and_dead_vortons(vg, oracle):
    """72-byte records (angular velocity at float 9, birth time as int32 at
    14); vortons with a negative birth time are dead."""
    pos, om, radius = scenarios.vortex_ring(4096, seed=53)
    case("PSE on 72-byte records with dead vortons")
    n = len(pos)
    rng = np.random.default_rng(53)
    rec = rng.normal(size=(n, 18)).astype(F32)
    rec[:, 0:3] = pos
    rec[:, 9:12] = om
    birth = np.where(rng.random(n) < 0.2, -1, 5).astype(np.int32)
    rec[:, 14] = birth.view(F32)
    d = upload(vg, rec)
    og = oracle.grid_for(pos, radius)
    g = vg.GridDesc.from_fields(og)
    v = vg.Vortons()
    v.pos, v.angvel, v.size = d.ptr.value, d.offset(9).value, None
    v.stride, v.n, v.uniform_size = 18, n, 2 * radius
    t = vg.Tree()
    t.build(g, v)
    t.pse_diffuse(d.offset(9), 0.5, 0.2, ang_stride=18, birth=d.offset(14), birth_stride=18)
    vg.synchronize()
    out = d.numpy()
    want = oracle.pse_diffuse(og, pos, om, 2 * radius, 0.5, 0.2, alive=birth >= 0)
    assert np.array_equal(out[:, 9:12], want)
    d2 = upload(vg, rec)   # the simultaneous form on the same records
    t.pse_diffuse(d2.offset(9), 0.5, 0.2, ang_stride=18, birth=d2.offset(14), birth_stride=18,
                  order=1)
    vg.synchronize()
    want2 = oracle.pse_diffuse(og, pos, om, 2 * radius, 0.5, 0.2, alive=birth >= 0,
                               simultaneous=True)
    assert np.array_equal(d2.numpy()[:, 9:12], want2)
    keep = np.ones(18, bool)
    keep[9:12] = False
    assert np.array_equal(out.view(np.int32)[:, keep], rec.view(np.int32)[:, keep])


def test_pse_1m_vortons(vg, oracle):
    pos, om, radius = scenarios.random_rings(1 << 20, seed=54)
    case(f"PSE N={len(pos)} (64 random rings)")
    og, g, t, (dp, dw) = build_tree(vg, oracle, pos, om, radius)
    t.pse_diffuse(dw, 0.01, 1.0 / 30.0)
    vg.synchronize()
    want = oracle.pse_diffuse(og, pos, om, 2 * radius, 0.01, 1.0 / 30.0)
    assert np.array_equal(dw.numpy(), want)
    dw.upload(om)
    t.pse_diffuse(dw, 0.01, 1.0 / 30.0, order=1)
    vg.synchronize()
    got = dw.numpy()
    assert np.array_equal(got, oracle.pse_diffuse(og, pos, om, 2 * radius, 0.01, 1.0 / 30.0,
                                                  simultaneous=True))
    # cells of up to ~70 vortons: hundreds of partners per vorton make the
    # in-place order matter at the 1e-3 level (8k-vorton rings: 1e-5)
    dev = np.linalg.norm(got - want) / np.linalg.norm(want)
    print(f"simultaneous vs reference order rel-L2 {dev:.3e}")
    assert dev < 1e-2


def test_sim_pse_diffuse(vg, oracle):
    """vg_sim_pse_diffuse after vg_sim_update: the partition of the frame."""
    pos, om, radius = scenarios.colliding_rings(8192, seed=55)
    case("sim PSE after the update")
    n = len(pos)
    rec = np.zeros((n, 18), F32)
    rec[:, 0:3] = pos
    rec[:, 9:12] = om
    rec[:, 13] = 2 * radius
    d = upload(vg, rec)
    v = vg.Vortons()
    v.pos, v.angvel, v.size = d.ptr.value, d.offset(9).value, d.offset(13).value
    v.stride, v.n, v.uniform_size = 18, n, 0.0
    sim = vg.Sim(vg.sim_config(technique=1))
    sim.set_vortons(v)
    sim.find_bounding_box()
    sim.update_bounding_box(pos[0], pos[0], True)
    sim.update(1.0 / 30.0, 0)
    sim.pse_diffuse(0.3, 0.1, d.offset(9), 18)
    vg.synchronize()
    gd = sim.grid_template()
    og = oracle.Grid()
    for f in ("min", "extent", "cell", "cpe", "np"):
        for a in range(3):
            getattr(og, f)[a] = getattr(gd, f)[a]
    want = oracle.pse_diffuse(og, pos, om, 2 * radius, 0.3, 0.1)
    assert np.array_equal(d.numpy()[:, 9:12], want)

"""Vortex stretching and tilting (vortonSim.cpp:2748-2820) on the GPU: the
Jacobian grid, the per-vorton update from it and the fused update straight
from the velocity grid, all bit-exact against the CPU oracle and against
each other; the 2-D early return; the simulation entry point on 72-byte
Particle records."""
import numpy as np
import pytest

from vortegrid_amd import scenarios

pytestmark = pytest.mark.gpu
F32 = np.float32


def case(msg):
    print(f"[case] {msg}", flush=True)


def upload(vg, a):
    return vg.DeviceArray.from_numpy(np.ascontiguousarray(a, F32))


@pytest.fixture(scope="module")
def ring_field(oracle):
    """A ring's treecode velocity grid (the field stretching sees in a frame)."""
    pos, om, radius = scenarios.vortex_ring(4096, seed=3)
    og = oracle.grid_for(pos, radius)
    tree = oracle.InfluenceTree(og, pos, om, 2 * radius)
    vel = oracle.velocity_grid(tree, og, pos, om, 2 * radius, technique=1, threads=8)
    return pos, om, og, vel


def test_jacobian_bit_exact(vg, oracle, ring_field):
    pos, om, og, vel = ring_field
    case(f"jacobian grid={og.shape}")
    g = vg.GridDesc.from_fields(og)
    jac = vg.DeviceArray.zeros((og.capacity, 3, 3))
    vg.jacobian(g, upload(vg, vel), jac)
    vg.synchronize()
    assert np.array_equal(jac.numpy(), oracle.jacobian(og, vel))


@pytest.mark.parametrize("dt", [1.0 / 30.0, 0.25])
def test_stretch_tilt_bit_exact(vg, oracle, ring_field, dt):
    pos, om, og, vel = ring_field
    case(f"stretch/tilt N={len(pos)} grid={og.shape} dt={dt}")
    g = vg.GridDesc.from_fields(og)
    want = oracle.stretch_tilt(og, oracle.jacobian(og, vel), pos, om, dt)
    assert not np.array_equal(want, om)
    dvel, dpos = upload(vg, vel), upload(vg, pos)
    jac = vg.DeviceArray.zeros((og.capacity, 3, 3))
    vg.jacobian(g, dvel, jac)
    w1 = upload(vg, om)
    vg.stretch_tilt(g, jac, dpos, w1, len(pos), dt)
    w2 = upload(vg, om)
    vg.stretch_tilt_from_velocity(g, dvel, dpos, w2, len(pos), dt)
    vg.synchronize()
    assert np.array_equal(w1.numpy(), want)
    assert np.array_equal(w2.numpy(), want)


def test_stretch_tilt_particle_records(vg, oracle, ring_field):
    """18-float (72-byte) Particle records: position at 0, angular velocity
    at 9; every other field is left alone."""
    pos, om, og, vel = ring_field
    case("stretch/tilt on 72-byte records")
    n = len(pos)
    rec = np.random.default_rng(31).normal(size=(n, 18)).astype(F32)
    rec[:, 0:3] = pos
    rec[:, 9:12] = om
    d = upload(vg, rec)
    g = vg.GridDesc.from_fields(og)
    vg.stretch_tilt_from_velocity(g, upload(vg, vel), d.ptr, d.offset(9), n, 0.1,
                                  pos_stride=18, ang_stride=18)
    vg.synchronize()
    out = d.numpy()
    want = oracle.stretch_tilt(og, oracle.jacobian(og, vel), pos, om, 0.1)
    assert np.array_equal(out[:, 9:12], want)
    keep = np.ones(18, bool)
    keep[9:12] = False
    assert np.array_equal(out[:, keep], rec[:, keep])


def test_stretch_tilt_flat_domain_is_a_no_op(vg, oracle, ring_field):
    pos, om, og, vel = ring_field
    case("stretch/tilt with a zero z extent")
    g = vg.GridDesc.from_fields(og)
    g.extent[2] = 0.0
    w = upload(vg, om)
    vg.stretch_tilt_from_velocity(g, upload(vg, vel), upload(vg, pos), w, len(pos), 0.1)
    vg.synchronize()
    assert np.array_equal(w.numpy(), om)


def test_stretch_tilt_large_fused_equals_two_step(vg, oracle):
    """1M vortons of 64 rings on a 129^3-class grid: the fused update equals
    the Jacobian-grid route bit for bit everywhere, and the oracle on a
    sample (the oracle on the Jacobian grid the GPU produced)."""
    pos, om, radius = scenarios.random_rings(1 << 20, seed=5)
    og = oracle.grid_for(pos, radius)
    case(f"stretch/tilt N={len(pos)} grid={og.shape}: fused vs two-step")
    g = vg.GridDesc.from_fields(og)
    rng = np.random.default_rng(32)
    # a smooth field: sum of a few Fourier modes over the grid points
    pts = np.stack(np.meshgrid(*[np.arange(og.np[a]) for a in (2, 1, 0)], indexing="ij")[::-1],
                   -1).reshape(-1, 3).astype(np.float64)
    vel = np.zeros((og.capacity, 3))
    for _ in range(4):
        k = rng.normal(size=3) * 0.2
        vel += np.sin(pts @ k + rng.uniform(0, 6))[:, None] * rng.normal(size=3)
    vel = vel.astype(F32)
    dvel, dpos = upload(vg, vel), upload(vg, pos)
    jac = vg.DeviceArray.zeros((og.capacity, 3, 3))
    vg.jacobian(g, dvel, jac)
    w1, w2 = upload(vg, om), upload(vg, om)
    vg.stretch_tilt(g, jac, dpos, w1, len(pos), 1.0 / 30.0)
    vg.stretch_tilt_from_velocity(g, dvel, dpos, w2, len(pos), 1.0 / 30.0)
    # any processing order gives the same values: the tree's cell order and a
    # random permutation
    cell = vg.DeviceArray.from_numpy(vg.cell_order(g, pos))
    perm = vg.DeviceArray.from_numpy(rng.permutation(len(pos)).astype(np.uint32))
    w3, w4 = upload(vg, om), upload(vg, om)
    vg.stretch_tilt_from_velocity(g, dvel, dpos, w3, len(pos), 1.0 / 30.0, order=cell)
    vg.stretch_tilt(g, jac, dpos, w4, len(pos), 1.0 / 30.0, order=perm)
    vg.synchronize()
    a, b = w1.numpy(), w2.numpy()
    assert np.array_equal(a, b)
    assert np.array_equal(w3.numpy(), a)
    assert np.array_equal(w4.numpy(), a)
    idx = rng.choice(len(pos), 4096, replace=False)
    want = oracle.stretch_tilt(og, jac.numpy(), pos[idx], om[idx], 1.0 / 30.0)
    assert np.array_equal(a[idx], want)


@pytest.mark.parametrize("technique", [1, 2])
def test_sim_stretch_tilt(vg, oracle, technique):
    """vg_sim_stretch_tilt after vg_sim_update equals the oracle's update on
    the simulation's own velocity grid."""
    pos, om, radius = scenarios.colliding_rings(8192, seed=33)
    case(f"sim stretch/tilt technique={technique}")
    n = len(pos)
    rec = np.zeros((n, 18), F32)
    rec[:, 0:3] = pos
    rec[:, 9:12] = om
    rec[:, 13] = 2 * radius
    d = upload(vg, rec)
    v = vg.Vortons()
    v.pos, v.angvel, v.size = d.ptr.value, d.offset(9).value, d.offset(13).value
    v.stride, v.n, v.uniform_size = 18, n, 0.0
    sim = vg.Sim(vg.sim_config(technique=technique))
    sim.set_vortons(v)
    sim.find_bounding_box()
    sim.update_bounding_box(pos[0], pos[0], True)
    sim.update(1.0 / 30.0, 0)
    sim.stretch_tilt(1.0 / 30.0, d.offset(9), 18)
    vg.synchronize()
    g, p = sim.velocity_grid()
    og = oracle.Grid()
    for f in ("min", "extent", "cell", "cpe", "np"):
        for a in range(3):
            getattr(og, f)[a] = getattr(g, f)[a]
    vel = np.empty((og.capacity, 3), F32)
    vg.call("vg_memcpy_dtoh", vel.ctypes.data_as(vg.C.c_void_p), p, vg.C.c_size_t(vel.nbytes))
    want = oracle.stretch_tilt(og, oracle.jacobian(og, vel), pos, om, 1.0 / 30.0)
    assert np.array_equal(d.numpy()[:, 9:12], want)

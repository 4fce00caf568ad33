"""vg_sort_particles_by_cell: a stable counting sort of particle records by
grid cell, checked against numpy's stable argsort of the same cell keys;
advecting the sorted tracers equals advecting them unsorted and permuting."""
import numpy as np
import pytest

from vortegrid_amd import scenarios

pytestmark = pytest.mark.gpu
F32 = np.float32


def cell_keys(g, pos):
    """Leaf cell of each position: truncation of (p - min) * cellsPerExtent,
    clamped to the grid's cells (x fastest)."""
    idx = []
    for a in range(3):
        f = (pos[:, a].astype(F32) - F32(g.min[a])) * F32(g.cpe[a])
        i = np.where(f > F32(-1.0), np.trunc(f), -1.0).astype(np.int64)
        idx.append(np.clip(i, 0, int(g.np[a]) - 2))
    return idx[0] + int(g.np[0]) * (idx[1] + int(g.np[1]) * idx[2])


@pytest.mark.parametrize("stride", [3, 18])
def test_sort_particles_by_cell(vg, oracle, stride):
    print(f"[case] cell sort of 300k records, stride {stride}", flush=True)
    pos, om, radius = scenarios.colliding_rings(16384, seed=2)
    og = oracle.grid_for(pos, radius)
    g = vg.GridDesc.from_fields(og)
    rng = np.random.default_rng(9)
    mn, mx = np.array(og.min), np.array(og.min) + np.array(og.extent)
    n = 300_000
    rec = rng.normal(size=(n, stride)).astype(F32)
    rec[:, 0:3] = scenarios.tracers_in_box(n, mn - 0.1, mx + 0.1, seed=10)  # some outside
    d = vg.DeviceArray.from_numpy(rec)
    perm = vg.DeviceArray.zeros((n,), np.uint32)
    vg.sort_particles_by_cell(g, d, n, stride=stride, perm=perm)
    vg.synchronize()
    want = np.argsort(cell_keys(og, rec[:, 0:3]), kind="stable")
    assert np.array_equal(perm.numpy().astype(np.int64), want)
    assert np.array_equal(d.numpy(), rec[want])


def test_sorted_advection_is_a_permutation(vg, oracle):
    print("[case] advection of cell-sorted tracers", flush=True)
    pos, om, radius = scenarios.colliding_rings(16384, seed=2)
    og = oracle.grid_for(pos, radius)
    g = vg.GridDesc.from_fields(og)
    rng = np.random.default_rng(11)
    field = rng.normal(size=(og.capacity, 3)).astype(F32)
    mn, mx = np.array(og.min), np.array(og.min) + np.array(og.extent)
    tr = scenarios.tracers_in_box(200_000, mn + 1e-3, mx - 1e-3, seed=12)
    dfield = vg.DeviceArray.from_numpy(field)
    a = vg.DeviceArray.from_numpy(tr)
    b = vg.DeviceArray.from_numpy(tr)
    perm = vg.DeviceArray.zeros((len(tr),), np.uint32)
    vg.sort_particles_by_cell(g, b, len(tr), perm=perm)
    vg.advect_tracers(g, dfield, a, None, len(tr), 1.0 / 30.0, 1.0)
    vg.advect_tracers(g, dfield, b, None, len(tr), 1.0 / 30.0, 1.0)
    vg.synchronize()
    p = perm.numpy().astype(np.int64)
    assert np.array_equal(b.numpy(), a.numpy()[p])
    assert np.array_equal(b.numpy(), oracle.advect(og, field, tr, np.zeros_like(tr), 1.0 / 30.0)[0][p])

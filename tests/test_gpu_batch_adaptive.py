"""Batched effects with grids defined on the device (vg_batch_config.cells = 0):
each effect's grid is UniformGridGeometry::DefineShape(vorton count, padded
box, power-of-two cells) every frame, computed inside the frame, and the
walk takes (effect, brick) items from a device work queue.

Checked here: the device's grid definition equals the host's (and so the
oracle's, which tests/test_abi.py pins) bit for bit; every effect's grid equals
the single-tree walk on that geometry (a one-vorton effect's single-cell
grid included), and so does every grid of the reference opening rule
(near_k = -1) against the single-tree reference walk; the max_cells limit is
counted; graph replays equal eager frames; and a one-effect batch runs the
reference frame of vg_sim (tracers with the previous grid, vorton and tracer
bounds, UpdateBoundingBox + DefineShape, tree, walk, vorton advection) bit for
bit with no host round trip, with either walk."""
import numpy as np
import pytest

from vortegrid_amd import scenarios

from test_gpu_batch import DT, F32, case, make_effects

pytestmark = pytest.mark.gpu


def adaptive_batch(vg, counts, seed, max_cells=64, **kw):
    vp, vw, tp, radii = make_effects(counts, seed)
    b = vg.Batch([len(p) for p in vp], [len(t) for t in tp], radii, cells=0,
                 max_cells=max_cells, **kw)
    b.upload(np.concatenate(vp), np.concatenate(vw), np.concatenate(tp))
    return b, vp, vw, tp, radii


def host_shape(vg, pos_all, n_vortons, radius, max_cells):
    """DefineShape on the host over the bounds padded as UpdateBoundingBox
    pads them (vortonSim.cpp:1069-1089); returns (desc, clamped)."""
    pad = F32(1.0 + 2.0) * F32(radius)
    mn = (pos_all.min(0) - pad).astype(F32)
    mx = (pos_all.max(0) + pad).astype(F32)
    g = vg.define_shape(n_vortons, mn, mx, pow2=True)
    return g, any(g.np[a] > max_cells + 1 for a in range(3))


def same_desc(a, b):
    return (list(a.np) == list(b.np) and
            all(np.float32(a.min[i]) == np.float32(b.min[i]) for i in range(3)) and
            all(np.float32(a.extent[i]) == np.float32(b.extent[i]) for i in range(3)) and
            all(np.float32(a.cell[i]) == np.float32(b.cell[i]) for i in range(3)) and
            all(np.float32(a.cpe[i]) == np.float32(b.cpe[i]) for i in range(3)))


def single_tree_grid(vg, g, pos, om, radius, spread, circ, near_k, order, expand_from):
    dp, dw = vg.DeviceArray.from_numpy(pos), vg.DeviceArray.from_numpy(om)
    v = vg.vortons_view(dp, dw, len(pos), uniform_size=2 * radius)
    t = vg.Tree()
    t.build(g, v)
    out = vg.DeviceArray.zeros((g.capacity, 3))
    t.velocity_grid_mac(g, out, near_k, order, spread, circ, policy=1, expand_from=expand_from)
    vg.synchronize()
    return out.numpy()


@pytest.mark.parametrize("expand_from", [-1, 2])
def test_adaptive_shapes_equal_host_define_shape_and_single_trees(vg, expand_from):
    counts = [4096, 700, 0, 3, 2500, 1]
    case(f"adaptive batch: device DefineShape vs host, grids vs single trees, effects {counts}, "
         f"expand_from={expand_from}")
    b, vp, vw, tp, radii = adaptive_batch(vg, counts, seed=23, near_k=2, order=1,
                                          expand_from=expand_from)
    for frame in range(2):
        b.step(DT)
        vg.synchronize()
        vpos, tpos = b.download("vpos"), b.download("tpos")
        vo = np.cumsum([0] + [len(p) for p in vp])
        to = np.cumsum([0] + [len(t) for t in tp])
        for e, n in enumerate(counts):
            g, field = b.grid(e)
            spread, circ = b.mollify(e)
            # the frame's grid was defined before its Evolve: bounds of the
            # positions at the start of the frame, after the tracers moved
            pv = prev_v[e] if frame else vp[e]
            pt = tpos[to[e]:to[e + 1]]
            if n == 0:
                # no vortons: 4 cells per axis over the padded tracer box, no field
                pad = F32(1.0 + 2.0) * F32(radii[e])
                mn, mx = (pt.min(0) - pad).astype(F32), (pt.max(0) + pad).astype(F32)
                assert list(g.np) == [5, 5, 5]
                assert np.array_equal(np.float32(g.min), mn)
                assert np.array_equal(np.float32(g.extent), (mx - mn).astype(F32))
                assert not np.any(field)
                continue
            hg, _ = host_shape(vg, np.concatenate([pv, pt]), n, radii[e], 64)
            assert same_desc(g, hg), (frame, e, list(g.np), list(hg.np))
            if n == 1:
                assert list(g.np) == [2, 2, 2]  # one cell: a tree of one layer
            ref = single_tree_grid(vg, g, pv, vw[e], radii[e], spread, circ, 2, 1, expand_from)
            assert np.array_equal(field, ref), (frame, e)
        prev_v = [vpos[vo[e]:vo[e + 1]] for e in range(len(counts))]
    assert b.clamped() == 0


def test_adaptive_reference_rule_equals_single_trees(vg):
    """near_k = -1: every effect's grid by the reference opening rule
    (k_tree_grid on the batched trees, whose empty leaf cells are then
    written too) equals the single-tree reference walk on the effect's
    geometry bit for bit, including an effect without vortons and a
    one-vorton effect (a one-layer tree: the direct sum over its cell)."""
    counts = [4096, 700, 0, 3, 2500, 1]
    case(f"adaptive batch, reference opening rule vs single trees, effects {counts}")
    b, vp, vw, tp, radii = adaptive_batch(vg, counts, seed=41, near_k=-1)
    vo = np.cumsum([0] + counts)
    for frame in range(2):
        b.step(DT)
        vg.synchronize()
        vpos = b.download("vpos")
        for e, n in enumerate(counts):
            g, field = b.grid(e)
            if n == 0:
                assert not np.any(field)
                continue
            spread, circ = b.mollify(e)
            pv = prev_v[e] if frame else vp[e]
            dp, dw = vg.DeviceArray.from_numpy(pv), vg.DeviceArray.from_numpy(vw[e])
            v = vg.vortons_view(dp, dw, n, uniform_size=2 * radii[e])
            t = vg.Tree()
            t.build(g, v)
            ref = vg.DeviceArray.zeros((g.capacity, 3))
            t.velocity_grid(g, ref, spread, circ, 1)
            vg.synchronize()
            assert np.array_equal(field, ref.numpy()), (frame, e, list(g.np))
        prev_v = [vpos[vo[e]:vo[e + 1]] for e in range(len(counts))]
    assert b.clamped() == 0


def test_adaptive_max_cells_limit(vg):
    """A shape beyond max_cells on an axis is limited there and counted; its
    grid still equals the single-tree walk on the limited geometry."""
    counts = [20000, 900]
    case(f"adaptive batch: max_cells 16 limits a 20000-vorton effect, effects {counts}")
    b, vp, vw, tp, radii = adaptive_batch(vg, counts, seed=29, max_cells=16, near_k=2, order=1)
    b.step(DT)
    vg.synchronize()
    tpos = b.download("tpos")
    to = np.cumsum([0] + [len(t) for t in tp])
    clamped = 0
    for e, n in enumerate(counts):
        g, field = b.grid(e)
        hg, over = host_shape(vg, np.concatenate([vp[e], tpos[to[e]:to[e + 1]]]), n, radii[e], 16)
        clamped += over
        assert all(g.np[a] == min(hg.np[a], 17) for a in range(3)), (e, list(g.np), list(hg.np))
        spread, circ = b.mollify(e)
        ref = single_tree_grid(vg, g, vp[e], vw[e], radii[e], spread, circ, 2, 1, -1)
        assert np.array_equal(field, ref), e
    assert clamped >= 1 and b.clamped() == clamped


def test_adaptive_graph_replay_matches_eager(vg):
    case("adaptive batch: hipGraph replays vs eager frames, 3 frames")
    counts = [1500, 3300, 600, 2100]
    a, *_ = adaptive_batch(vg, counts, seed=31, near_k=2, order=1, expand_from=2)
    b, *_ = adaptive_batch(vg, counts, seed=31, near_k=2, order=1, expand_from=2)
    s = vg.Stream()
    b.capture(DT, s)
    for f in range(3):
        a.step(DT)
        b.launch(s)
        s.sync()
        vg.synchronize()
        for key in ("vpos", "vvel", "tpos"):
            assert np.array_equal(a.download(key), b.download(key)), (f, key)
        for e in range(len(counts)):
            ga, fa = a.grid(e)
            gb, fb = b.grid(e)
            assert same_desc(ga, gb) and np.array_equal(fa, fb), (f, e)


@pytest.mark.parametrize("near_k", [2, -1])
def test_one_effect_batch_is_the_sim_frame(vg, near_k):
    """The reference frame with host geometry (vg_sim: tracers with the
    previous grid, tracer + vorton bounds, UpdateBoundingBox + DefineShape on
    the host, tree, near-field walk or, with near_k = -1, the reference
    opening rule, vortons AssignVelocityFromField + Evolve) equals a
    one-effect adaptive batch frame, which defines the grid on the device,
    bit for bit over 3 frames."""
    rule = "near-field" if near_k >= 0 else "reference-rule"
    case(f"one-effect adaptive batch vs vg_sim TREE {rule} frame, colliding rings 16384, 3 frames")
    pos, om, radius = scenarios.colliding_rings(16384, seed=37)
    tr = scenarios.tracers_in_box(65536, pos.min(0), pos.max(0), seed=38)
    order, expand_from = 1, 2
    s = vg.Stream()
    d_pos, d_om = vg.DeviceArray.from_numpy(pos), vg.DeviceArray.from_numpy(om)
    d_vel = vg.DeviceArray.zeros(pos.shape)
    d_tr = vg.DeviceArray.from_numpy(tr)
    v = vg.vortons_view(d_pos, d_om, len(pos), uniform_size=2 * radius)
    sim = vg.Sim(vg.sim_config(technique=vg.TECH_TREE, rsqrt_policy=1, tree_near_k=near_k,
                               tree_order=order, tree_expand_from=expand_from), s)
    sim.set_vortons(v)
    b = vg.Batch([len(pos)], [len(tr)], [radius], cells=0, max_cells=256, near_k=near_k,
                 order=order, expand_from=expand_from)
    b.upload(pos, om, tr)
    prev = None
    for f in range(3):
        if prev is not None:
            vg.advect_tracers(prev[0], prev[1], d_tr, None, len(tr), float(DT), 1.0, stream=s)
        tmn, tmx = sim.find_bounding_boxes(d_tr, len(tr))
        sim.update_bounding_box(tmn, tmx, True)
        sim.update(float(DT), 0, d_vel)
        vg.evolve(d_pos, d_vel, len(pos), float(DT), stream=s)
        g, p = sim.velocity_grid()
        prev = (g, p)
        b.step(DT)
        s.sync()
        vg.synchronize()
        gb, fb = b.grid(0)
        assert same_desc(g, gb), (f, list(g.np), list(gb.np))
        assert b.mollify(0) == sim.spreading(), f
        assert np.array_equal(vg.fetch(p, 3 * g.capacity).reshape(-1, 3), fb), f
        assert np.array_equal(d_pos.numpy(), b.download("vpos")), f
        assert np.array_equal(d_tr.numpy(), b.download("tpos")), f
    assert b.clamped() == 0

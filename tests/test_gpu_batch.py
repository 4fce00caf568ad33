"""Batched effects: every effect of a batch frame equals the single-tree
path on the same geometry (bit for bit), the recorded hipGraph frame equals
the eager frame over several replays, and autotuning changes no result."""
import numpy as np
import pytest

from vortegrid_amd import scenarios

pytestmark = pytest.mark.gpu
F32 = np.float32
DT = F32(1.0 / 30.0)


def case(msg):
    print(f"[case] {msg}", flush=True)


def make_effects(counts, seed):
    """Rings and plumes of the given vorton counts, 4 tracers per vorton."""
    vp, vw, tp, radii = [], [], [], []
    for i, n in enumerate(counts):
        if n == 0:
            pos = np.zeros((0, 3), F32)
            om = np.zeros((0, 3), F32)
            radius = 0.05
            tr = scenarios.tracers_in_box(64, np.float32([-1, -1, -1]), np.float32([1, 1, 1]), seed=seed + i)
        else:
            gen = scenarios.vortex_ring if i % 2 == 0 else scenarios.jet_plume
            pos, om, radius = gen(n, seed=seed + i)
            pos = pos + F32(3.0 * i)  # effects live in different places
            tr = scenarios.tracers_in_box(4 * n, pos.min(0), pos.max(0), seed=seed + 100 + i)
        vp.append(pos)
        vw.append(om)
        tp.append(tr)
        radii.append(radius)
    return vp, vw, tp, radii


def new_batch(vg, counts, seed=1, **kw):
    vp, vw, tp, radii = make_effects(counts, seed)
    b = vg.Batch([len(p) for p in vp], [len(t) for t in tp], radii, **kw)
    b.upload(np.concatenate(vp) if vp else np.zeros((0, 3), F32), np.concatenate(vw),
             np.concatenate(tp))
    return b, vp, vw, tp, radii


COUNTS = [700, 2500, 0, 1800, 300, 4096]


@pytest.mark.parametrize("near_k,expand_from", [(1, -1), (1, 2), (-1, -1)])
def test_batch_frame_matches_single_trees(vg, near_k, expand_from):
    """Every effect's grid (the near-field walk, or with near_k = -1 the
    reference opening rule), the vortons' velocities and positions after the
    frame and the tracers moved with the zero initial grids equal the same
    steps on each effect alone, bit for bit."""
    method = "reference rule" if near_k < 0 else f"k={near_k} dipole, expand_from={expand_from}"
    case(f"batched frame vs single trees, effects {COUNTS}, cells 16, {method}")
    b, vp, vw, tp, radii = new_batch(vg, COUNTS, cells=16, near_k=near_k, order=1,
                                     expand_from=expand_from)
    b.step(DT)
    vg.synchronize()
    vpos1, vvel1, tpos1 = b.download("vpos"), b.download("vvel"), b.download("tpos")
    vo = np.cumsum([0] + [len(p) for p in vp])
    to = np.cumsum([0] + [len(t) for t in tp])
    grids = []
    for e in range(len(COUNTS)):
        g, field = b.grid(e)
        spread, circ = b.mollify(e)
        grids.append((g, field))
        # geometry: bounds of vortons u tracers padded by 3 radii, 16^3 cells
        pts = np.concatenate([vp[e], tp[e]])
        pad = F32(3.0) * F32(radii[e])
        mn, mx = pts.min(0) - pad, pts.max(0) + pad
        assert list(g.np) == [17, 17, 17]
        assert np.array_equal(np.float32(g.min), mn)
        assert np.array_equal(np.float32(g.extent), (mx - mn).astype(F32))
        # the same tree and walk on this effect alone
        n = len(vp[e])
        if n:
            dp, dw = vg.DeviceArray.from_numpy(vp[e]), vg.DeviceArray.from_numpy(vw[e])
            v = vg.vortons_view(dp, dw, n, uniform_size=2 * radii[e])
            t = vg.Tree()
            t.build(g, v)
            out = vg.DeviceArray.zeros((g.capacity, 3))
            if near_k < 0:
                t.velocity_grid(g, out, spread, circ, 1)
            else:
                t.velocity_grid_mac(g, out, near_k, 1, spread, circ, policy=1,
                                    expand_from=expand_from)
            vg.synchronize()
            assert np.array_equal(out.numpy(), field), e
            # vortons: AssignVelocityFromField + Evolve with that grid
            fld = vg.DeviceArray.from_numpy(field)
            pv = vg.DeviceArray.from_numpy(vp[e])
            vv = vg.DeviceArray.zeros((n, 3))
            vg.assign_velocity_from_field(g, fld, pv, vv, n)
            vg.evolve(pv, vv, n, float(DT))
            vg.synchronize()
            assert np.array_equal(vv.numpy(), vvel1[vo[e]:vo[e + 1]]), e
            assert np.array_equal(pv.numpy(), vpos1[vo[e]:vo[e + 1]]), e
        else:
            assert not np.any(field)
        # frame 1 tracers moved with the zero initial grids
        assert np.array_equal(tpos1[to[e]:to[e + 1]], tp[e])
    # frame 2: tracers advect with each effect's frame-1 grid
    b.step(DT)
    vg.synchronize()
    tpos2 = b.download("tpos")
    for e in range(len(COUNTS)):
        g, field = grids[e]
        tr = vg.DeviceArray.from_numpy(tp[e])
        vg.advect_tracers(g, vg.DeviceArray.from_numpy(field), tr, None, len(tp[e]), float(DT), 1.0)
        vg.synchronize()
        assert np.array_equal(tr.numpy(), tpos2[to[e]:to[e + 1]]), e


def test_graph_replay_matches_eager(vg):
    case("batched frame: hipGraph replays vs eager frames, 4 frames")
    counts = [1200, 3000, 500, 2200]
    a, *_ = new_batch(vg, counts, seed=7, cells=16, near_k=2, order=1)
    b, *_ = new_batch(vg, counts, seed=7, cells=16, near_k=2, order=1)
    s = vg.Stream()
    b.capture(DT, s)
    for f in range(4):
        a.step(DT)
        b.launch(s)
        s.sync()
        vg.synchronize()
        for key in ("vpos", "vvel", "tpos"):
            assert np.array_equal(a.download(key), b.download(key)), (f, key)
        for e in range(len(counts)):
            assert np.array_equal(a.grid(e)[1], b.grid(e)[1]), (f, e)


def test_autotune_keeps_results(vg):
    case("batched frame: autotune changes launch shapes only")
    counts = [900, 2600, 1500]
    a, *_ = new_batch(vg, counts, seed=11, cells=16, near_k=1, order=1)
    b, *_ = new_batch(vg, counts, seed=11, cells=16, near_k=1, order=1)
    a.step(DT)
    b.step(DT)
    report = b.autotune()
    print(report)
    assert "chosen" in report
    b.set_launch(1, 512, 4)
    assert b.get_launch() == (1, 512, 4)
    for _ in range(2):
        a.step(DT)
        b.step(DT)
    vg.synchronize()
    for key in ("vpos", "vvel", "tpos"):
        assert np.array_equal(a.download(key), b.download(key)), key


def test_solve_equals_a_still_step_and_set_method(vg):
    """vg_batch_solve computes the grids of the current particles: equal to a
    frame with dt = 0 (which moves nothing).  set_method switches the walk:
    the grids then equal a batch created with that method."""
    case("batched solve vs step(dt=0); set_method vs a batch created with the method")
    counts = [1500, 700, 2600]
    a, *_ = new_batch(vg, counts, seed=13, cells=16, near_k=1, order=1)
    b, *_ = new_batch(vg, counts, seed=13, cells=16, near_k=1, order=1)
    a.step(F32(0.0))
    b.solve()
    vg.synchronize()
    for e in range(len(counts)):
        assert np.array_equal(a.grid(e)[1], b.grid(e)[1]), e
    for key in ("vpos", "tpos"):
        assert np.array_equal(a.download(key), b.download(key)), key
    c, *_ = new_batch(vg, counts, seed=13, cells=16, near_k=2, order=1, expand_from=2)
    b.set_method(2, 1, 2)
    b.solve()
    c.solve()
    vg.synchronize()
    for e in range(len(counts)):
        assert np.array_equal(b.grid(e)[1], c.grid(e)[1]), e


def test_method_autotuner_meets_its_target(vg, tmp_path):
    """The autotuner picks the fastest method within the accuracy target,
    leaves the batch set to it, and its JSON applies to another batch."""
    case("batched autotuner: method against an accuracy target, JSON round trip")
    from vortegrid_amd import autotune

    counts = [4096, 1200, 2800, 600, 3500, 2000]
    b, vp, vw, tp, radii = new_batch(vg, counts, seed=17, cells=16, near_k=1, order=1)
    b.step(DT)
    path = tmp_path / "choice.json"
    rec = autotune.tune_batch(b, np.concatenate(vw), radii, target=2e-2,
                              candidates=((1, -1), (2, -1), (2, 2)), sample_points=256,
                              path=str(path))
    print(rec["trials"], rec["method"])
    assert rec["met_target"] and rec["chosen"]["rel_l2"] <= 2e-2
    fastest_ok = min((t for t in rec["trials"] if t["meets_target"]), key=lambda t: t["solve_ms"])
    assert rec["chosen"] == fastest_ok
    errs = {(t["near_k"], t["expand_from"]): t["rel_l2"] for t in rec["trials"]}
    assert errs[(2, -1)] < errs[(1, -1)]
    c, *_ = new_batch(vg, counts, seed=17, cells=16, near_k=1, order=1)
    loaded = autotune.load_choice(c, str(path))
    assert c.get_launch() == (loaded["launch"]["walk_waves"], loaded["launch"]["advect_block"],
                              loaded["launch"]["advect_items"])


def test_sort_tracers_is_per_effect_cell_order(vg):
    """vg_batch_sort_tracers equals the single-grid stable cell sort of each
    effect's tracers on that effect's grid, keeps every tracer in its effect,
    and the next frame equals the frame of an unsorted twin up to that
    permutation, bit for bit (grids included)."""
    case(f"batched tracer cell sort vs per-effect single-grid sort, effects {COUNTS}, cells 16")
    a, vp, vw, tp, radii = new_batch(vg, COUNTS, seed=19, cells=16, near_k=1, order=1)
    b, *_ = new_batch(vg, COUNTS, seed=19, cells=16, near_k=1, order=1)
    with pytest.raises(vg.VGError):
        b.sort_tracers()  # no grids yet
    a.step(DT)
    b.step(DT)
    vg.synchronize()
    before = b.download("tpos")
    b.sort_tracers()
    vg.synchronize()
    after = b.download("tpos")
    to = np.cumsum([0] + [len(t) for t in tp])
    perm = np.empty(len(before), np.int64)
    for e in range(len(COUNTS)):
        g, _ = b.grid(e)
        seg = vg.DeviceArray.from_numpy(before[to[e]:to[e + 1]])
        p = vg.DeviceArray.zeros((len(tp[e]),), np.uint32)
        vg.sort_particles_by_cell(g, seg, len(tp[e]), perm=p)
        vg.synchronize()
        assert np.array_equal(seg.numpy(), after[to[e]:to[e + 1]]), e
        perm[to[e]:to[e + 1]] = to[e] + p.numpy().astype(np.int64)
    assert np.array_equal(before[perm], after)
    a.step(DT)
    b.step(DT)
    vg.synchronize()
    assert np.array_equal(a.download("tpos")[perm], b.download("tpos"))
    for key in ("vpos", "vvel"):
        assert np.array_equal(a.download(key), b.download(key)), key
    for e in range(len(COUNTS)):
        assert np.array_equal(a.grid(e)[1], b.grid(e)[1]), e

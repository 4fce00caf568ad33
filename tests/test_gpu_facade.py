"""The C++ VortonSim facade (vortegrid_amd/host/vortonsim.h) wired with the
reference's own statements (tests/cpp/facade_frame.cpp) and driven frame by
frame on 72-byte Particle records: tracers AssignVelocityFromField (from the
VortonSim's velocity grid object), Evolve, FindBoundingBox; then vortons
PclOpVortonSim (on the tracer box), Evolve, PopulateVelocityGrid (into the
same grid object, bounded through GetMinCorner()).  Every frame is checked
against the oracle, starting from the facade's state at the previous frame."""
import os
import subprocess

import numpy as np
import pytest

from vortegrid_amd import scenarios

pytestmark = pytest.mark.gpu
F32 = np.float32
DT = 1.0 / 30.0
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "facade_frame")
GRID_DT = np.dtype([("min", F32, 3), ("extent", F32, 3), ("cell", F32, 3), ("cpe", F32, 3),
                    ("np", np.uint32, 3)])


def records(pos, om=None, size=0.0):
    r = np.zeros((len(pos), 18), F32)
    r[:, 0:3] = pos
    if om is not None:
        r[:, 9:12] = om
    r[:, 13] = size
    return r


def rel_l2(a, b):
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def read_frames(path, frames, nv, nt):
    raw = open(path, "rb").read()
    off = 0
    out = []

    def grid():
        nonlocal off
        gd = np.frombuffer(raw[off:off + GRID_DT.itemsize], GRID_DT)[0]
        off += GRID_DT.itemsize
        cap = int(np.prod(gd["np"])) if gd["np"][0] else 0
        field = np.frombuffer(raw[off:off + 12 * cap], F32).reshape(-1, 3)
        off += 12 * cap
        return gd, field

    for _ in range(frames):
        sim_g, sim_f = grid()
        pop_g, pop_f = grid()
        v = np.frombuffer(raw[off:off + 72 * nv], F32).reshape(-1, 18)
        off += 72 * nv
        t = np.frombuffer(raw[off:off + 72 * nt], F32).reshape(-1, 18)
        off += 72 * nt
        out.append({"sim": (sim_g, sim_f), "pop": (pop_g, pop_f), "v": v, "t": t})
    assert off == len(raw)
    return out


def oracle_grid(oracle, gd):
    g = oracle.Grid()
    for f in ("min", "extent", "cell", "cpe", "np"):
        for a in range(3):
            getattr(g, f)[a] = gd[f][a]
    return g


def run_driver(tmp_path, technique, vrec, trec, frames, vpm=False, viscosity=0.0):
    if not os.path.exists(EXE):
        from vortegrid_amd import build

        build.build_facade_driver()
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([len(vrec), len(trec)], np.uint64).tofile(f)
        vrec.tofile(f)
        trec.tofile(f)
    args = [EXE, str(technique), str(inp), str(outp), str(frames),
            "vpm" if vpm else "hybrid", str(viscosity)]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    return read_frames(outp, frames, len(vrec), len(trec))


def check_frame(oracle, technique, radius, prev_v, prev_t, prev_pop, got, vpm, viscosity,
                exact_vorticity):
    """One frame from the state (prev_v, prev_t, prev_pop) against the oracle."""
    pos, om = prev_v[:, 0:3], prev_v[:, 9:12]
    size = 2 * radius
    # tracers: AssignVelocityFromField with last frame's grid, Evolve, box
    tr, tv = prev_t[:, 0:3], prev_t[:, 3:6]
    if prev_pop is not None:
        pg = oracle_grid(oracle, prev_pop[0])
        tp, tv = oracle.advect(pg, prev_pop[1], tr, tv, DT, 1.0)
    else:  # empty grid: zero extent, velocities kept
        tp = (tr + tv * F32(DT)).astype(F32)
    assert rel_l2(got["t"][:, 3:6], tv) < 1e-5
    assert rel_l2(got["t"][:, 0:3], tp) < 1e-6
    tmn, tmx = oracle.find_bbox(got["t"][:, 0:3])
    # vortons: PclOpVortonSim on the tracer box, Evolve
    vmn, vmx = oracle.find_bbox(pos)
    og, pmn, pmx = oracle.finalize_bbox_corners(np.minimum(vmn, tmn), np.maximum(vmx, tmx),
                                                radius, len(pos))
    gd, grid = got["sim"]
    assert [int(x) for x in gd["np"]] == list(og.np)
    ot = oracle.InfluenceTree(og, pos, om, size)
    if technique == 1:
        s, c = oracle.auto_mollify(og, radius)
        want = oracle.velocity_grid(ot, og, pos, om, size, 1, s, c, 0, threads=8)
    elif technique == 2:
        negw = oracle.splat_vorticity(og, pos, om, radius, -1.0)
        a0 = oracle.vecpot_boundary(ot, og, pos, om, size, 1.0, 1.0, 0, True, threads=8)
        want = oracle.curl(og, oracle.mg_cascade(og, a0, negw, 16))
    assert rel_l2(grid, want) < 1e-5
    vout = got["v"]
    wp, wv = oracle.advect(og, want, pos, np.zeros_like(pos), DT, 1.0)
    assert rel_l2(vout[:, 3:6], wv) < 1e-5
    assert rel_l2(vout[:, 0:3], wp) < 1e-6
    # vorticity terms, before Evolve: StretchAndTiltVortons on the frame's own
    # velocity grid (vortex particle method only), then PSE diffusion on the
    # frame's partition (any technique, when the viscosity is not zero)
    fg = oracle_grid(oracle, gd)
    w = om
    if vpm:
        w = oracle.stretch_tilt(fg, oracle.jacobian(fg, grid), pos, w, DT)
    if viscosity:
        w = oracle.pse_diffuse(fg, pos, w, size, viscosity, DT)
    if exact_vorticity:
        assert np.array_equal(vout[:, 9:12], w)
    else:
        assert rel_l2(vout[:, 9:12], w) < 1e-5
    assert np.array_equal(vout[:, 9:12], om) == (not vpm and not viscosity)
    # PopulateVelocityGrid into the VortonSim's grid object: the sim's padded
    # box (read through &GetMinCorner(), two adjacent corners) grown by the
    # moved vortons and padded by the first vorton's radius
    mn2, mx2 = oracle.find_bbox(vout[:, 0:3])
    mn2 = np.minimum(mn2, pmn) - F32(0.5) * vout[0, 13]
    mx2 = np.maximum(mx2, pmx) + F32(0.5) * vout[0, 13]
    pg2 = oracle.define_shape(len(pos), mn2.astype(F32), mx2.astype(F32), pow2=True)
    pgd, pfield = got["pop"]
    assert [int(x) for x in pgd["np"]] == list(pg2.np)
    assert np.array_equal(np.asarray(pgd["min"], F32), np.asarray(pg2.min, F32))
    wpop = oracle.populate_velocity_grid(pg2, vout[:, 0:3], vout[:, 3:6])
    assert rel_l2(pfield, wpop) < 1e-6


def make_inputs(seed=41):
    pos, om, radius = scenarios.colliding_rings(8192, seed=seed)
    mn, mx = pos.min(0), pos.max(0)
    tr = scenarios.tracers_in_box(20000, mn - 0.2, mx + 0.2, seed=seed + 1)
    return pos, om, radius, records(pos, om, 2 * radius), records(tr)


@pytest.mark.parametrize("technique,vpm,viscosity", [(1, False, 0.0), (2, False, 0.0),
                                                     (1, True, 0.0), (2, True, 0.05),
                                                     (1, False, 0.05)])
def test_facade_frame(vg, oracle, tmp_path, technique, vpm, viscosity):
    print(f"[case] C++ facade frame technique={technique} vpm={vpm} viscosity={viscosity}",
          flush=True)
    pos, om, radius, vrec, trec = make_inputs()
    got = run_driver(tmp_path, technique, vrec, trec, 1, vpm, viscosity)
    check_frame(oracle, technique, radius, vrec, trec, None, got[0], vpm, viscosity, True)


@pytest.mark.parametrize("technique", [1, 2])
def test_facade_three_frames(vg, oracle, tmp_path, technique):
    """Three frames with the reference wiring: from the second frame on the
    tracers move with the grid PopulateVelocityGrid left in the VortonSim's
    velocity grid object the frame before."""
    print(f"[case] C++ facade, 3 frames, technique={technique}", flush=True)
    pos, om, radius, vrec, trec = make_inputs(seed=43)
    got = run_driver(tmp_path, technique, vrec, trec, 3)
    prev_v, prev_t, prev_pop = vrec, trec, None
    for f in range(3):
        check_frame(oracle, technique, radius, prev_v, prev_t, prev_pop, got[f], False, 0.0,
                    True)
        prev_v, prev_t, prev_pop = got[f]["v"], got[f]["t"], got[f]["pop"]
    # the tracers did move with the populated grids
    assert not np.array_equal(got[2]["t"][:, 0:3], trec[:, 0:3])

"""The C++ VortonSim facade (vortegrid_amd/host/vortonsim.h) driving one
reference-ordered frame on 72-byte Particle records, checked against the
oracle pipeline on the same inputs."""
import os
import subprocess

import numpy as np
import pytest

from vortegrid_amd import scenarios

pytestmark = pytest.mark.gpu
F32 = np.float32
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
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("technique,vpm,viscosity", [(1, False, 0.0), (2, False, 0.0),
                                                     (1, True, 0.0), (2, True, 0.05),
                                                     (1, False, 0.05)])
def test_facade_frame(vg, oracle, tmp_path, technique, vpm, viscosity):
    print(f"[case] C++ facade frame technique={technique} vpm={vpm} viscosity={viscosity}",
          flush=True)
    if not os.path.exists(EXE):
        from vortegrid_amd import build

        build.build_facade_driver()
    pos, om, radius = scenarios.colliding_rings(8192, seed=41)
    mn, mx = pos.min(0), pos.max(0)
    tr = scenarios.tracers_in_box(20000, mn - 0.2, mx + 0.2, seed=42)
    vrec, trec = records(pos, om, 2 * radius), records(tr)
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([len(pos), len(tr)], np.uint64).tofile(f)
        vrec.tofile(f)
        trec.tofile(f)
    extra = (["vpm" if vpm else "hybrid", str(viscosity)] if (vpm or viscosity) else [])
    r = subprocess.run([EXE, str(technique), str(inp), str(outp)] + extra,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    raw = open(outp, "rb").read()
    gd = np.frombuffer(raw[:GRID_DT.itemsize], GRID_DT)[0]
    cap = int(np.prod(gd["np"]))
    off = GRID_DT.itemsize
    grid = np.frombuffer(raw[off:off + 12 * cap], F32).reshape(-1, 3)
    off += 12 * cap
    vout = np.frombuffer(raw[off:off + 72 * len(pos)], F32).reshape(-1, 18)
    off += 72 * len(pos)
    tout = np.frombuffer(raw[off:off + 72 * len(tr)], F32).reshape(-1, 18)
    # oracle pipeline: bounds of vortons and tracers, tree, velocity grid
    og = oracle.grid_for(pos, radius, extra_pos=tr)
    assert [int(x) for x in gd["np"]] == list(og.np)
    ot = oracle.InfluenceTree(og, pos, om, 2 * radius)
    if technique == 1:
        s, c = oracle.auto_mollify(og, radius)
        want = oracle.velocity_grid(ot, og, pos, om, 2 * radius, 1, s, c, 0, threads=8)
    else:
        negw = oracle.splat_vorticity(og, pos, om, radius, -1.0)
        a0 = oracle.vecpot_boundary(ot, og, pos, om, 2 * radius, 1.0, 1.0, 0, True, threads=8)
        want = oracle.curl(og, oracle.mg_cascade(og, a0, negw, 16))
    assert rel_l2(grid, want) < 1e-5
    dt = 1.0 / 30.0
    wp, wv = oracle.advect(og, want, pos, np.zeros_like(pos), dt, 1.0)
    assert rel_l2(vout[:, 3:6], wv) < 1e-5
    assert rel_l2(vout[:, 0:3], wp) < 1e-6
    # vorticity terms, before Evolve: StretchAndTiltVortons on the frame's own
    # velocity grid (vortex particle method only), then PSE diffusion on the
    # frame's partition (any technique, when the viscosity is not zero)
    fg = oracle.Grid()
    for f in ("min", "extent", "cell", "cpe", "np"):
        for a in range(3):
            getattr(fg, f)[a] = gd[f][a]
    w = om
    if vpm:
        w = oracle.stretch_tilt(fg, oracle.jacobian(fg, grid), pos, w, dt)
    if viscosity:
        w = oracle.pse_diffuse(fg, pos, w, 2 * radius, viscosity, dt)
    assert np.array_equal(vout[:, 9:12], w)
    assert np.array_equal(vout[:, 9:12], om) == (not vpm and not viscosity)
    tp, tv = oracle.advect(og, want, tr, np.zeros_like(tr), dt, 1.0)
    assert rel_l2(tout[:, 3:6], tv) < 1e-5
    assert rel_l2(tout[:, 0:3], tp) < 1e-6

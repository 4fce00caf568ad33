"""Velocity-from-vorticity frame benchmark on one MI355X.

Workload (per step = one frame, reference frame order: tracers first,
particleSystemConfiguration.cpp:282-296):
  1. tracers: AssignVelocityFromField(previous grid) + Evolve (fused)
  2. tracer AABB + vorton AABB -> UpdateBoundingBox(final) -> DefineShape
  3. VortonSim::Update, TREE technique: influence tree, auto-mollified
     treecode velocity at every grid point, vortons get grid velocity
  4. vortons: Evolve
By default every stage runs on the device with nothing read back (a
one-effect vg_batch whose grid DefineShape defines on the device); the same
frames through vg_sim, which reads the two boxes back and defines the grid
on the host, are re-run on the same input, timed and compared bit for bit
(--host-geometry makes that form the headline).
on N = 1,048,576 vortons (64 random rings) and 4,194,304 tracers, inputs
resident in HBM.  value = treecode velocity evaluations (grid points) per
second over the whole frame; rel-L2 of the treecode against an FP64 direct
sum on a seeded 1% sample of grid points is reported beside it.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--vortons N] [--tracers T]
               [--near-k K] [--order O] [--expand-from L] [--no-cpu] [--no-accuracy]
Other workloads (one JSON line each): --suite (direct sums, a 65k-vorton
treecode frame with 1M tracers, the vector-Poisson solve), --batched (256
effects in one recorded frame launch), --mlp (the learned far-field
correction), --copy-probe (counter calibration).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FLOP_PER_PAIR = 26.0   # monopole / direct pair (FMA = 2, rsqrt = 1)
FLOP_DIPOLE = 42.0     # first-moment correction of one cluster interaction
# brick-centred local expansion of one cluster (FMA = 2): monopole to second
# order (offset and distance powers 20, Omega x d 9, value 6, gradient 42,
# Hessian 98: the 10 distinct components of the symmetric second-derivative
# tensor 26, Omega x its six columns 72) and first moment to first order
# (moments and t x d 51, gradient 99); per grid point, the 30-coefficient
# wave sum (180) and the polynomial (60)
FLOP_EXPAND_MONO = 175.0
FLOP_EXPAND_DIPOLE = 150.0
FLOP_EXPAND_EVAL = 240.0
BF16_PEAK_TFLOPS = 2500.0
MLP_FLOP = 2.0 * (16 * 64 + 64 * 64 + 64 * 64 + 64 * 3)  # per (point, cluster) pair


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class HostFrame:
    """The frame through vg_sim, as the reference wires it: tracers, the two
    bounding boxes read back to the host, UpdateBoundingBox + DefineShape on
    the host, VortonSim::Update (tree, near-field walk), vortons Evolve."""

    walk_kernels = ("vg_tree_velocity_grid_mac", "vg_tree_mac_expand")
    tracer_kernel = "vg_advect_tracers"
    sort_prefix = "vg_sort_"

    def __init__(self, vg, args, pos, om, radius, tr, stream):
        self.vg, self.args, self.stream, self.dt = vg, args, stream, 1.0 / 30.0
        self.n, self.nt = len(pos), len(tr)
        self.d_pos = vg.DeviceArray.from_numpy(pos)
        self.d_om = vg.DeviceArray.from_numpy(om)
        self.d_vel = vg.DeviceArray.zeros(pos.shape)
        self.d_tr = vg.DeviceArray.from_numpy(tr)
        self.v = vg.vortons_view(self.d_pos, self.d_om, len(pos), uniform_size=2 * radius)
        self.sim = vg.Sim(vg.sim_config(technique=vg.TECH_TREE, rsqrt_policy=args.policy,
                                        tree_near_k=args.near_k, tree_order=args.order,
                                        tree_expand_from=args.expand_from), stream)
        self.sim.set_vortons(self.v)
        self.grid = None
        self.frames = 0
        self.evals = 0

    def frame(self):
        vg, a = self.vg, self.args
        # tracers first, with the previous frame's grid; every K frames they
        # are put back into that grid's cell order first (they drift slowly)
        if self.grid is not None:
            g, p = self.grid
            # sorts before frames 1, 1 + K, 1 + 2K, ... (the first as soon as
            # a grid exists, so no timed frame runs on never-sorted tracers)
            if a.sort_tracers > 0 and self.frames % a.sort_tracers == 1 % a.sort_tracers:
                vg.sort_particles_by_cell(g, self.d_tr, self.nt, stream=self.stream)
            vg.advect_tracers(g, p, self.d_tr, None, self.nt, self.dt, 1.0, stream=self.stream)
        self.frames += 1
        # tracer box and vorton box in one pass and one wait
        tmn, tmx = self.sim.find_bounding_boxes(self.d_tr, self.nt)
        self.sim.update_bounding_box(tmn, tmx, True)
        self.sim.update(self.dt, 0, self.d_vel)
        vg.evolve(self.d_pos, self.d_vel, self.n, self.dt, stream=self.stream)
        self.grid = self.sim.velocity_grid()
        self.evals += self.grid[0].capacity

    def solve(self):
        """One more velocity solve on the current particles (nothing moves);
        returns (grid, field pointer, (spread, circ))."""
        vg = self.vg
        tmn, tmx = vg.find_bounding_box(self.d_tr, self.nt, stream=self.stream)
        self.sim.find_bounding_box()
        self.sim.update_bounding_box(tmn, tmx, True)
        self.sim.update(self.dt, 0, None)
        self.stream.sync()
        g, p = self.sim.velocity_grid()
        return g, p, self.sim.spreading()

    def vortons(self):
        return self.d_pos.numpy()

    def tracers(self):
        return self.d_tr.numpy()


class DeviceFrame:
    """The same frame with nothing read back to the host: a one-effect
    vg_batch whose grid DefineShape defines on the device every frame (tracer
    pass with the tracer box, vorton box, UpdateBoundingBox padding +
    DefineShape + NestedGrid layers, tree, near-field walk, vortons
    AssignVelocityFromField + Evolve), bit for bit the host-geometry frame."""

    walk_kernels = ("vg_batch_velocity_grid", "vg_batch_expand")
    tracer_kernel = "vg_batch_tracers"
    sort_prefix = "vg_batch_sort"

    def __init__(self, vg, args, pos, om, radius, tr, stream):
        self.vg, self.args, self.stream, self.dt = vg, args, stream, 1.0 / 30.0
        self.b = vg.Batch([len(pos)], [len(tr)], [radius], cells=0,
                          max_cells=args.graph_max_cells, near_k=args.near_k, order=args.order,
                          policy=args.policy, expand_from=args.expand_from)
        self.b.upload(pos, om, tr)
        self.frames = 0

    def frame(self):
        a = self.args
        if (self.frames > 0 and a.sort_tracers > 0
                and self.frames % a.sort_tracers == 1 % a.sort_tracers):
            self.b.sort_tracers(self.stream)
        self.b.step(self.dt, self.stream)
        self.frames += 1

    def solve(self):
        self.b.solve(self.stream)
        self.stream.sync()
        g, p = self.b.grid_ptr(0)
        return g, p, self.b.mollify(0)

    def vortons(self):
        return self.b.download("vpos")

    def tracers(self):
        return self.b.download("tpos")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--vortons", type=int, default=1 << 20)
    ap.add_argument("--tracers", type=int, default=1 << 22)
    ap.add_argument("--tracer-order", choices=("random", "emitted"), default="random",
                    help="tracer memory order: uniform random, or the reference emitter's "
                         "cell-by-cell jittered lattice (pclOpEmit.cpp:77-172)")
    ap.add_argument("--sort-tracers", type=int, default=8,
                    help="reorder the tracers into the velocity grid's cell order every K frames "
                         "inside the timed frame (vg_sort_particles_by_cell, with --batched "
                         "vg_batch_sort_tracers; 0 = never)")
    ap.add_argument("--policy", type=int, default=1, help="0 reference rsqrt, 1 accurate")
    ap.add_argument("--near-k", type=int, default=3,
                    help="near-field radius of the accurate treecode; -1 = reference opening rule")
    ap.add_argument("--order", type=int, default=1, help="cluster order: 0 monopole, 1 dipole")
    ap.add_argument("--expand-from", type=int, default=2,
                    help="near-field treecode: first layer applied through brick-centred local "
                         "expansions (-1 = none)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-correction leg of the default line (training + 3 "
                         "solves at its near radius, against the analytic treecodes)")
    ap.add_argument("--learned-near-k", type=int, default=1)
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--suite", action="store_true",
                    help="per-kernel records of the direct, treecode+tracer and multigrid "
                         "workloads instead of the frame line")
    ap.add_argument("--mlp", action="store_true",
                    help="learned far-field correction on the monopole near-field treecode "
                         "(trained on the GPU before the timed frames)")
    ap.add_argument("--batched", action="store_true",
                    help="batched effects: 256 independent effects, one recorded frame launch")
    ap.add_argument("--effects", type=int, default=256)
    ap.add_argument("--batch-cells", type=int, default=32,
                    help="cells per axis of every effect's grid; 0 = DefineShape per effect on "
                         "the device every frame")
    ap.add_argument("--batch-max-cells", type=int, default=64,
                    help="with --batch-cells 0: at most this many cells per axis")
    ap.add_argument("--graph-max-cells", type=int, default=128,
                    help="--graph-frame: storage bound of the device-defined grid, cells per "
                         "axis (DefineShape gives 128^3 for the default scene)")
    ap.add_argument("--batch-near-k", type=int, default=2)
    ap.add_argument("--batch-expand-from", type=int, default=2)
    ap.add_argument("--autotune-json", default=None,
                    help="write the batched autotuner's choice and trials to this file")
    ap.add_argument("--batch-target", type=float, default=2.5e-2,
                    help="accuracy target (rel-L2 vs FP64 direct) of the batched autotuner; "
                         "<= 0 keeps --batch-near-k / --batch-expand-from")
    ap.add_argument("--host-geometry", action="store_true",
                    help="run the frame through vg_sim (bounding boxes read back, DefineShape on "
                         "the host) instead of the device-geometry frame")
    ap.add_argument("--no-host-compare", action="store_true",
                    help="skip re-running the frames with host geometry for the comparison record")
    ap.add_argument("--graph-frame", action="store_true",
                    help="the default frame workload (1M vortons, 4M tracers) as one effect of the "
                         "batched path: bounding box, grid definition, tree, walk and Euler steps "
                         "on the device, replayed as one hipGraph launch per frame")
    ap.add_argument("--copy-probe", action="store_true",
                    help="one 256 MiB device-to-device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.suite:
        suite(args)
        return
    if args.copy_probe:
        copy_probe()
        return
    if args.batched:
        batched(args)
        return
    if args.graph_frame:
        graph_frame(args)
        return

    from vortegrid_amd import scenarios, vg

    vg.require_gpu()
    dt = 1.0 / 30.0
    log(f"generating {args.vortons} vortons (random rings) + {args.tracers} tracers")
    pos, om, radius = scenarios.random_rings(args.vortons, seed=5)
    mn, mx = pos.min(0), pos.max(0)
    tr = (scenarios.tracers_emitted(args.tracers, mn, mx, seed=6)
          if args.tracer_order == "emitted" else scenarios.tracers_in_box(args.tracers, mn, mx, seed=6))
    stream = vg.Stream()
    # the device-geometry frame runs the near-field treecode; the reference
    # rule and the learned correction go through vg_sim
    device = not (args.host_geometry or args.mlp or args.near_k < 0)
    eng = (DeviceFrame if device else HostFrame)(vg, args, pos, om, radius, tr, stream)
    mlp_info = None
    if args.mlp:
        if args.near_k < 0:
            raise SystemExit("--mlp needs --near-k >= 0")
        from vortegrid_amd import learned

        log("training the far-field correction (untimed)")
        mlp, mlp_info = learned.train_far_field(near_k=args.near_k)
        log(f"trained: {mlp_info}")
        eng.sim.set_mlp(mlp)

    log(f"warmup ({'device' if device else 'host'} geometry)")
    for _ in range(args.warmup):
        eng.frame()
    stream.sync()
    if device:
        g_start = eng.b.desc(0) if eng.frames else None
    else:
        eng.evals = 0
    # HIP events bracket only the launches the record reports (the solve's
    # walk and expansions, the tracer pass, the tracer sort): an event pair
    # around every launch of the frame costs ~0.25 ms per frame
    timed = list(eng.walk_kernels) + [eng.tracer_kernel, eng.sort_prefix]
    if args.mlp:
        timed.append("vg_tree_velocity_grid_mlp")
    elif args.near_k < 0:
        timed.append("vg_tree_velocity_grid")
    vg.profile_select(timed)
    vg.profile_reset()
    vg.profile_enable(True)
    log(f"timing {args.steps} frames")
    stream.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.frame()
    stream.sync()
    t1 = time.perf_counter()
    vg.profile_enable(False)
    prof = vg.profile_dump()
    vg.profile_select(None)
    elapsed = t1 - t0
    ms_per_frame = 1e3 * elapsed / args.steps
    shape_note = None
    if device:
        g_end = eng.b.desc(0)
        # the grid is defined on the device each frame; this scene keeps one
        # shape, which the first and last timed frames confirm
        steady = g_start is not None and tuple(g_start.np) == tuple(g_end.np)
        value = g_end.capacity * args.steps / elapsed
        shape_note = {"before_timed_frames": list(g_start.np) if g_start is not None else None,
                      "last_timed_frame": list(g_end.np),
                      "steady": steady, "clamped_grids": eng.b.clamped()}
        snapshot = (eng.vortons(), eng.tracers(), eng.frames)
    else:
        value = eng.evals / elapsed
    log(f"{ms_per_frame:.3f} ms/frame, {value:.4g} evals/s")
    # every kernel of a few more frames, each launch bracketed by events (a
    # breakdown, not the timed frames)
    vg.profile_reset()
    vg.profile_enable(True)
    for _ in range(4):
        eng.frame()
    stream.sync()
    vg.profile_enable(False)
    breakdown = vg.profile_dump()

    # ---- one more velocity solve on the current vortons (no Evolve), so the
    # grid, the interaction count, the FP64 reference and the CPU leg all see
    # the same state
    g, gptr, (spread, circ) = eng.solve()
    cur = eng.vortons()
    d_cur = vg.DeviceArray.from_numpy(cur)
    d_om = vg.DeviceArray.from_numpy(om)
    v = vg.vortons_view(d_cur, d_om, len(cur), uniform_size=2 * radius)
    tr_now = eng.tracers()
    tmn, tmx = tr_now.min(0), tr_now.max(0)

    # ---- roofline of the dominant kernel (treecode at grid points): kernel
    # time averaged over the timed launches, interactions counted on the final
    # frame's tree (the count drifts by well under 1% frame to frame)
    t = vg.Tree()
    t.build(g, v, stream)
    peak, note = FP32_PEAK_TFLOPS, "FP32 VALU kernel (no MFMA); peak is the published FP32 vector peak"
    if args.mlp:
        kname = "vg_tree_velocity_grid_mlp"
        clusters, _, _ = t.mac_work(g, args.near_k, -1, stream)
        pairs = expansions = 0
        flop_cluster = MLP_FLOP
        peak, note = BF16_PEAK_TFLOPS, "BF16 MFMA with FP32 accumulation; dense BF16 peak"
    elif args.near_k < 0:
        kname = "vg_tree_velocity_grid"
        clusters, pairs, expansions = t.interaction_count(g, stream), 0, 0
        flop_cluster = FLOP_PER_PAIR
    else:
        kname = eng.walk_kernels[0]
        clusters, pairs, expansions = t.mac_work(g, args.near_k, args.expand_from, stream)
        flop_cluster = FLOP_PER_PAIR + (FLOP_DIPOLE if args.order else 0.0)
    flop_expand = FLOP_EXPAND_MONO + (FLOP_EXPAND_DIPOLE if args.order else 0.0)
    flops = flop_cluster * clusters + FLOP_PER_PAIR * pairs + flop_expand * expansions
    if expansions:
        flops += FLOP_EXPAND_EVAL * g.capacity
    cnt, ms = prof.get(kname, (0, 0.0))
    if kname == eng.walk_kernels[0] and expansions:
        # the expansion launch belongs to the same solve
        ms += prof.get(eng.walk_kernels[1], (0, 0.0))[1]
        kname = f"{eng.walk_kernels[0]} + {eng.walk_kernels[1]}"
    avg_s = (ms / max(cnt, 1)) * 1e-3
    achieved = flops / avg_s / 1e12 if avg_s > 0 else 0.0
    roofline = {"kernel": kname, "bound": "mfma" if args.mlp else "fp32", "achieved": achieved,
                "peak": peak,
                "unit": "TFLOP/s", "frac": achieved / peak,
                "traffic": measured_traffic(args),
                "avg_launch_ms": ms / max(cnt, 1), "cluster_interactions_per_launch": clusters,
                "direct_pairs_per_launch": pairs, "flop_per_cluster": flop_cluster,
                "clusters_expanded_per_launch": expansions, "flop_per_expansion": flop_expand,
                "flop_per_pair": FLOP_PER_PAIR, "note": note}
    # tracer advection (HBM bound): 24 B per tracer (pos read + write)
    acnt, ams = prof.get(eng.tracer_kernel, (0, 0.0))
    adv = None
    if acnt:
        adv_s = ams / acnt * 1e-3
        gbs = 24.0 * len(tr) / adv_s / 1e9
        sort_ms = sum(m for k, (c, m) in prof.items() if k.startswith(eng.sort_prefix))
        adv = {"kernel": eng.tracer_kernel + (" (advection + the tracer bounding box)"
                                              if device else ""),
               "tracer_order": args.tracer_order + (f", cell-sorted every {args.sort_tracers} "
                                                    f"frames" if args.sort_tracers > 0 else ""),
               "avg_launch_ms": ams / acnt, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS,
               "sort_ms_per_frame": sort_ms / args.steps}
        # the same launch on the same number of tracers in other memory
        # orders (the gathers of spatially ordered tracers hit the caches)
        for other in ("random", "emitted"):
            tr_o = (scenarios.tracers_emitted(len(tr), mn, mx, seed=6) if other == "emitted"
                    else scenarios.tracers_in_box(len(tr), mn, mx, seed=6))
            d_tro = vg.DeviceArray.from_numpy(tr_o)
            vg.profile_reset()
            vg.profile_enable(True)
            for _ in range(10):
                vg.advect_tracers(g, gptr, d_tro, None, len(tr_o), dt, 1.0, stream=stream)
            stream.sync()
            vg.profile_enable(False)
            ocnt, oms = vg.profile_dump().get("vg_advect_tracers", (0, 0.0))
            if ocnt:
                ogbs = 24.0 * len(tr_o) / (oms / ocnt * 1e-3) / 1e9
                adv[f"{other}_order_unsorted"] = {"avg_launch_ms": oms / ocnt, "achieved_GBs": ogbs,
                                                  "frac": ogbs / HBM_PEAK_GBS}
            del d_tro

    # ---- accuracy: treecode vs FP64 direct on a 1% sample of grid points
    rel = None
    sample = None
    if not args.no_accuracy:
        cap = g.capacity
        rng = np.random.default_rng(1234)
        idx = np.sort(rng.choice(cap, max(1, cap // 100), replace=False))
        gp = vg.fetch(gptr, 3 * cap).reshape(-1, 3)
        # the sample's query positions exactly as the kernel forms them
        nudge = np.float32(1.0) - np.float32(2.0) * np.float32(np.finfo(np.float32).eps)
        ix = idx % g.np[0]
        iy = (idx // g.np[0]) % g.np[1]
        iz = idx // (g.np[0] * g.np[1])
        q = np.stack([np.float32(g.min[a]) + np.asarray(i, np.float32) * (np.float32(g.cell[a]) * nudge)
                      for a, i in enumerate((ix, iy, iz))], 1).astype(np.float32)
        d_q = vg.DeviceArray.from_numpy(q)
        d_ref = vg.DeviceArray.zeros((len(q), 3), np.float64)
        tq = time.perf_counter()
        vg.direct_velocity_f64(v, d_q, len(q), d_ref, spread, circ, stream=stream)
        stream.sync()
        ref = d_ref.numpy()
        got = gp[idx].astype(np.float64)
        rel = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        log(f"rel-L2 vs FP64 direct on {len(q)} grid points: {rel:.4e} "
            f"(fp64 direct {time.perf_counter() - tq:.2f}s)")
        sample = (idx, ref)

    # ---- learned far-field correction at its own near radius, against the
    # analytic treecodes it sits between (dipole = its fallback, monopole),
    # on the same state and the same FP64 sample
    learned_rec = mlp_info
    if not args.mlp and not args.no_learned and not args.no_accuracy and args.near_k >= 0:
        try:
            learned_rec = learned_leg(args, vg, v, g, tmn, tmx, sample, stream,
                                      main_solve_ms=roofline["avg_launch_ms"], main_rel=rel)
        except Exception as e:  # the learned leg never blocks the headline
            log(f"learned leg skipped: {e!r}")

    # ---- CPU baseline: the oracle restatement of the same frame, timed on
    # bounded samples and scaled to the whole frame
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(args, g, (spread, circ), cur, om, radius, tr, dt,
                               vg.fetch(gptr, 3 * g.capacity))
            log(f"cpu baseline {cpu['ms_per_frame']:.1f} ms/frame on {cpu['cores']} threads, "
                f"{cpu['one_thread']['ms_per_frame']:.1f} ms/frame on 1; reference rule "
                f"{cpu['reference_rule']['ms_per_frame']:.1f} ms/frame")
        except Exception as e:  # the CPU leg never blocks the GPU number
            log(f"cpu baseline skipped: {e!r}")

    # ---- the same frames with the host defining the grid (vg_sim, the
    # bounding boxes read back every frame): timed, and checked bit for bit
    # against the device-geometry frames on the same input
    host_rec = None
    if device and not args.no_host_compare:
        try:
            del d_cur
            h = HostFrame(vg, args, pos, om, radius, tr, stream)
            vs, ts, nframes = snapshot
            for _ in range(nframes - args.steps):
                h.frame()
            stream.sync()
            th = time.perf_counter()
            for _ in range(args.steps):
                h.frame()
            stream.sync()
            hms = (time.perf_counter() - th) * 1e3 / args.steps
            same_v = bool(np.array_equal(h.vortons(), vs))
            same_t = bool(np.array_equal(h.tracers(), ts))
            host_rec = {"ms_per_step": hms, "frames": nframes,
                        "vortons_bit_identical": same_v, "tracers_bit_identical": same_t,
                        "note": "vg_sim frames (host UpdateBoundingBox + DefineShape after a "
                                "readback of the two bounding boxes) on the same input and "
                                "sort schedule, timed without per-kernel events"}
            log(f"host-geometry frame {hms:.3f} ms/frame; after {nframes} frames vortons "
                f"{'equal' if same_v else 'DIFFER'}, tracers {'equal' if same_t else 'DIFFER'}")
            del h
        except Exception as e:  # the comparison never blocks the headline
            log(f"host-geometry comparison skipped: {e!r}")

    out = {
        "metric": "velocity evals/s + ms/frame at 1M vortons; rel-L2 vs FP64 direct",
        "value": value,
        "unit": "velocity evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_frame,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": ("treecode frame (TREE technique, " + (
                       "reference opening rule" if args.near_k < 0 else
                       f"near-field treecode near_k={args.near_k} monopole clusters + learned "
                       f"correction" if args.mlp else
                       f"near-field treecode near_k={args.near_k} "
                       f"{'dipole' if args.order else 'monopole'} clusters" +
                       (f", local expansions from layer {args.expand_from}"
                        if args.expand_from >= 1 else "")) +
                       ", auto-mollified) + 4M-tracer advection" +
                       ("; bounding boxes, DefineShape and every stage on the device, one "
                        "stream, no host round trip" if device else
                        "; bounding boxes read back, DefineShape on the host")),
                   "frame_geometry": "device" if device else "host",
                   "vortons": len(pos), "tracers": len(tr), "grid": list(g.shape),
                   "scenario": "64 random rings seed 5", "rsqrt_policy": args.policy,
                   "near_k": args.near_k, "order": args.order, "expand_from": args.expand_from,
                   "parallelism": "single GPU"},
        "rel_l2_vs_fp64_direct": rel,
        "grid_shape": shape_note,
        "host_geometry_frame": host_rec,
        "learned_correction": learned_rec,
        "roofline": roofline,
        "tracer_advection": adv,
        "cpu_baseline": cpu,
        "kernels_ms": {k: {"count": c, "total_ms": m} for k, (c, m) in prof.items()},
        "kernels_ms_4_frames_all_timed": {k: {"count": c, "total_ms": m}
                                          for k, (c, m) in breakdown.items()},
    }
    print(json.dumps(out))


def learned_leg(args, vg, v, g, tmn, tmx, sample, stream, main_solve_ms, main_rel):
    """Trains the far-field network on the ring, colliding-ring and plume
    distributions (untimed), then times the velocity solve on the frame's
    final vortons three ways at near radius K: monopole clusters, dipole
    clusters (the network's fallback), dipole + the learned correction."""
    from vortegrid_amd import learned

    k = args.learned_near_k
    log(f"learned leg: training the far-field correction for near_k={k}")
    mlp, info = learned.train_far_field(near_k=k)
    idx, ref = sample
    out = {"near_k": k, "training": info,
           "main_path": {"solve_ms": main_solve_ms, "rel_l2_vs_fp64_direct": main_rel}}
    for name, order, net in (("monopole", 0, None), ("dipole", 1, None), ("learned", 1, mlp)):
        sim = vg.Sim(vg.sim_config(technique=vg.TECH_TREE, rsqrt_policy=args.policy,
                                   tree_near_k=k, tree_order=order, tree_expand_from=-1), stream)
        sim.set_vortons(v)
        if net is not None:
            sim.set_mlp(net)
        times = []
        for rep in range(3):
            sim.find_bounding_box()
            sim.update_bounding_box(tmn, tmx, True)
            stream.sync()
            t0 = time.perf_counter()
            sim.update(1.0 / 30.0, 0, None)
            stream.sync()
            times.append((time.perf_counter() - t0) * 1e3)
        g2, p2 = sim.velocity_grid()
        assert tuple(g2.np) == tuple(g.np)
        got = vg.fetch(p2, 3 * g2.capacity).reshape(-1, 3)[idx].astype(np.float64)
        err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        out[name] = {"update_ms": min(times), "rel_l2_vs_fp64_direct": err}
        log(f"learned leg: {name:8s} k={k}: update {min(times):.2f} ms, rel-L2 {err:.3e}")
    out["note"] = ("update_ms is the whole vg_sim_update (tree build + velocity grid) on the "
                   "frame's final vortons; main_path.solve_ms is the default path's velocity "
                   "solve launches alone")
    return out


def cpu_threads():
    """Threads for the CPU legs: the cores this process may run on (its
    affinity set), at most 16, the host share of one GPU on the machines
    this runs on."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(args, g, mollify, cur, om, radius, tr, dt, grid_field):
    """The oracle restatement of one frame on the host cores: tree build
    (serial in the oracle, as MakeBaseVortonGrid is), the velocity grid by
    the same method as the GPU line on a band of bricks (all threads, and a
    smaller band on one thread), the reference opening rule on a band of z
    slices (all threads and one thread), and advection of a tracer subset;
    each timed part is scaled to the whole frame.  About 10-30 s in all."""
    from oracle import oracle as O

    O.build()
    cores = cpu_threads()
    spread, circ = mollify
    og = O.Grid()
    for k in ("min", "extent", "cell", "cpe", "np"):
        for a in range(3):
            getattr(og, k)[a] = getattr(g, k)[a]
    cap = g.capacity
    nx, ny, nz = (int(v) for v in g.np)
    size = 2 * radius
    tb = time.perf_counter()
    ot = O.InfluenceTree(og, cur, om, size)
    t_build = time.perf_counter() - tb
    ntr = min(len(tr), 1 << 20)
    ta = time.perf_counter()
    O.advect(og, grid_field, tr[:ntr], np.zeros((ntr, 3), np.float32), dt, 1.0)
    t_adv = (time.perf_counter() - ta) * (len(tr) / ntr)

    def timed(fn):
        t0 = time.perf_counter()
        fn()
        return time.perf_counter() - t0

    # reference opening rule on the lowest z slices
    slices = max(1, min(nz, 16))
    sub = O.Grid()
    for k in ("min", "extent", "cell", "cpe", "np"):
        for a in range(3):
            getattr(sub, k)[a] = getattr(og, k)[a]
    sub.np[2] = slices
    ref_pts = nx * ny * slices
    ref_all = timed(lambda: O.velocity_grid(ot, sub, cur, om, size, 1, spread, circ, args.policy,
                                            threads=cores)) * cap / ref_pts
    sub.np[2] = 2
    ref_one = timed(lambda: O.velocity_grid(ot, sub, cur, om, size, 1, spread, circ, args.policy,
                                            threads=1)) * cap / (nx * ny * 2)
    rec = {"unit": "velocity evals/s", "kind": "port"}
    if args.near_k >= 0:
        # the GPU line's method on one z row of bricks a quarter of the way up
        # the grid, on all threads and on one thread
        nbx = (nx + 3) // 4
        bz0 = (nz + 3) // 8
        row_pts = nx * ny * 4
        walk = lambda th: O.mac_velocity_grid_box(ot, og, cur, om, size, args.near_k, args.order,
                                                  args.expand_from, spread, circ, args.policy, th,
                                                  (0, nbx), (bz0, bz0 + 1))
        w_all = timed(lambda: walk(cores)) * cap / row_pts
        w_one = timed(lambda: walk(1)) * cap / row_pts
        method = (f"near-field treecode near_k={args.near_k} order={args.order} expand_from="
                  f"{args.expand_from} on z brick row {bz0} ({row_pts} points)")
    else:
        w_all, w_one = ref_all, ref_one
        method = "reference opening rule"
    frame_all = t_build + w_all + t_adv
    frame_one = t_build + w_one + t_adv
    rec.update({
        "value": cap / frame_all, "cores": cores, "host_cpus": os.cpu_count(),
        "ms_per_frame": frame_all * 1e3,
        "one_thread": {"value": cap / frame_one, "ms_per_frame": frame_one * 1e3},
        "reference_rule": {"value": cap / (t_build + ref_all + t_adv),
                           "ms_per_frame": (t_build + ref_all + t_adv) * 1e3,
                           "one_thread_ms_per_frame": (t_build + ref_one + t_adv) * 1e3,
                           "velocity_grid_ms": ref_all * 1e3,
                           "velocity_grid_one_thread_ms": ref_one * 1e3},
        "parts_ms": {"tree_build_1_thread": t_build * 1e3, "velocity_grid": w_all * 1e3,
                     "velocity_grid_1_thread": w_one * 1e3, "advection_1_thread": t_adv * 1e3},
        "sample": (f"oracle on the frame's final state: tree build on all {len(cur)} vortons "
                   f"(serial), {method}, reference rule on {slices} of {nz} z-slices "
                   f"({cores} threads) and 2 slices (1 thread), advection of {ntr} of {len(tr)} "
                   f"tracers (serial); each scaled to the full {nx}x{ny}x{nz} frame; {cores} "
                   f"threads = this process's CPU share (affinity, at most 16)"),
    })
    return rec


# HBM-side bytes per launch of the dominant kernel, from the two counter
# runs (FETCH_SIZE, WRITE_SIZE) summarised under profiles/ for the default
# workload; None for other settings.
TRAFFIC_FILES = {(3, 1, 2, 1 << 20): "tree_mac_k3_e2_1M_traffic.txt"}


def measured_traffic(args):
    if args.mlp:
        return None
    name = TRAFFIC_FILES.get((args.near_k, args.order, args.expand_from, args.vortons))
    path = os.path.join(ROOT, "profiles", name) if name else None
    if not path or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def effect_fields(n_effects, seed=4):
    """Vorton counts log-uniform in [512, 32768], each a ring or a plume
    from its own seed, placed on a lattice; 16 tracers per vorton."""
    from vortegrid_amd import scenarios

    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(np.log(512), np.log(32768), n_effects)).astype(np.int64)
    vp, vw, tp, radii = [], [], [], []
    for i, n in enumerate(counts):
        gen = scenarios.vortex_ring if rng.random() < 0.5 else scenarios.jet_plume
        pos, om, radius = gen(int(n), seed=1000 + i)
        pos = pos + np.float32(8.0) * np.float32([i % 16, i // 16, 0])
        vp.append(pos)
        vw.append(om)
        tp.append(scenarios.tracers_in_box(16 * int(n), pos.min(0), pos.max(0), seed=5000 + i))
        radii.append(radius)
    return vp, vw, tp, radii


def batched(args):
    from vortegrid_amd import vg

    vg.require_gpu()
    dt = np.float32(1.0 / 30.0)
    log(f"generating {args.effects} effects")
    vp, vw, tp, radii = effect_fields(args.effects)
    nk = args.batch_near_k
    b = vg.Batch([len(p) for p in vp], [len(t) for t in tp], radii, cells=args.batch_cells,
                 near_k=nk, order=args.order, policy=args.policy,
                 expand_from=args.batch_expand_from, max_cells=args.batch_max_cells)
    b.upload(np.concatenate(vp), np.concatenate(vw), np.concatenate(tp))
    s = vg.Stream()
    b.step(dt, s)
    s.sync()
    tuned = None
    if args.batch_target > 0:
        from vortegrid_amd import autotune

        tuned = autotune.tune_batch(b, np.concatenate(vw), radii, target=args.batch_target,
                                    order=args.order, stream=s, path=args.autotune_json)
        nk = tuned["method"]["near_k"]
        args.batch_expand_from = tuned["method"]["expand_from"]
        log("autotune: " + json.dumps({k: tuned[k] for k in ("method", "launch", "chosen")}))
    else:
        log("autotune:\n" + b.autotune(s))
    # every K frames each effect's tracers go back into its grid's cell order
    # (inside the timed loop; the first call allocates, so it comes first)
    K = args.sort_tracers
    if K > 0:
        b.sort_tracers(s)
    b.capture(dt, s)
    for i in range(args.warmup):
        if K > 0 and i % K == 0:
            b.sort_tracers(s)
        b.launch(s)
    s.sync()
    vg.profile_reset()
    t0 = time.perf_counter()
    for i in range(args.steps):
        if K > 0 and i % K == 0:
            b.sort_tracers(s)
        b.launch(s)
    s.sync()
    el = time.perf_counter() - t0
    ms = el * 1e3 / args.steps
    pts = sum(b.desc(e).capacity for e in range(args.effects))
    log(f"{ms:.3f} ms/frame for {args.effects} effects ({b.nv} vortons, {b.nt} tracers, "
        f"{pts} grid points)")
    # per-kernel times of one eager frame, and of one tracer sort
    vg.profile_enable(True)
    b.step(dt, s)
    s.sync()
    vg.profile_enable(False)
    prof = vg.profile_dump()
    sort_ms = None
    if K > 0:
        vg.profile_reset()
        vg.profile_enable(True)
        b.sort_tracers(s)
        s.sync()
        vg.profile_enable(False)
        sort_ms = sum(m for k, (c, m) in vg.profile_dump().items() if k.startswith("vg_batch_sort"))
        log(f"tracer sort {sort_ms:.3f} ms (every {K} frames)")
    # accuracy on 8 effects: 1% of their grid points vs the FP64 direct sum
    # over the effect's current vortons (the grid of the last frame was
    # computed before its Evolve, so re-run one frame with dt = 0 first)
    b.step(np.float32(0.0), s)
    s.sync()
    cur = b.download("vpos")
    off = np.cumsum([0] + [len(p) for p in vp])
    rng = np.random.default_rng(3)
    num = den = 0.0
    for e in rng.choice(args.effects, 8, replace=False):
        g, field = b.grid(int(e))
        spread, circ = b.mollify(int(e))
        idx = np.sort(rng.choice(g.capacity, max(1, g.capacity // 100), replace=False))
        nudge = np.float32(1) - np.float32(2) * np.float32(np.finfo(np.float32).eps)
        ijk = (idx % g.np[0], (idx // g.np[0]) % g.np[1], idx // (g.np[0] * g.np[1]))
        q = np.stack([np.float32(g.min[a]) + np.asarray(i, np.float32) * (np.float32(g.cell[a]) * nudge)
                      for a, i in enumerate(ijk)], 1).astype(np.float32)
        dp = vg.DeviceArray.from_numpy(cur[off[e]:off[e + 1]])
        dw = vg.DeviceArray.from_numpy(vw[e])
        v = vg.vortons_view(dp, dw, len(vw[e]), uniform_size=2 * radii[e])
        ref = vg.DeviceArray.zeros((len(q), 3), np.float64)
        vg.direct_velocity_f64(v, vg.DeviceArray.from_numpy(q), len(q), ref, spread, circ)
        vg.synchronize()
        r = ref.numpy()
        num += float(np.sum((field[idx] - r) ** 2))
        den += float(np.sum(r ** 2))
    rel = (num / den) ** 0.5
    log(f"rel-L2 vs FP64 direct (8 effects, 1% of their grid points): {rel:.3e}")
    cnt, wms = prof.get("vg_batch_velocity_grid", (1, 0.0))
    tcnt, tms = prof.get("vg_batch_tracers", (1, 0.0))
    # CPU baseline: the oracle on 8 sampled effects (tree build, the same
    # near-field walk with expansions, vorton and tracer advection), scaled to
    # all effects by vortons (build, vorton advection), tracers (tracer
    # advection) and effect count (walk on the fixed-size grids)
    cpu = None
    if not args.no_cpu:
        try:
            from oracle import oracle as O

            O.build()
            cores = cpu_threads()
            tpos = b.download("tpos")
            toff = np.cumsum([0] + [len(t) for t in tp])
            t_build = t_walk = t_adv = 0.0
            nv_s = nt_s = 0
            sample = rng.choice(args.effects, 8, replace=False)
            for e in sample:
                g, field = b.grid(int(e))
                spread, circ = b.mollify(int(e))
                og = O.Grid()
                for k in ("min", "extent", "cell", "cpe", "np"):
                    for a in range(3):
                        getattr(og, k)[a] = getattr(g, k)[a]
                pe = cur[off[e]:off[e + 1]]
                t0 = time.perf_counter()
                ot = O.InfluenceTree(og, pe, vw[e], 2 * radii[e])
                t1 = time.perf_counter()
                vf = O.mac_velocity_grid_ex(ot, og, pe, vw[e], 2 * radii[e], nk, args.order,
                                            args.batch_expand_from, spread, circ, args.policy,
                                            cores)
                t2 = time.perf_counter()
                O.advect(og, vf, pe, np.zeros_like(pe), float(dt), 1.0)
                O.advect(og, vf, tpos[toff[e]:toff[e + 1]],
                         np.zeros((toff[e + 1] - toff[e], 3), np.float32), float(dt), 1.0)
                t3 = time.perf_counter()
                t_build += t1 - t0
                t_walk += t2 - t1
                t_adv += t3 - t2
                nv_s += len(pe)
                nt_s += toff[e + 1] - toff[e]
            frame_s = (t_build * b.nv / nv_s + t_walk * args.effects / len(sample) +
                       t_adv * (b.nv + b.nt) / (nv_s + nt_s))
            cpu = {"value": pts / frame_s, "unit": "velocity evals/s", "cores": cores,
                   "kind": "port", "ms_per_frame": frame_s * 1e3,
                   "sample": (f"oracle on {len(sample)} of {args.effects} effects: tree build "
                              f"(1 thread), near-field walk ({cores} threads), vorton + tracer "
                              f"advection (1 thread); scaled by vortons, effects and particles")}
            log(f"cpu baseline {cpu['ms_per_frame']:.1f} ms/frame")
        except Exception as ex:  # the CPU leg never blocks the GPU number
            log(f"cpu baseline skipped: {ex!r}")
    out = {
        "metric": "batched effects: ms/frame and grid velocity evals/s, one recorded frame launch",
        "value": pts / (ms * 1e-3), "unit": "velocity evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.effects} effects, N log-uniform [512, 32768], rings or "
                               f"plumes, 16 tracers per vorton (uniform in each effect's box"
                               + (f", cell-sorted every {K} frames" if K > 0 else "") + "), "
                               + (f"{args.batch_cells}^3-cell grid each" if args.batch_cells else
                                  f"grids from DefineShape on the device (at most "
                                  f"{args.batch_max_cells} cells per axis)") +
                               f", near-field "
                               f"treecode near_k={nk} {'dipole' if args.order else 'monopole'}"
                               + (f", local expansions from layer {args.batch_expand_from}"
                                  if args.batch_expand_from >= 1 else ""),
                   "vortons": b.nv, "tracers": b.nt, "effects": args.effects,
                   "grid_points": pts, "clamped_grids": b.clamped()},
        "rel_l2_vs_fp64_direct": rel,
        "launch_shapes": dict(zip(("walk_waves", "advect_block", "advect_items"), b.get_launch())),
        "autotune": tuned,
        "walk_ms": wms / max(cnt, 1), "tracers_ms": tms / max(tcnt, 1),
        "tracer_sort": ({"every_frames": K, "ms": sort_ms} if K > 0 else None),
        "tracers_GBs": 24.0 * b.nt / (tms / max(tcnt, 1) * 1e-3) / 1e9,
        "cpu_baseline": cpu,
        "kernels_ms_one_frame": {k: {"count": c, "total_ms": m} for k, (c, m) in prof.items()},
    }
    print(json.dumps(out))


def copy_probe():
    """A plain copy of known size: 256 MiB read + 256 MiB written.  Profiled
    with FETCH_SIZE / WRITE_SIZE it gives the factor between those counters
    and bytes moved to and from HBM."""
    import torch

    n = 64 << 20
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_each_way": 4 * n}))


def _timed(vg, name, fn, reps):
    """Run fn reps times under the HIP-event profiler; average ms of kernel name."""
    fn()
    vg.synchronize()
    vg.profile_reset()
    vg.profile_enable(True)
    for _ in range(reps):
        fn()
    vg.synchronize()
    vg.profile_enable(False)
    prof = vg.profile_dump()
    cnt, ms = prof.get(name, (0, 0.0))
    return ms / max(cnt, 1), prof


def graph_frame(args):
    """The default frame's workload with no host round trip: one effect of the
    batched path whose grid DefineShape defines on the device every frame
    (the same frame as the default one, bit for bit: tested), eager and as a
    captured hipGraph."""
    from vortegrid_amd import scenarios, vg

    vg.require_gpu()
    pos, om, radius = scenarios.random_rings(args.vortons, seed=5)
    tr = scenarios.tracers_in_box(args.tracers, pos.min(0), pos.max(0), seed=6)
    b = vg.Batch([len(pos)], [len(tr)], [radius], cells=0, max_cells=args.graph_max_cells,
                 near_k=args.near_k, order=args.order, policy=args.policy,
                 expand_from=args.expand_from)
    b.upload(pos, om, tr)
    dt = 1.0 / 30.0
    s = vg.Stream()
    K = args.sort_tracers
    b.step(dt, s)  # a first frame: grids to sort the tracers by
    if K > 0:
        b.sort_tracers(s)  # the first call allocates
    res = {}
    for mode in ("eager", "graph"):
        if mode == "graph":
            b.capture(dt, s)
        run = (lambda: b.step(dt, s)) if mode == "eager" else (lambda: b.launch(s))
        for i in range(args.warmup):
            if K > 0 and i % K == 0:
                b.sort_tracers(s)
            run()
        s.sync()
        t0 = time.perf_counter()
        for i in range(args.steps):
            if K > 0 and i % K == 0:
                b.sort_tracers(s)
            run()
        s.sync()
        res[mode] = (time.perf_counter() - t0) * 1e3 / args.steps
    vg.profile_reset()
    vg.profile_enable(True)
    b.step(dt, s)
    s.sync()
    vg.profile_enable(False)
    prof = vg.profile_dump()
    g = b.desc(0)
    log("one eager frame: " + ", ".join(f"{k} {m:.3f}" for k, (c, m) in
                                        sorted(prof.items(), key=lambda kv: -kv[1][1])))
    print(json.dumps({
        "workload": "default frame workload as one batched effect (grid defined on the device by "
                    "DefineShape, no host round trip): 1M vortons (64 random rings), 4M tracers"
                    "%s, near_k=%d, order=%d, expand_from=%d" % (
                        f" cell-sorted every {K} frames" if K > 0 else "", args.near_k,
                        args.order, args.expand_from),
        "grid": list(g.np), "eager_ms_per_frame": res["eager"],
        "graph_ms_per_frame": res["graph"], "steps": args.steps, "clamped_grids": b.clamped(),
        "kernels_ms_one_frame": {k: {"count": c, "total_ms": m} for k, (c, m) in prof.items()}}),
        flush=True)


def suite(args):
    """One JSON line per workload: the direct sum (4096-vorton ring and a
    65,536-vorton pair of rings), the monopole treecode + 1M-tracer frame on
    65^3, and the vector-Poisson solve on 129^3 with per-kernel achieved
    FLOP/s or GB/s (algorithmic bytes: each input read once, each output
    written once)."""
    from vortegrid_amd import scenarios, vg

    vg.require_gpu()
    recs = []

    def gbs(nbytes, ms):
        return nbytes / (ms * 1e-3) / 1e9

    # ---- direct Biot-Savart (K1): 26 flop per pair
    for label, (pos, om, radius), grid_pts in (
            ("direct ring N=4096 (vortons + 33^3 grid)", scenarios.vortex_ring(4096, seed=1), True),
            ("direct rings N=65536 (at vortons)", scenarios.colliding_rings(65536, seed=2), False)):
        n = len(pos)
        dp, dw = vg.DeviceArray.from_numpy(pos), vg.DeviceArray.from_numpy(om)
        v = vg.vortons_view(dp, dw, n, uniform_size=2 * radius)
        out = vg.DeviceArray.zeros(pos.shape)
        ms, _ = _timed(vg, "vg_direct_velocity",
                       lambda: vg.direct_velocity(v, dp, n, out, policy=args.policy), 10)
        pairs = float(n) * n
        rec = {"workload": label, "kernel": "vg_direct_velocity", "avg_launch_ms": ms,
               "pairs": pairs, "pairs_per_s": pairs / (ms * 1e-3),
               "achieved_TFLOPs": FLOP_PER_PAIR * pairs / (ms * 1e-3) / 1e12}
        rec["frac_fp32_peak"] = rec["achieved_TFLOPs"] / FP32_PEAK_TFLOPS
        if grid_pts:
            mn, mx = vg.find_bounding_box(dp, n)
            g = vg.finalize_bounding_box(mn, mx, radius, n)
            g = vg.define_shape(32 ** 3, g.min, [g.min[a] + g.extent[a] for a in range(3)])
            og = vg.DeviceArray.zeros((g.capacity, 3))
            gms, _ = _timed(vg, "vg_direct_velocity_grid",
                            lambda: vg.direct_velocity_grid(v, g, og, policy=args.policy), 10)
            gp = float(n) * g.capacity
            rec.update({"grid": list(g.shape), "grid_launch_ms": gms,
                        "grid_TFLOPs": FLOP_PER_PAIR * gp / (gms * 1e-3) / 1e12})
        recs.append(rec)

    # ---- monopole treecode frame (reference opening rule) + 1M tracers, 65^3
    pos, om, radius = scenarios.colliding_rings(65536, seed=2)
    mn, mx = pos.min(0), pos.max(0)
    tr = scenarios.tracers_in_box(1 << 20, mn, mx, seed=3)
    dp, dw = vg.DeviceArray.from_numpy(pos), vg.DeviceArray.from_numpy(om)
    dvel, dtr = vg.DeviceArray.zeros(pos.shape), vg.DeviceArray.from_numpy(tr)
    v = vg.vortons_view(dp, dw, len(pos), uniform_size=2 * radius)
    sim = vg.Sim(vg.sim_config(technique=vg.TECH_TREE, rsqrt_policy=args.policy))
    sim.set_vortons(v)
    state = {"g": None, "f": None}
    dt = 1.0 / 30.0

    def frame():
        if state["g"] is not None:
            vg.advect_tracers(state["g"], state["f"], dtr, None, len(tr), dt, 1.0)
        tmn, tmx = sim.find_bounding_boxes(dtr, len(tr))
        sim.update_bounding_box(tmn, tmx, True)
        sim.update(dt, 0, dvel)
        vg.evolve(dp, dvel, len(pos), dt)
        state["g"], state["f"] = sim.velocity_grid()

    for _ in range(3):
        frame()
    vg.synchronize()
    steps = 100
    # timed without per-launch events, then 20 more frames with them for the
    # per-kernel figures
    t0 = time.perf_counter()
    for _ in range(steps):
        frame()
    vg.synchronize()
    el = time.perf_counter() - t0
    vg.profile_reset()
    vg.profile_enable(True)
    psteps = 20
    for _ in range(psteps):
        frame()
    vg.synchronize()
    vg.profile_enable(False)
    prof = vg.profile_dump()
    g = state["g"]
    c, ms = prof.get("vg_advect_tracers", (1, 0.0))
    frames_sim = 3 + steps + psteps
    # The same frames on the device: a one-effect batch with the reference
    # opening rule (near_k = -1) defines the grid on the GPU (no readback),
    # eagerly and as a replayed hipGraph; after as many eager frames as the
    # vg_sim run above, vortons and tracers must be bit-identical to it.
    b = vg.Batch([len(pos)], [len(tr)], [radius], cells=0, max_cells=64, near_k=-1,
                 policy=args.policy)
    b.upload(pos, om, tr)
    for _ in range(3):
        b.step(dt)
    vg.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.step(dt)
    vg.synchronize()
    dev_el = time.perf_counter() - t0
    for _ in range(psteps):
        b.step(dt)
    vg.synchronize()
    same = (b.clamped() == 0 and np.array_equal(dp.numpy(), b.download("vpos")) and
            np.array_equal(dtr.numpy(), b.download("tpos")))
    gs = vg.Stream()
    b.capture(dt, gs)
    for _ in range(3):
        b.launch(gs)
    gs.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.launch(gs)
    gs.sync()
    graph_el = time.perf_counter() - t0
    device_frame = {"ms_per_step": dev_el * 1e3 / steps, "graph_ms_per_step": graph_el * 1e3 / steps,
                    "bit_identical_to_vg_sim_after_frames": frames_sim if same else False,
                    "note": "one-effect vg_batch, near_k = -1 (reference opening rule), DefineShape "
                            "on the device; graph: the frame recorded once and replayed"}
    recs.append({"workload": "treecode frame, 2 colliding rings N=65536 + 1M tracers, 100 steps",
                 "device_frame": device_frame,
                 "grid": list(g.shape), "ms_per_step": el * 1e3 / steps,
                 "evals_per_s": g.capacity * steps / el,
                 "tree_grid_avg_ms": prof.get("vg_tree_velocity_grid", (1, 0.0))[1] / psteps,
                 "kernel_sum_ms_per_step": sum(m for (c, m) in prof.values()) / psteps,
                 "advect_avg_ms": ms / max(c, 1),
                 "advect_GBs": gbs(24.0 * len(tr), ms / max(c, 1)),
                 "kernels_ms": {k: {"count": c, "total_ms": m} for k, (c, m) in prof.items()},
                 "kernels_ms_frames": psteps})

    # ---- vector Poisson on 129^3: plume vorticity, treecode faces, V-cycles
    pos, om, radius = scenarios.jet_plume(262144, seed=3)
    n = len(pos)
    dp, dw = vg.DeviceArray.from_numpy(pos), vg.DeviceArray.from_numpy(om)
    v = vg.vortons_view(dp, dw, n, uniform_size=2 * radius)
    mn, mx = vg.find_bounding_box(dp, n)
    g0 = vg.finalize_bounding_box(mn, mx, radius, n)
    g = vg.define_shape(128 ** 3, g0.min, [g0.min[a] + g0.extent[a] for a in range(3)])
    t = vg.Tree()
    t.build(g, v)
    cap = g.capacity
    negw, a0 = vg.DeviceArray.zeros((cap, 3)), vg.DeviceArray.zeros((cap, 3))
    ms_splat, _ = _timed(vg, "vg_splat_vorticity", lambda: t.splat_vorticity(radius, -1.0, negw), 5)
    ms_faces, _ = _timed(vg, "vg_tree_vecpot_faces", lambda: t.vecpot_boundary(g, a0), 5)
    a = vg.DeviceArray.zeros((cap, 3))
    a0_host = a0.numpy()
    a.upload(a0_host)
    mg = vg.Multigrid()
    vg.synchronize()
    # cold: the first solve allocates the hierarchy and touches it first
    t0 = time.perf_counter()
    mg.vcycle_solve(g, a, negw, tol=1e-6, max_cycles=40)
    vg.synchronize()
    cold_ms = (time.perf_counter() - t0) * 1e3
    # warm: the same solver object again, as a simulation reuses it per frame
    a.upload(a0_host)
    vg.synchronize()
    t0 = time.perf_counter()
    rel, cycles = mg.vcycle_solve(g, a, negw, tol=1e-6, max_cycles=40)
    vg.synchronize()
    solve_ms = (time.perf_counter() - t0) * 1e3
    # once more with per-launch event timing (which adds its own overhead)
    a.upload(a0_host)
    vg.synchronize()
    vg.profile_reset()
    vg.profile_enable(True)
    mg.vcycle_solve(g, a, negw, tol=1e-6, max_cycles=40)
    vg.synchronize()
    vg.profile_enable(False)
    mgprof = vg.profile_dump()
    vel = vg.DeviceArray.zeros((cap, 3))
    ms_curl, _ = _timed(vg, "vg_curl", lambda: vg.curl(g, a, vel), 10)
    ms_res, _ = _timed(vg, "vg_residual_norm", lambda: vg.residual_norm(g, a, negw), 10)
    soln = vg.DeviceArray.zeros((cap, 3))
    ms_sor, _ = _timed(vg, "vg_sor_rows", lambda: vg.sor_solve(g, soln, negw, 1), 10)
    lo = vg.GridDesc()
    vg.check(vg.lib().vg_decimate(vg.C.byref(lo), vg.C.byref(g), 2), "decimate")
    flo = vg.DeviceArray.zeros((lo.capacity, 3))
    ms_inj, _ = _timed(vg, "vg_restrict_inject", lambda: vg.downsample_inject(lo, flo, g, a), 10)
    ms_pro, _ = _timed(vg, "vg_prolong_interior", lambda: vg.upsample_interior(g, soln, lo, flo), 10)
    interior = (g.np[0] - 2) * (g.np[1] - 2) * (g.np[2] - 2)
    # the sor_solve launch count per step: one launch per colour
    recs.append({
        "workload": "vector Poisson 129^3, jet plume N=262144, treecode Dirichlet faces, "
                    "V-cycles to 1e-6 (FP64 coarsest level and norms)",
        "grid": list(g.shape), "vcycle_solve_ms": solve_ms, "vcycle_solve_cold_ms": cold_ms,
        "cycles": cycles,
        "final_rel_residual": rel,
        "kernels": {
            "vg_splat_vorticity": {"ms": ms_splat, "GBs": gbs(24.0 * n + 12.0 * cap, ms_splat)},
            "vg_tree_vecpot_faces": {"ms": ms_faces},
            "vg_curl": {"ms": ms_curl, "GBs": gbs(24.0 * cap, ms_curl)},
            "vg_residual_norm": {"ms": ms_res, "GBs": gbs(24.0 * cap, ms_res)},
            "vg_sor_rows (one colour)": {"ms": ms_sor, "GBs": gbs(18.0 * interior, ms_sor)},
            "vg_restrict_inject": {"ms": ms_inj, "GBs": gbs(24.0 * lo.capacity, ms_inj)},
            "vg_prolong_interior": {"ms": ms_pro, "GBs": gbs(12.0 * interior, ms_pro)},
        },
        "vcycle_kernels_ms": {k: {"count": c, "total_ms": m} for k, (c, m) in mgprof.items()},
    })
    # ---- the reference's default technique: a whole POISSON update (tree,
    # vorticity splat, treecode faces, the reference cascade with 16 SOR
    # steps per level visit, curl, vortons' grid velocity) on the plume
    psim = vg.Sim(vg.sim_config(technique=vg.TECH_POISSON, rsqrt_policy=args.policy))
    pvel = vg.DeviceArray.zeros((n, 3))
    psim.set_vortons(v)

    def poisson_update():
        psim.find_bounding_box()
        psim.update_bounding_box(mn, mx, True)
        psim.update(1.0 / 30.0, 0, pvel)

    poisson_update()
    vg.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        poisson_update()
    vg.synchronize()
    upd_ms = (time.perf_counter() - t0) * 1e3 / reps
    vg.profile_reset()
    vg.profile_enable(True)
    poisson_update()
    vg.synchronize()
    vg.profile_enable(False)
    pprof = vg.profile_dump()
    pg = psim.grid_template()
    sc, sms = pprof.get("vg_sor_rows", (0, 0.0))
    interior0 = (pg.np[0] - 2) * (pg.np[1] - 2) * (pg.np[2] - 2)
    recs.append({
        "workload": "POISSON technique update (the reference default): plume N=262144, reference "
                    "cascade (injection, overwrite prolongation, 16 SOR steps per level visit)",
        "grid": list(pg.shape), "update_ms": upd_ms, "sor_rows_launches": sc,
        "sor_rows_ms_total": sms,
        "kernels_ms": {k: {"count": c, "total_ms": m} for k, (c, m) in pprof.items()},
        "note": f"one SOR colour pass at the finest level moves 18 B per interior point "
                f"({interior0} points)"})
    # ---- stretching and tilting: 1M vortons (64 random rings) on their 129^3-class grid
    pos, om, radius = scenarios.random_rings(1 << 20, seed=5)
    n = len(pos)
    dp, dw = vg.DeviceArray.from_numpy(pos), vg.DeviceArray.from_numpy(om)
    mn, mx = vg.find_bounding_box(dp, n)
    g = vg.finalize_bounding_box(mn, mx, radius, n)
    cap = g.capacity
    rng = np.random.default_rng(32)
    ijk = np.stack(np.meshgrid(*[np.arange(g.np[a], dtype=np.float32) for a in (2, 1, 0)],
                               indexing="ij")[::-1], -1).reshape(-1, 3)
    vel = np.zeros((cap, 3), np.float32)
    for _ in range(4):  # a smooth synthetic velocity field
        k = rng.normal(size=3).astype(np.float32) * np.float32(0.2)
        vel += np.sin(ijk @ k + np.float32(rng.uniform(0, 6)))[:, None] * \
            rng.normal(size=3).astype(np.float32)
    dvel = vg.DeviceArray.from_numpy(vel)
    jac = vg.DeviceArray.zeros((cap, 3, 3))
    dt = 1.0 / 30.0
    order = vg.DeviceArray.from_numpy(vg.cell_order(g, pos))  # the influence tree's order
    ms_fused, _ = _timed(vg, "vg_stretch_tilt_fused",
                         lambda: vg.stretch_tilt_from_velocity(g, dvel, dp, dw, n, dt), 20)
    ms_fused_o, _ = _timed(vg, "vg_stretch_tilt_fused",
                           lambda: vg.stretch_tilt_from_velocity(g, dvel, dp, dw, n, dt,
                                                                 order=order), 20)
    ms_jac, _ = _timed(vg, "vg_jacobian", lambda: vg.jacobian(g, dvel, jac), 20)
    ms_interp, _ = _timed(vg, "vg_stretch_tilt", lambda: vg.stretch_tilt(g, jac, dp, dw, n, dt), 20)
    ms_interp_o, _ = _timed(vg, "vg_stretch_tilt",
                            lambda: vg.stretch_tilt(g, jac, dp, dw, n, dt, order=order), 20)
    recs.append({
        "workload": "stretch/tilt: 1M vortons (64 random rings), Jacobian of the velocity grid "
                    "interpolated at each vorton, Omega += 0.5 (Omega J) dt",
        "grid": list(g.shape), "vortons": n,
        "fused_ms": ms_fused,
        "fused_GBs": gbs(36.0 * n + 12.0 * cap, ms_fused),
        "fused_cell_order_ms": ms_fused_o,
        "fused_cell_order_GBs": gbs(40.0 * n + 12.0 * cap, ms_fused_o),
        "two_step_ms": ms_jac + ms_interp,
        "jacobian_ms": ms_jac, "jacobian_GBs": gbs(48.0 * cap, ms_jac),
        "interp_ms": ms_interp, "interp_GBs": gbs(36.0 * n + 36.0 * cap, ms_interp),
        "interp_cell_order_ms": ms_interp_o,
        "bytes_note": "algorithmic: per vorton pos + Omega read, Omega written (36 B, +4 B of "
                      "order index); each grid read once (12 B/point velocity, 36 B/point "
                      "Jacobian)",
    })

    # ---- PSE viscous diffusion on the same 1M vortons (their tree's partition)
    t = vg.Tree()
    v = vg.vortons_view(dp, dw, n, uniform_size=2 * radius)
    t.build(g, v)
    pse_labels = ("vg_pse_gather", "vg_pse_frac_size", "vg_scan_tile", "vg_scan_add",
                  "vg_pse_fractions", "vg_pse_schedule", "vg_pse_phase", "vg_pse_scatter")
    t.pse_diffuse(dw, 0.01, dt)
    vg.synchronize()
    vg.profile_reset()
    vg.profile_enable(True)
    reps = 5
    for _ in range(reps):
        t.pse_diffuse(dw, 0.01, dt)
    vg.synchronize()
    vg.profile_enable(False)
    prof = vg.profile_dump()
    pse = {k: prof.get(k, (0, 0.0))[1] / reps for k in pse_labels}
    vg.profile_reset()
    vg.profile_enable(True)
    for _ in range(reps):
        t.pse_diffuse(dw, 0.01, dt, order=1)
    vg.synchronize()
    vg.profile_enable(False)
    prof = vg.profile_dump()
    pse_sim = {k: prof.get(k, (0, 0.0))[1] / reps
               for k in ("vg_pse_gather", "vg_pse_simultaneous", "vg_pse_scatter")}
    recs.append({
        "workload": "PSE viscous diffusion: 1M vortons (64 random rings), nu = 0.01, dt = 1/30, "
                    "reference threaded-build exchange order reproduced bit for bit",
        "grid": list(g.shape), "vortons": n,
        "ms": sum(pse.values()), "phases_ms": pse["vg_pse_phase"],
        "schedule_ms": pse["vg_pse_schedule"],
        "fractions_ms": pse["vg_pse_frac_size"] + pse["vg_scan_tile"] + pse["vg_scan_add"]
        + pse["vg_pse_fractions"],
        "gather_scatter_ms": pse["vg_pse_gather"] + pse["vg_pse_scatter"],
        "simultaneous_ms": sum(pse_sim.values()),
        "simultaneous_kernel_ms": pse_sim["vg_pse_simultaneous"],
    })

    for r in recs:
        print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()

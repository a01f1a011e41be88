#!/usr/bin/env python3
"""Per-step particle pipeline on one MI355X: whole FLIP steps per second.

Default workload (the largest single-GPU configuration): two colliding fluid
spheres (r = 0.15 at (0.3,.5,.5) moving +2 m/s and at (0.7,.5,.5) moving
-2 m/s) on a 512^3 MAC grid stored as 8^3 bricks, 8 jittered particles per
cell (Philox seed 4, 30.36M particles), dt = 1/240.  One step is the whole
particle step:
  bin -> 24+ bit key/value radix sort -> gather of the particle records ->
  per-cell counts -> exclusive scan -> trilinear MAC P2G -> wall faces and
  gravity -> active-brick Poisson system -> mixed-precision MIC(0)-PCG to an
  FP64 relative residual of 1e-6 (FP32 matvec and preconditioner, FP64 dot
  products, true residual and refinement) -> pressure-gradient update ->
  FLIP 0.95 G2P -> RK2 advection.
The particle state is resident in device memory before the timed region.

Prints one JSON line.  value = particles x steps / wall seconds of the timed
region (bracketed by device synchronisation).  roofline: the kernel with the
largest share of an untimed profiled step; in the timed region only that
kernel's launches are bracketed by HIP events on the simulation's stream,
and its algorithmic bytes per launch / mean launch time give `achieved`.
Extra keys: the same step on the 64^3 dam break (c1), the learned
preconditioner against plain PCG on the 128^3 dam break with a sphere (c4,
the committed network), and the 2^24-particle / 256^3 transfer microbench
(c2) with per-kernel rates.  cpu_baseline: the oracle's whole serial step
(FP64, Jacobi-PCG) on one host thread over the same scene at 128^3.
"""
import argparse
import glob
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s; ms/frame pressure solve @1e-6 rel. residual, learned vs PCG"
PEAK_HBM_GBPS = 8000.0
TIMING_SAMPLE = 16  # the timed steps record one launch in this many of the dominant kernel
NET_PATH = os.path.join(ROOT, "flip2_amd", "data", "c4_unet_16_32_64.npy")

# Algorithmic bytes of the solver kernels per fluid cell (inputs read once,
# outputs written once; coefficients are u16, Krylov vectors FP32, the
# solution and right-hand side FP64).
SOLVER_BYTES_PER_FLUID_CELL = {
    "cg_direction_matvec_kernel": 4 + 4 + 2 + 4 + 4,  # z, p_old, coef -> p, q
    "cg_update_jacobi_kernel": 4 + 4 + 4 + 4 + 2 + 4 + 4 + 4,  # p, q, e, r, coef -> e, r, z
    "cg_update_mic_kernel": 4 + 4 + 4 + 4 + 2 + 4 + 4 + 4 + 4,  # ... and the MIC(0) factors
    "matvec_dot_kernel": 4 + 2 + 4,              # p, coef -> q
    "update_kernel": 4 + 4 + 2 + 4 + 4 + 4 + 4,  # p, q, coef, e, r -> e, r
    "refine_start_kernel": 8 + 2 + 4 + 4,        # r64, coef -> r, e
    "true_residual_kernel": 8 + 8 + 2 + 8,       # b, x, coef -> r64
    "accumulate_kernel": 4 + 2 + 8 + 8,          # e, coef, x -> x
}


def kernel_alg_bytes(name, n, cells, fluid_cells, jacobi=False):
    """Algorithmic bytes of one launch of `name` in a FLIP step (n particles,
    `cells` grid cells, `fluid_cells` fluid cells); None if not modelled."""
    if name in SOLVER_BYTES_PER_FLUID_CELL:
        return SOLVER_BYTES_PER_FLUID_CELL[name] * fluid_cells
    per = {
        "bin_rec_kernel": 16 * n,                    # position in, key out
        "gather_rec_kernel": 52 * n,                 # perm + 24 B state in, 24 B out
        "p2g_brick_kernel": 24 * n + 16 * cells,     # state in, cell starts in, 3 faces out
        "prepare_faces_kernel": 36 * cells,          # 3 faces in, 6 faces out
        "g2p_flip_kernel": 24 * n + 12 * n,          # position + velocity in, velocity out
        "advect_rk2_kernel": 12 * n + 12 * n,        # position in and out
    }
    return per.get(name)


def transfer_alg_bytes(name, n, cells):
    """Bytes a transfer-microbench kernel must move (reads + writes), per launch."""
    if name == "bin":
        return n * (12 + 4)
    if name == "sort_setup":
        return n * 4
    if name.startswith("sort_pass"):
        p = int(name[len("sort_pass"):])
        return n * (12 if p == 0 else 16)  # first pass generates the payload
    if name == "gather":
        return n * (4 + 24 + 24)
    if name == "counts":
        return 4 * n + 4 * (cells + 1)  # sorted keys read, every count written once
    if name == "scan":
        return 8 * (cells + 1)
    if name == "p2g":
        return n * 24 + 4 * cells + 12 * cells
    raise KeyError(name)


def profiled_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 counter summaries (lines 'TRAFFIC workload=... kernel=...
    bytes_per_launch=...' in profiles/*.md: FETCH_SIZE + WRITE_SIZE per
    dispatch, scaled by the same counters on a copy of known size); None if
    no summary covers it."""
    pat = re.compile(r"TRAFFIC workload=(\S+) kernel=(\S+) bytes_per_launch=([0-9.eE+]+)")
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as f:
            for line in f:
                m = pat.search(line)
                if m and m.group(1) == workload and m.group(2) == kernel:
                    return float(m.group(3))
    return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(grid_n=128, target_seconds=15.0):
    """The oracle's whole serial step (FP64; bin, 1-bit LSD radix sort,
    gather, counts, scan, P2G, walls, assembly, Jacobi-PCG to 1e-6, pressure
    gradient, FLIP G2P, RK2) on the colliding-spheres scene at grid_n^3, one
    host thread, steps after an untimed first step."""
    from flip2_amd import scenes
    from oracle import pyoracle

    (x, y, z), vel = scenes.colliding_spheres(grid_n, seed=4)
    st = pyoracle.NsStep(x, y, z, *vel, (0.0, 0.0, 0.0, grid_n, grid_n, grid_n, 1.0 / grid_n))
    st.run()  # untimed: the first step's right-hand side is zero
    reps, its, t0 = 0, [], time.perf_counter()
    while True:
        st.run()
        its.append(st.iterations)
        reps += 1
        el = time.perf_counter() - t0
        if el >= target_seconds or reps >= 50:
            break
    st.close()
    return {"value": x.size * reps / el, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"{reps} serial steps (after one untimed step) of the oracle's whole FP64 step (transfer with "
                      f"a 1-bit-split radix sort, walls, assembly, Jacobi-PCG to 1e-6 ({int(np.mean(its))} "
                      f"iterations on average), pressure gradient, FLIP G2P, RK2) on the colliding-spheres scene at "
                      f"{grid_n}^3 ({x.size} particles, the same geometry, seed and 8 particles per cell as the "
                      f"512^3 workload), {el:.1f} s on 1 host thread of {os.cpu_count()} ({cpu_model()})"}


def copy_bandwidth(nbytes=1 << 30, reps=20):
    """Measured device-to-device copy rate (bytes read + written per second)
    of the library's 16-byte copy kernel, as a practical HBM ceiling."""
    import flip2_amd as F
    src, dst = F.DeviceBuffer(nbytes), F.DeviceBuffer(nbytes)
    F.check(F.lib().flip2_memset(src.ptr, 1, nbytes))
    F.check(F.lib().flip2_copy(dst.ptr, src.ptr, nbytes, None))
    F.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        F.check(F.lib().flip2_copy(dst.ptr, src.ptr, nbytes, None))
    F.synchronize()
    return 2.0 * nbytes * reps / (time.perf_counter() - t0) / 1e9


def _summary(solves):
    it = [s["iterations"] for s in solves]
    ms = [s["solve_ms"] for s in solves]
    return {"mean_iterations": round(float(np.mean(it)), 1), "mean_solve_ms": round(float(np.mean(ms)), 3),
            "max_rel_residual": float(max(s["rel_residual"] for s in solves)),
            "all_converged": bool(all(s["converged"] for s in solves)),
            "fluid_cells": int(solves[-1]["fluid_cells"]), "active_bricks": int(solves[-1]["active_bricks"])}


def make_scene(cfg, grid_c3=512):
    """(sim arguments, particles, velocities, solid, description) of a FLIP configuration."""
    from flip2_amd import scenes
    if cfg == "c1":
        n, prec = 64, "jacobi"
        x, y, z = scenes.dam_break(n, seed=1)
        return n, prec, (x, y, z), None, None, (
            "dam break x<0.4, y<0.6 on a 64^3 MAC grid (8^3 bricks), 8 jittered particles per cell (seed 1), "
            "dt 1/240, Jacobi-PCG to an FP64 relative residual of 1e-6")
    if cfg == "c3":
        n, prec = grid_c3, "mic0"
        (x, y, z), vel = scenes.colliding_spheres(n, seed=4)
        return n, prec, (x, y, z), vel, None, (
            f"colliding spheres: r=0.15 at (0.3,.5,.5) +2 m/s and (0.7,.5,.5) -2 m/s, {n}^3 MAC grid as 8^3 "
            "bricks, 8 jittered particles per cell (seed 4), dt 1/240; whole step (bin, key/value radix sort, "
            "gather, counts, exclusive scan, trilinear MAC P2G, walls + gravity, active-brick system, "
            "mixed-precision MIC(0)-PCG to an FP64 relative residual of 1e-6, pressure gradient, FLIP 0.95 G2P, "
            "RK2)")
    raise SystemExit(f"unknown config {cfg}")


def fluid_steps(cfg, steps, warmup, grid_c3=512, roofline=False):
    """Whole FLIP steps of configuration cfg; returns the result dict."""
    import flip2_amd as F

    n, prec, (x, y, z), vel, solid, desc = make_scene(cfg, grid_c3)
    grid = (0.0, 0.0, 0.0, n, n, n, 1.0 / n)
    print(f"[bench] {cfg}: {x.size} particles, {n}^3", file=sys.stderr, flush=True)
    sim = F.FluidSim(x.size, grid, preconditioner=prec)
    zero = np.zeros_like(x)
    sim.set_particles(x, y, z, *(vel if vel is not None else (zero, zero, zero)))
    if solid is not None:
        sim.set_solid(solid)
    for _ in range(warmup):  # untimed: allocation, first sort
        sim.step()
    F.synchronize()
    out = {"workload": desc, "particles": int(x.size), "grid": n, "preconditioner": prec}
    dom = None
    if roofline:
        # one untimed profiled step: every kernel bracketed by events
        sim.enable_timing()
        sim.step()
        prof = sim.timing()
        sim.disable_timing()
        prof_solve = sim.last_solve()
        total = sum(ms for ms, _ in prof.values())
        dom = max(prof, key=lambda k: prof[k][0])
        out["profiled_step"] = {
            "ms_all_kernels": round(total, 3), "solve_iterations": prof_solve["iterations"],
            "kernels": {k: {"ms": round(ms, 4), "launches": c} for k, (ms, c) in
                        sorted(prof.items(), key=lambda kv: -kv[1][0])}}
        # the timed region brackets one in TIMING_SAMPLE launches of the
        # dominant kernel (the events cost the stream ~5 us each)
        sim.enable_timing([dom], sample_every=TIMING_SAMPLE)
    print(f"[bench] {cfg}: timing {steps} steps", file=sys.stderr, flush=True)
    solves = []
    F.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
        solves.append(sim.last_solve())
    F.synchronize()
    wall = time.perf_counter() - t0
    out.update({"steps": steps, "value": x.size * steps / wall, "unit": "particle-steps/s",
                "ms_per_step": round(wall / steps * 1e3, 3), "solve": _summary(solves)})
    if roofline:
        t = sim.timing()
        sim.disable_timing()
        ms, launches = t[dom]
        avg_ms = ms / launches
        its = np.array([s["iterations"] for s in solves], np.float64)
        fcs = np.array([s["fluid_cells"] for s in solves], np.float64)
        fc = float((its * fcs).sum() / max(its.sum(), 1.0))  # fluid cells per solver launch
        b = kernel_alg_bytes(dom, int(x.size), n ** 3, fc, jacobi=(prec == "jacobi"))
        achieved = b / (avg_ms * 1e-3) / 1e9 if b else None
        traffic = profiled_traffic(f"{cfg}_{n}", dom)
        # launches of the dominant kernel in the timed steps, from the
        # profiled step's launches per solver iteration
        per_it = prof[dom][1] / max(prof_solve["iterations"], 1)
        est_launches = per_it * float(its.sum())
        out["roofline"] = {
            "bound": "hbm", "kernel": dom, "achieved": round(achieved, 1) if achieved else None,
            "peak": PEAK_HBM_GBPS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBPS, 4) if achieved else None,
            "traffic": traffic, "alg_bytes_per_launch": int(b) if b else None,
            "timed_launches": launches, "timing_sample_every": TIMING_SAMPLE, "avg_launch_ms": round(avg_ms, 5),
            "share_of_timed_region": round(avg_ms * est_launches * 1e-3 / wall, 4),
            "fluid_cells_per_launch": int(fc)}
    sim.close()
    return out


def learned_vs_pcg(frames=200, compare=8, grid_n=128, fallback=0, flex=8, net_path=NET_PATH, no_graph=False,
                   smoother="ic0", smooth_steps=1):
    """Learned FCG vs Jacobi-PCG and MIC(0)-PCG on the 128^3 dam break with a
    sphere: ms per pressure solve to the 1e-6 FP64 residual on the same
    systems, the `compare` frames after `frames` Jacobi-PCG steps."""
    import flip2_amd as F
    from flip2_amd import scenes

    n = grid_n
    (x, y, z), solid = scenes.dam_break_sphere(n, seed=5)
    grid = (0.0, 0.0, 0.0, n, n, n, 1.0 / n)
    sim = F.FluidSim(x.size, grid, preconditioner="jacobi")
    sim.set_particles(x, y, z, *(np.zeros_like(x),) * 3)
    sim.set_solid(solid)
    net = F.UNet(n, seed=2)
    net.set_params(np.load(net_path, allow_pickle=False))
    print(f"[bench] c4: {x.size} particles; {frames} frames of Jacobi-PCG steps", file=sys.stderr, flush=True)
    for _ in range(frames):
        sim.step()
    sim.set_unet(net, fallback_iterations=fallback)
    sim.set_flex_directions(flex)
    sim.set_learned_smoother(smoother)
    sim.set_learned_smooth_steps(smooth_steps)
    if no_graph:
        sim.set_learned_graph(False)
    res = compare_solvers(sim, compare)
    out = {"workload": f"dam break + static sphere r=0.1 at (0.6,0.2,0.5), {n}^3, 8 ppc (seed 5); U-Net 16/32/64, "
                       f"BF16 weights, FP32 accumulation ({os.path.relpath(net_path, ROOT)}); the {compare} frames "
                       f"after frame {frames}",
           "particles": int(x.size), "grid": n, "frames": compare}
    out.update(res)
    sim.close()
    return out


def compare_solvers(sim, compare, paddle=False):
    """Solve each frame's system with Jacobi-PCG, MIC(0)-PCG and learned FCG
    (untimed first use), then advance the frame with the Jacobi solution."""
    from flip2_amd import scenes
    res = {"jacobi": [], "mic0": [], "learned": []}
    t_sim, dt = 0.0, 1.0 / 240.0
    for s in range(compare):
        if paddle:
            xp, up = scenes.paddle_position(t_sim)
            sim.set_solid_box((0.0, 0.0, 0.0), (xp, 1.0, 1.0), (up, 0.0, 0.0))
        for st in ("transfer", "prepare", "system"):
            sim.stage(st)
        if s == 0:
            for name in ("jacobi", "mic0", "learned"):
                sim.solve_with(name)
        for name in ("jacobi", "mic0", "learned"):
            res[name].append(sim.solve_with(name))
        sim.solve_with("jacobi")
        for st in ("project", "g2p", "advect"):
            sim.stage(st)
        t_sim += dt
    summ = {k: _summary(v) for k, v in res.items()}
    summ["learned"]["used_fallback"] = int(sum(s["used_fallback"] for s in res["learned"]))
    return {"ms_per_frame": {k: summ[k]["mean_solve_ms"] for k in summ},
            "iterations": {k: summ[k]["mean_iterations"] for k in summ}, "solves": summ,
            "speedup_vs_jacobi": round(summ["jacobi"]["mean_solve_ms"] / summ["learned"]["mean_solve_ms"], 3),
            "speedup_vs_mic0": round(summ["mic0"]["mean_solve_ms"] / summ["learned"]["mean_solve_ms"], 3)}


def transfer_microbench(steps=20, warmup=3, n=1 << 24, gn=256):
    """bin -> radix sort -> gather -> counts -> scan -> P2G on 2^24 particles
    uniform-random in a 256^3 grid; every step starts from the same unsorted
    (fully shuffled) input, the gather's worst case."""
    import flip2_amd as F
    from flip2_amd import scenes

    grid = (0.0, 0.0, 0.0, gn, gn, gn, 1.0 / gn)
    print(f"[bench] c2: generating {n} particles", file=sys.stderr, flush=True)
    x, y, z = scenes.uniform_random(n, seed=2)
    vx, vy, vz = scenes.uniform_velocities(n, seed=3)
    rec = np.zeros((n, 8), np.float32)
    for col, a in enumerate((x, y, z, vx, vy, vz)):
        rec[:, col] = a
    records = F.DeviceBuffer.from_array(rec)
    del rec
    cells = gn ** 3
    out = [F.DeviceBuffer(4 * cells) for _ in range(3)]
    pl = F.P2GPipeline(n, grid)
    for _ in range(warmup):
        pl.run_records(n, records, out)
    F.synchronize()
    pl.enable_timing(steps)
    F.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        pl.run_records(n, records, out)
    F.synchronize()
    wall = time.perf_counter() - t0
    kernels, runs = pl.timing()
    table = {}
    for name, total in kernels.items():
        ms = total / runs
        b = transfer_alg_bytes(name, n, cells)
        table[name] = {"ms": round(ms, 5), "alg_bytes": int(b), "GBps": round(b / (ms * 1e-3) / 1e9, 1),
                       "frac": round(b / (ms * 1e-3) / 1e9 / PEAK_HBM_GBPS, 4)}
    pl.close()
    return {"workload": "2^24 particles uniform-random in a 256^3 grid (8^3 bricks), FP32 velocities U[-1,1); "
                        "bin -> 24-bit key/value radix sort -> gather -> counts + exclusive scan over 16,777,216 "
                        "counts -> trilinear MAC P2G; every step re-sorts the same fully shuffled input",
            "particles": n, "grid": gn, "steps": steps, "ms_per_step": round(wall / steps * 1e3, 4),
            "value": n * steps / wall, "unit": "particle-transfers/s", "kernels": table}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid-c3", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="headline line only (no c1 / c4 / c2 keys)")
    ap.add_argument("--config", choices=["c1", "c1rc", "c2", "c3", "c4", "c5"], default=None,
                    help="one configuration on its own line instead of the default headline")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB --steps times (reference traffic for counters)")
    ap.add_argument("--rc-frames", type=int, default=3, help="c1rc: timed frames")
    ap.add_argument("--grid-c4", type=int, default=128)
    ap.add_argument("--grid-c5", type=int, default=256)
    ap.add_argument("--frames", type=int, default=200, help="c4: Jacobi-PCG steps before the compared frames")
    ap.add_argument("--compare-frames", type=int, default=8)
    ap.add_argument("--fallback", type=int, default=0, help="learned: Jacobi-PCG fallback iteration cap (0: none)")
    ap.add_argument("--flex", type=int, default=8, help="learned: earlier directions flexible CG keeps (1..16)")
    ap.add_argument("--smoother", choices=["jacobi", "ic0"], default="ic0",
                    help="learned: smoothing step around the network")
    ap.add_argument("--smooth-steps", type=int, default=1, help="learned: IC(0) smoothing steps (1 or 2)")
    ap.add_argument("--no-graph", action="store_true",
                    help="learned: launch the preconditioner's kernels one by one instead of replaying a HIP graph")
    ap.add_argument("--load-net", default=NET_PATH, help="c4/c5: the network to use")
    # training (c4 --train): capture residuals of Jacobi-PCG solves and train a network
    ap.add_argument("--train", action="store_true", help="c4: train a network instead of loading one")
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--capture-its", default="0,2,4,8,16,32,64,128")
    ap.add_argument("--save-net", default=None)
    ap.add_argument("--init-net", default=None)
    ap.add_argument("--lr", type=float, default=1e-3)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this pipeline has no multi-GPU "
              "path: see DESIGN.md for how it would shard)", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        import flip2_amd as F
        nbytes = args.calibrate_copy << 20
        src, dst = F.DeviceBuffer(nbytes), F.DeviceBuffer(nbytes)
        F.check(F.lib().flip2_memset(src.ptr, 1, nbytes))
        for _ in range(args.steps):
            F.check(F.lib().flip2_copy(dst.ptr, src.ptr, nbytes, None))
        F.synchronize()
        print(json.dumps({"copy_bytes": nbytes, "launches": args.steps}), flush=True)
        return
    if args.config in ("c1", "c3"):
        r = fluid_steps(args.config, args.steps, args.warmup, args.grid_c3, roofline=True)
        r["config"] = args.config
        print(json.dumps(r), flush=True)
        return
    if args.config == "c1rc":
        return rc_run(args)
    if args.config == "c2":
        print(json.dumps(transfer_microbench(args.steps, args.warmup)), flush=True)
        return
    if args.config == "c4":
        if args.train:
            return train_c4(args)
        r = learned_vs_pcg(args.frames, args.compare_frames, args.grid_c4, args.fallback, args.flex, args.load_net,
                           args.no_graph, args.smoother, args.smooth_steps)
        print(json.dumps(r), flush=True)
        return
    if args.config == "c5":
        return c5_run(args)

    head = fluid_steps("c3", args.steps, args.warmup, args.grid_c3, roofline=True)
    result = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": head["workload"], "particles": head["particles"], "grid": head["grid"], "brick": 8,
                   "seed": 4, "precision": "FP32 particle state, P2G, matvec and MIC(0) apply; FP64 dot products, "
                                           "right-hand side, solution, true residual and refinement"},
        "roofline": dict(head["roofline"], copy_GBps=round(copy_bandwidth(), 1)),
        "solve": head["solve"],
        "profiled_step": head["profiled_step"],
        "cpu_baseline": None,
    }
    if not args.no_extras:
        c1 = fluid_steps("c1", 10, 2, roofline=True)
        result["c1"] = {k: c1[k] for k in ("workload", "particles", "value", "unit", "ms_per_step", "solve",
                                           "roofline")}
        result["learned_vs_pcg"] = learned_vs_pcg(smoother=args.smoother)
        result["learned_vs_pcg"]["metric"] = "ms/frame pressure solve @1e-6 rel. residual, learned vs PCG"
        result["transfer_c2"] = transfer_microbench()
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline()
    print(json.dumps(result), flush=True)


def rc_run(args):
    """The reference-compatible path (Particles::solveFrame: binning, 1-bit-split
    order sort, voxelisation, FP64 P2G, brick-local red-black SOR, G2P, Euler)
    on the C1 dam break: --rc-frames frames at 240 frames/s (one JSON line)."""
    import flip2_amd as F
    from flip2_amd import scenes

    n, frames, fps = 64, args.rc_frames, 240.0
    x, y, z = scenes.dam_break(n, seed=1)
    print(f"[bench] c1rc: {x.size} particles", file=sys.stderr, flush=True)
    p = F.Particles(x.size)
    p.setDomain(0.0, 0.0, 0.0, n, n, n, 1.0 / n)
    p.set_positions(x.astype(np.float64), y.astype(np.float64), z.astype(np.float64))
    zero = np.zeros(x.size)
    p.set_velocities(zero, zero, zero)
    p.initialize()
    t0 = time.perf_counter()
    p.solveFrame(fps)  # warm-up frame
    F.synchronize()
    print(f"[bench] c1rc: warm-up frame {time.perf_counter() - t0:.2f} s", file=sys.stderr, flush=True)
    sor, subs = [], 0
    t0 = time.perf_counter()
    for f in range(frames):
        p.solveFrame(fps)
        st = p.stats()
        sor.append(st["lastSorIterations"])
        subs += st["substeps"]
        print(f"[bench] c1rc: frame {f + 1}: {st['substeps']} substeps, last SOR {st['lastSorIterations']} "
              f"iterations, {time.perf_counter() - t0:.2f} s", file=sys.stderr, flush=True)
    F.synchronize()
    wall = time.perf_counter() - t0
    st = p.stats()
    out = {"metric": "particle-frames/s", "config": "c1rc",
           "workload": "reference-compatible path (FP64, SOR pressure) on the C1 dam break: x<0.4, y<0.6, 64^3, "
                       f"8 jittered particles per cell, {frames} frames at 240 frames/s",
           "particles": int(x.size), "grid": n, "frames": frames, "value": x.size * frames / wall,
           "ms_per_frame": round(wall / frames * 1e3, 3), "substeps": subs, "sor_iterations_last": sor,
           "used_nodes": st["numUsedGridNodes"], "used_voxels": st["numUsedVoxels"]}
    print(json.dumps(out), flush=True)


def c5_run(args):
    """The C4 network, without retraining, on the 256^3 wave tank with a
    moving paddle: learned FCG vs plain PCG per frame (one JSON line)."""
    import flip2_amd as F
    from flip2_amd import scenes

    n = args.grid_c5
    x, y, z = scenes.wave_tank(n, seed=6)
    grid = (0.0, 0.0, 0.0, n, n, n, 1.0 / n)
    sim = F.FluidSim(x.size, grid, preconditioner="jacobi")
    sim.set_particles(x, y, z, *(np.zeros_like(x),) * 3)
    net = F.UNet(n, seed=2)
    net.set_params(np.load(args.load_net, allow_pickle=False))
    sim.set_unet(net, fallback_iterations=args.fallback)
    sim.set_flex_directions(args.flex)
    sim.set_learned_smoother(args.smoother)
    sim.set_learned_smooth_steps(args.smooth_steps)
    if args.no_graph:
        sim.set_learned_graph(False)
    print(f"[bench] c5: {x.size} particles", file=sys.stderr, flush=True)
    out = {"metric": "ms/frame pressure solve @1e-6 rel. residual, learned vs PCG", "config": "c5",
           "workload": f"wave tank depth 0.35 with paddle x = 0.05 + 0.05 sin(2 pi t), {n}^3, 8 ppc (seed 6); "
                       f"network {os.path.relpath(args.load_net, ROOT)} trained on c4, no retraining",
           "particles": int(x.size), "grid": n, "frames": args.compare_frames}
    out.update(compare_solvers(sim, args.compare_frames, paddle=True))
    print(json.dumps(out), flush=True)


def train_c4(args):
    """Capture residuals of --frames Jacobi-PCG solves of the C4 scene and
    train the network on them (--save-net writes it)."""
    import flip2_amd as F
    from flip2_amd import scenes

    its = [int(v) for v in args.capture_its.split(",")]
    n = args.grid_c4
    (x, y, z), solid = scenes.dam_break_sphere(n, seed=5)
    grid = (0.0, 0.0, 0.0, n, n, n, 1.0 / n)
    sim = F.FluidSim(x.size, grid, preconditioner="jacobi")
    sim.set_particles(x, y, z, *(np.zeros_like(x),) * 3)
    sim.set_solid(solid)
    net = F.UNet(n, seed=2)
    if args.init_net:
        net.set_params(np.load(args.init_net, allow_pickle=False))
    print(f"[bench] c4: {x.size} particles; capturing {args.frames} frames", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    net.capture_residuals(sim, its)
    for _ in range(args.frames):
        sim.step()
    net.capture_residuals(sim, [])
    nf = net.num_frames()
    capture_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    losses = []
    for e in range(args.epochs):
        lr = args.lr * 0.3 ** (e / max(1, args.epochs - 1))
        losses.append(net.train(1, lr, 0, nf, seed=e))
        print(f"[bench] epoch {e + 1}: loss {losses[-1]:.4f}", file=sys.stderr, flush=True)
    train_s = time.perf_counter() - t0
    if args.save_net:
        np.save(args.save_net, net.params(), allow_pickle=False)
    sim.set_unet(net, fallback_iterations=args.fallback)
    sim.set_flex_directions(args.flex)
    sim.set_learned_smoother(args.smoother)
    out = {"config": "c4-train", "frames_captured": nf, "epochs": args.epochs, "capture_s": round(capture_s, 1),
           "train_s": round(train_s, 1), "final_loss": round(float(losses[-1]), 5) if losses else None}
    out.update(compare_solvers(sim, args.compare_frames))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Primitive microbench of the per-step particle pipeline on one MI355X.

Workload (the configuration the headline metric is quoted on): 16,777,216
particles uniform-random in the unit cube (Philox seed 2), FP32 velocities
U[-1,1)^3 (seed 3), a 256^3 grid stored as 8^3 bricks.  One step:
  bin (brick-tiled u32 cell keys) -> key/value radix sort (24 key bits, u32
  payload = particle index) -> gather of the six particle arrays -> per-cell
  counts -> exclusive scan over the 16,777,216 cell counts -> trilinear MAC P2G
  onto the bricks.
Inputs (the particle state as 32-byte records, the layout the FLIP step
keeps) are resident in device memory before the timed region; every step
redoes the whole transfer from the same unsorted input.

Prints one JSON line.  value = particles x steps / wall seconds of the timed
region (bracketed by device synchronisation).  roofline: the kernel with the
largest share of the timed region, its algorithmic bytes per launch / its
mean duration from HIP events recorded around every kernel of the timed
region.  cpu_baseline: the oracle's serial C transfer step on this host's
CPU (1 thread) over a sample of the same workload (2^21 particles in a 128^3
grid: 1/8 of the particles on 1/8 of the cells, same density).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s; ms/frame pressure solve @1e-6 rel. residual, learned vs PCG"
PEAK_HBM_GBPS = 8000.0

# HBM bytes per dispatch of the default workload (2^24 particles, 256^3),
# from rocprofv3 FETCH_SIZE / WRITE_SIZE counter runs scaled by the same
# counters on a plain copy of known size (profiles/transfer_16M_256_pmc.md).
TRAFFIC_BYTES = {"bin": 604.0e6, "sort_pass": 284.6e6, "gather": 2748.0e6, "scan": 135.9e6, "p2g": 962.4e6}


def algorithmic_bytes(name, n, cells, occupied, passes):
    """Bytes a kernel must move at minimum (reads + writes), per launch."""
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


def cpu_baseline(target_seconds=10.0):
    from flip2_amd import scenes
    from oracle import pyoracle

    n, grid_n = 1 << 21, 128
    x, y, z = scenes.uniform_random(n, seed=2)
    vx, vy, vz = scenes.uniform_velocities(n, seed=3)
    step = pyoracle.NsTransferStep(x, y, z, vx, vy, vz, (0.0, 0.0, 0.0, grid_n, grid_n, grid_n, 1.0 / grid_n))
    reps, t0 = 0, time.perf_counter()
    while True:
        step.run()
        reps += 1
        el = time.perf_counter() - t0
        if el >= target_seconds or reps >= 50:
            break
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": n * reps / el, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"{reps} serial steps of the oracle's C transfer step (bin, 21-bit 1-bit-split radix sort, "
                      f"gather, counts, exclusive scan, FP64 trilinear MAC P2G) on 2^21 uniform-random particles "
                      f"in a 128^3 grid, {el:.1f} s on 1 host thread of {os.cpu_count()} ({model})"}


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


def _fluid_step_stats(sim, steps):
    """Run `steps` FLIP steps; return wall seconds and per-step solve stats."""
    import flip2_amd as F
    F.synchronize()
    solves = []
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.step()
        solves.append(sim.last_solve())
    F.synchronize()
    return time.perf_counter() - t0, solves


def _summary(solves):
    it = [s["iterations"] for s in solves]
    ms = [s["solve_ms"] for s in solves]
    return {"mean_iterations": round(float(np.mean(it)), 1), "mean_solve_ms": round(float(np.mean(ms)), 3),
            "max_rel_residual": float(max(s["rel_residual"] for s in solves)),
            "all_converged": bool(all(s["converged"] for s in solves)),
            "fluid_cells": int(solves[-1]["fluid_cells"])}


def config_run(args):
    """Whole FLIP steps on the configurations the solver targets (one JSON line)."""
    import flip2_amd as F
    from flip2_amd import scenes

    cfg = args.config
    if cfg == "c1":
        n, steps, prec = 64, 10, "jacobi"
        x, y, z = scenes.dam_break(n, seed=1)
        vel = None
        solid = None
        desc = "dam break x<0.4, y<0.6 on a 64^3 MAC grid, 8 jittered particles per cell, 10 steps, Jacobi-PCG 1e-6"
    elif cfg == "c3":
        n, steps, prec = args.grid_c3, 3, "mic0"
        (x, y, z), vel = scenes.colliding_spheres(n, seed=4)
        solid = None
        desc = (f"two fluid spheres r=0.15 at (0.3,.5,.5) +2 m/s and (0.7,.5,.5) -2 m/s, {n}^3 as 8^3 bricks, "
                "8 ppc, MIC(0)-PCG 1e-6 (FP32 matvec/preconditioner, FP64 dots and refinement)")
    else:
        raise SystemExit(f"unknown config {cfg}")
    grid = (0.0, 0.0, 0.0, n, n, n, 1.0 / n)
    print(f"[bench] {cfg}: {x.size} particles", file=sys.stderr, flush=True)
    sim = F.FluidSim(x.size, grid, preconditioner=prec)
    zero = np.zeros_like(x)
    if vel is None:
        sim.set_particles(x, y, z, zero, zero, zero)
    else:
        sim.set_particles(x, y, z, *vel)
    if solid is not None:
        sim.set_solid(solid)
    _fluid_step_stats(sim, 1)  # warm-up step (allocation, first sort)
    wall, solves = _fluid_step_stats(sim, steps)
    out = {"metric": "particle-steps/s", "config": cfg, "workload": desc, "particles": int(x.size), "grid": n,
           "steps": steps, "value": x.size * steps / wall, "ms_per_step": round(wall / steps * 1e3, 3),
           "preconditioner": prec, "solve": _summary(solves)}
    print(json.dumps(out), flush=True)


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


def learned_run(args):
    """Learned preconditioner vs plain PCG, ms per frame of the pressure solve
    to the 1e-6 FP64 relative residual (c4: train then compare; c5: apply a
    saved network to the 256^3 wave tank with a moving paddle)."""
    import flip2_amd as F
    from flip2_amd import scenes

    its = [int(v) for v in args.capture_its.split(",")]
    if args.config == "c4":
        n = args.grid_c4
        (x, y, z), solid = scenes.dam_break_sphere(n, seed=5)
        grid = (0.0, 0.0, 0.0, n, n, n, 1.0 / n)
        sim = F.FluidSim(x.size, grid, preconditioner="jacobi")
        sim.set_particles(x, y, z, *(np.zeros_like(x),) * 3)
        sim.set_solid(solid)
        net = F.UNet(n, seed=2)
        print(f"[bench] c4: {x.size} particles; capturing {args.frames} frames", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        if not args.load_net:
            net.capture_residuals(sim, its)
        for _ in range(args.frames):
            sim.step()
        net.capture_residuals(sim, [])
        nf = net.num_frames()
        capture_s = time.perf_counter() - t0
        if args.load_net:
            # a network trained by an earlier run of this configuration
            net.set_params(np.load(args.load_net, allow_pickle=False))
        elif args.init_net:
            # continue the training of an earlier run's network
            net.set_params(np.load(args.init_net, allow_pickle=False))
        epochs = 0 if args.load_net else args.epochs
        print(f"[bench] training {epochs} epochs on {nf} frames", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        losses = []
        for e in range(epochs):
            # learning rate from --lr down to 0.3 --lr, geometrically over the epochs
            lr = args.lr * 0.3 ** (e / max(1, args.epochs - 1))
            losses.append(net.train(1, lr, 0, nf, seed=e))
            print(f"[bench] epoch {e + 1}: loss {losses[-1]:.4f}", file=sys.stderr, flush=True)
        train_s = time.perf_counter() - t0
        if args.save_net:
            np.save(args.save_net, net.params(), allow_pickle=False)
        trained = (f"network from {args.load_net}" if args.load_net else
                   f"trained on {nf} residual frames of {args.frames} Jacobi-PCG solves ({args.epochs} epochs"
                   + (f" after {args.init_net}" if args.init_net else "") + ")")
        desc = (f"dam break + static sphere r=0.1 at (0.6,0.2,0.5), {n}^3, 8 ppc; U-Net 16/32/64, BF16/FP32, "
                f"{trained}; compared on the {args.compare_frames} frames after frame {args.frames}")
        extra = {"capture_s": round(capture_s, 1), "train_s": round(train_s, 1),
                 "final_loss": round(float(losses[-1]), 5) if losses else None}
    elif args.config == "c5":
        n = args.grid_c5
        x, y, z = scenes.wave_tank(n, seed=6)
        grid = (0.0, 0.0, 0.0, n, n, n, 1.0 / n)
        sim = F.FluidSim(x.size, grid, preconditioner="jacobi")
        sim.set_particles(x, y, z, *(np.zeros_like(x),) * 3)
        net = F.UNet(n, seed=2)
        if not args.load_net:
            raise SystemExit("c5 applies a network trained on c4: pass --load-net")
        net.set_params(np.load(args.load_net, allow_pickle=False))
        desc = (f"wave tank depth 0.35 with paddle x = 0.05 + 0.05 sin(2 pi t), {n}^3, 8 ppc; network trained "
                "on c4, no retraining")
        extra = {}
    else:
        raise SystemExit(f"unknown config {args.config}")
    sim.set_unet(net, fallback_iterations=args.fallback)
    if args.no_graph:
        sim.set_learned_graph(False)  # one launch per kernel (for kernel tracing)
    sim.set_flex_directions(args.flex)
    res = {"jacobi": [], "mic0": [], "learned": []}
    t_sim = 0.0
    dt = 1.0 / 240.0
    for s in range(args.compare_frames):
        if args.config == "c5":
            xp, up = scenes.paddle_position(t_sim)
            sim.set_solid_box((0.0, 0.0, 0.0), (xp, 1.0, 1.0), (up, 0.0, 0.0))
        for st in ("transfer", "prepare", "system"):
            sim.stage(st)
        if s == 0:
            # untimed: first-use allocations (direction ring, network buffers)
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
    out = {"metric": "ms/frame pressure solve @1e-6 rel. residual, learned vs PCG", "config": args.config,
           "workload": desc, "particles": int(x.size), "grid": n, "frames": args.compare_frames,
           "solves": summ,
           "speedup_vs_jacobi": round(summ["jacobi"]["mean_solve_ms"] / summ["learned"]["mean_solve_ms"], 3),
           "speedup_vs_mic0": round(summ["mic0"]["mean_solve_ms"] / summ["learned"]["mean_solve_ms"], 3)}
    out.update(extra)
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--particles", type=int, default=1 << 24)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB --steps times (reference traffic for counters)")
    ap.add_argument("--config", choices=["c1", "c1rc", "c3", "c4", "c5"], default=None,
                    help="whole FLIP steps (c1, c3), the reference-compatible path on C1 (c1rc) or the "
                         "learned-vs-PCG solve comparison (c4, c5) instead of the transfer microbench")
    ap.add_argument("--grid-c3", type=int, default=512)
    ap.add_argument("--rc-frames", type=int, default=3, help="c1rc: timed frames")
    ap.add_argument("--grid-c4", type=int, default=128)
    ap.add_argument("--grid-c5", type=int, default=256)
    ap.add_argument("--frames", type=int, default=200, help="c4: frames of Jacobi-PCG solves captured for training")
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--capture-its", default="0,2,4,8,16,32,64,128",
                    help="c4: inner iterations of each Jacobi-PCG solve whose residuals become training frames")
    ap.add_argument("--compare-frames", type=int, default=8)
    ap.add_argument("--fallback", type=int, default=0, help="learned: Jacobi-PCG fallback iteration cap (0: none)")
    ap.add_argument("--flex", type=int, default=8, help="learned: earlier directions flexible CG keeps (1..16)")
    ap.add_argument("--no-graph", action="store_true",
                    help="learned: launch the preconditioner's kernels one by one instead of replaying a HIP graph")
    ap.add_argument("--save-net", default=None)
    ap.add_argument("--load-net", default=None, help="c4/c5: use this saved network (c4: no training)")
    ap.add_argument("--init-net", default=None, help="c4: start the training from this saved network")
    ap.add_argument("--lr", type=float, default=1e-3, help="c4: first epoch's learning rate")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this pipeline has no multi-GPU "
              "path: see DESIGN.md for how it would shard)", file=sys.stderr)
        sys.exit(2)

    if args.config in ("c1", "c3"):
        return config_run(args)
    if args.config == "c1rc":
        return rc_run(args)
    if args.config in ("c4", "c5"):
        return learned_run(args)

    import flip2_amd as F
    from flip2_amd import scenes

    if args.calibrate_copy:
        nbytes = args.calibrate_copy << 20
        src, dst = F.DeviceBuffer(nbytes), F.DeviceBuffer(nbytes)
        F.check(F.lib().flip2_memset(src.ptr, 1, nbytes))
        for _ in range(args.steps):
            F.check(F.lib().flip2_copy(dst.ptr, src.ptr, nbytes, None))
        F.synchronize()
        print(json.dumps({"copy_bytes": nbytes, "launches": args.steps}), flush=True)
        return

    n, gn = args.particles, args.grid
    grid = (0.0, 0.0, 0.0, gn, gn, gn, 1.0 / gn)
    print(f"[bench] generating {n} particles", file=sys.stderr, flush=True)
    x, y, z = scenes.uniform_random(n, seed=2)
    vx, vy, vz = scenes.uniform_velocities(n, seed=3)
    # the particle state as the FLIP step keeps it: 32-byte records
    rec = np.zeros((n, 8), np.float32)
    for col, a in enumerate((x, y, z, vx, vy, vz)):
        rec[:, col] = a
    records = F.DeviceBuffer.from_array(rec)
    del rec
    cells = gn ** 3
    out = [F.DeviceBuffer(4 * cells) for _ in range(3)]
    pl = F.P2GPipeline(n, grid)
    print(f"[bench] warmup {args.warmup}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        pl.run_records(n, records, out)
    F.synchronize()
    occupied = int(np.count_nonzero(pl.get("counts")))

    pl.enable_timing(args.steps)
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    F.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pl.run_records(n, records, out)
    F.synchronize()
    wall = time.perf_counter() - t0
    kernels, runs = pl.timing()
    assert runs == args.steps, (runs, args.steps)

    passes = sum(1 for k in kernels if k.startswith("sort_pass"))
    table = {}
    for name, total in kernels.items():
        ms = total / runs
        b = algorithmic_bytes(name, n, cells, occupied, passes)
        table[name] = {"ms": round(ms, 5), "alg_bytes": int(b), "GBps": round(b / (ms * 1e-3) / 1e9, 1)}
    dom = max(table, key=lambda k: table[k]["ms"])
    ms_step = wall / args.steps * 1e3
    achieved = table[dom]["GBps"]
    result = {
        "metric": METRIC,
        "value": n * args.steps / wall,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "primitive microbench: 16,777,216 particles uniform-random in a 256^3 grid "
                               "(8^3 bricks): bin -> 24-bit key/value radix sort -> gather -> per-cell counts "
                               "+ exclusive scan over 16,777,216 counts -> trilinear MAC P2G",
                   "particles": n, "grid": gn, "brick": 8, "seed": 2},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBPS, 4),
                     "traffic": (int(TRAFFIC_BYTES[dom]) if (n == 1 << 24 and gn == 256 and dom in TRAFFIC_BYTES)
                                 else None),
                     "alg_bytes": table[dom]["alg_bytes"],
                     "copy_GBps": round(copy_bandwidth(), 1)},
        "kernels": table,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

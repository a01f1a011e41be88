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
        return 4 * (cells + 1) + 4 * n + 4 * occupied
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
    return {"value": n * reps / el, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"{reps} serial steps of the oracle's C transfer step (bin, 21-bit 1-bit-split radix sort, "
                      f"gather, counts, exclusive scan, FP64 trilinear MAC P2G) on 2^21 uniform-random particles "
                      f"in a 128^3 grid, {el:.1f} s on 1 host thread"}


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
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (this pipeline has no multi-GPU "
              "path: see DESIGN.md for how it would shard)", file=sys.stderr)
        sys.exit(2)

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
                     "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": None},
        "kernels": table,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Autotuner of the launch-shape knobs (mcpm_set_tuning): priority bands of
the collision rounds, the round grid, the sort chunk.

Runs the chosen workload resident on the GPU, tries each setting for a few
steps (one knob at a time, the others at their best so far), and prints one
JSON object with the per-setting step times and the best setting.  No knob
changes any result, so the state keeps evolving across trials; the interface
workload is statistically stationary over these few steps.

    python autotune.py --workload interface --particles 20000000 --steps 4
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import mcpm  # noqa: E402
import workloads as W  # noqa: E402

SPACE = {
    mcpm.TUNE_BANDS: [8, 10, 12, 16],
    mcpm.TUNE_ROUND_GRID: [1024, 2048, 4096],
    mcpm.TUNE_SORT_CHUNK: [0, 1024, 2048, 4096, 8192],
}
NAMES = {mcpm.TUNE_BANDS: "bands", mcpm.TUNE_ROUND_GRID: "round_grid",
         mcpm.TUNE_SORT_CHUNK: "sort_chunk"}


def build(args):
    if args.workload == "interface":
        p, arrs, meta = W.interface(N=args.particles, n_cells=args.cells)
        sim = mcpm.Simulation(p)
        sim.upload(*arrs)
    else:
        p, arrs, meta = W.droplets(N=args.particles, nx=args.cells, ny=args.cells)
        sim = mcpm.Simulation(p)
        sim.upload_xy(*arrs)
    sim.cell_update()
    sim.init_omega_max(meta["n_gas"])
    return sim


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="interface", choices=["interface", "droplets"])
    ap.add_argument("--particles", type=int, default=20_000_000)
    ap.add_argument("--cells", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=3)
    args = ap.parse_args()
    sim = build(args)
    step = 0
    for _ in range(args.warmup):
        sim.step(step)
        step += 1
    sim.sync()
    best = {k: sim.get_tuning(k) for k in SPACE}
    trials = []
    for knob, values in SPACE.items():
        times = {}
        for v in values:
            try:
                sim.set_tuning(knob, v)
            except mcpm.McpmError:
                continue
            sim.step(step)  # one step to settle
            step += 1
            sim.sync()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                sim.step(step)
                step += 1
            sim.sync()
            ms = 1000.0 * (time.perf_counter() - t0) / args.steps
            times[v] = ms
            trials.append({"knob": NAMES[knob], "value": v, "ms_per_step": ms})
            print(f"{NAMES[knob]}={v}: {ms:.3f} ms/step", file=sys.stderr, flush=True)
        best[knob] = min(times, key=times.get)
        sim.set_tuning(knob, best[knob])
    print(json.dumps({"workload": args.workload, "particles": args.particles,
                      "best": {NAMES[k]: v for k, v in best.items()}, "trials": trials}))


if __name__ == "__main__":
    main()

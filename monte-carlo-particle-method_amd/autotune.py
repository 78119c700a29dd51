#!/usr/bin/env python3
"""Autotuner of the launch shapes: priority bands of the collision rounds,
the round grid (workgroups per round), the sort chunk (particles per
workgroup of the full-histogram sort), the operations per thread of a round
and the grid rounds of a captured step (the rounds before the one-workgroup
tail).

None of these changes a result, only the time.  The state of the system keeps
evolving while the tuner runs and the step cost drifts with it (omega_max
only grows, so the candidate count does), so every candidate is timed in
blocks that alternate with blocks of the best setting so far, and compared by
the ratio of the two; a candidate replaces the best only when it is faster by
more than `margin`.  Knobs are tuned one at a time in the order of SPACE.

    python autotune.py --workload interface --steps 4
    python autotune.py --workload replicas --replicas 256
    python autotune.py --workload dense --model fp
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

GRAPH_ROUNDS = "graph_rounds"  # not a mcpm_set_tuning knob: the capture's round count

SPACE = {
    mcpm.TUNE_BANDS: [8, 10, 12, 14, 16],
    mcpm.TUNE_ROUND_GRID: [512, 1024, 1536, 2048],
    mcpm.TUNE_SORT_CHUNK: [0, 2048, 4096, 8192],
    mcpm.TUNE_ROUND_UNROLL: [2, 4],  # 0 (the start): chosen from the band size
    GRAPH_ROUNDS: [2, 4, 6, 8],  # grid rounds beyond the band count
}
NAMES = {mcpm.TUNE_BANDS: "bands", mcpm.TUNE_ROUND_GRID: "round_grid",
         mcpm.TUNE_SORT_CHUNK: "sort_chunk", mcpm.TUNE_ROUND_UNROLL: "round_unroll",
         GRAPH_ROUNDS: "graph_rounds_past_bands"}


def tune(space, best, apply, time_block, pairs=2, margin=0.01, log=None, names=None):
    """Pick each knob's value in turn.

    space: {knob: [values]}; best: {knob: starting value} (updated and
    returned); apply(settings) makes a setting current; time_block() runs one
    block of steps under the current setting and returns its time.  Each
    candidate is timed in `pairs` block pairs ordered best, candidate,
    candidate, best, ... (an even count cancels a linear drift); the ratio
    of the candidate's total to the best's decides.  Returns (best, trials)."""
    trials = []
    for knob, values in space.items():
        for v in values:
            if v == best[knob]:
                continue
            cand = dict(best)
            cand[knob] = v
            t_best = t_cand = 0.0
            try:
                # best, candidate, candidate, best, ...: a drift linear in
                # time cancels between the two
                for q in range(pairs):
                    for first in ((True, False) if q % 2 == 0 else (False, True)):
                        apply(best if first else cand)
                        if first:
                            t_best += time_block()
                        else:
                            t_cand += time_block()
            except mcpm.McpmError:
                apply(best)
                continue
            ratio = t_cand / t_best
            trials.append({"knob": knob, "value": v, "ratio_to_best": ratio,
                           "best_value": best[knob]})
            if log:
                nm = (names or {}).get(knob, knob)
                log(f"{nm}={v}: {ratio:.4f} of {nm}={best[knob]}")
            if ratio < 1.0 - margin:
                best = cand
        apply(best)
    return best, trials


def build(args):
    if args.workload == "interface":
        p, arrs, meta = W.interface(N=args.particles, n_cells=args.cells)
        sim = mcpm.Simulation(p)
        sim.upload(*arrs)
        sim.cell_update()
        sim.init_omega_max(meta["n_gas"])
    elif args.workload == "dense":
        p, arrs, meta = W.dense_slab(n_density=args.density)
        if args.model == "fp":
            p.model = mcpm.MODEL_FP_DENSE
        sim = mcpm.Simulation(p)
        sim.upload(*arrs)
        sim.cell_update()
        sim.init_omega_max(meta["n_gas"])
    elif args.workload == "replicas":
        common, reps, rarrs, metas = W.replicas(n_rep=args.replicas, seed0=0)
        counts = [a[0].size for a in rarrs]
        sim = mcpm.Simulation.batch(common, reps)
        sim.upload_batch(counts, *[__import__("numpy").concatenate([a[q] for a in rarrs])
                                   for q in range(4)])
        sim.cell_update()
        sim.init_omega_max_batch([m["n_gas"] for m in metas])
    else:
        p, arrs, meta = W.droplets(N=args.particles, nx=args.cells, ny=args.cells)
        sim = mcpm.Simulation(p)
        sim.upload_xy(*arrs)
        sim.cell_update()
        sim.init_omega_max(meta["n_gas"])
    return sim


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="interface",
                    choices=["interface", "replicas", "dense", "droplets"])
    ap.add_argument("--model", default="esmc", choices=["esmc", "fp"],
                    help="dense workload: ESMC or FP_dense")
    ap.add_argument("--density", type=float, default=0.3, help="dense workload: n sigma^3")
    ap.add_argument("--particles", type=int, default=20_000_000)
    ap.add_argument("--cells", type=int, default=4096)
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--steps", type=int, default=3, help="steps per timed block")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-graph", action="store_true",
                    help="time one mcpm_step call per step (the plane always does)")
    args = ap.parse_args()
    sim = build(args)
    step = 0
    for _ in range(args.warmup):
        sim.step(step)
        step += 1
    sim.sync()
    fp = args.workload == "dense" and args.model == "fp"
    graph = not args.no_graph and args.workload != "droplets" and not fp
    space = dict(SPACE)
    if not graph:
        space.pop(GRAPH_ROUNDS)
    if fp:
        # no collisions; its particles drift past the windows of the
        # windowed sort, so the full-histogram sort's chunk is what counts
        space = {mcpm.TUNE_SORT_CHUNK: [0, 1024, 2048, 4096, 8192]}
    elif args.workload != "droplets":
        # the slabs sort by the windowed sort (a fixed chunk of particles)
        space.pop(mcpm.TUNE_SORT_CHUNK)
    best = {k: sim.get_tuning(k) for k in space if k != GRAPH_ROUNDS}
    if graph:
        best[GRAPH_ROUNDS] = 6
    state = {"applied": None}

    def apply(s):
        if state["applied"] == s:
            return
        for k, v in s.items():
            if k != GRAPH_ROUNDS:
                sim.set_tuning(k, v)
        if graph:
            sim.graph_capture(0, s[mcpm.TUNE_BANDS] + s[GRAPH_ROUNDS])
        state["applied"] = dict(s)

    def time_block():
        nonlocal step
        sim.sync()
        t0 = time.perf_counter()
        if graph:
            sim.graph_steps(step, args.steps)
            step += args.steps
        else:
            for _ in range(args.steps):
                sim.step(step)
                step += 1
        sim.sync()
        return time.perf_counter() - t0

    best, trials = tune(space, best, apply, time_block, names=NAMES,
                        log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps({"workload": args.workload, "graph_steps": graph,
                      "steps_per_block": args.steps,
                      "best": {NAMES[k]: v for k, v in best.items()},
                      "trials": [dict(t, knob=NAMES[t["knob"]]) for t in trials]}))


if __name__ == "__main__":
    main()

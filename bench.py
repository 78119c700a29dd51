#!/usr/bin/env python3
"""Throughput of the slab particle step on one MI355X.

Workload (default): the liquid-vapour planar interface — periodic 400 sigma
box, 4096 cells, 20M particles, n sigma^3 = 0.55 liquid slab in 0.02 vapour,
T = 0.7, Enskog collisions plus the screened-Poisson mean field, FP32
particles, FP64 field, dt = 0.02.  One step = collisions, field solve, kick,
streaming, periodic wrap and the stable re-sort into cells.

Prints one JSON line.  `value` is particle-steps per second with the
particles resident in HBM.  `roofline` describes the dominant kernel class,
timed with HIP events on the library's own stream; `cpu_baseline` is the CPU
oracle timed on this host on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "monte-carlo-particle-method_amd")
ORACLE = os.path.join(ROOT, "oracle")
sys.path.insert(0, PKG)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--particles", type=int, default=20_000_000)
    ap.add_argument("--cells", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--cpu-particles", type=int, default=20_000_000)
    ap.add_argument("--workload", default="interface",
                    choices=["interface", "droplets", "replicas", "dense"],
                    help="interface: the periodic 400 sigma slab the metric is quoted on; "
                         "droplets: the 512 x 512 periodic plane with 20 liquid discs; "
                         "replicas: 256 independent interface slabs in one batched context; "
                         "dense: the 20 sigma dense slab between diffuse walls at 1.0 and 1.5, "
                         "4000 cells x 2500 particles (10M)")
    ap.add_argument("--model", default="esmc", choices=["esmc", "fp"],
                    help="dense workload: Enskog collisions (ESMC) or the dense-gas "
                         "Fokker-Planck model (FP_dense)")
    ap.add_argument("--density", type=float, default=0.3, help="dense workload: n sigma^3")
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--unbatched-steps", type=int, default=2,
                    help="replicas: also time this many steps of the same replicas run one "
                         "context after the other")
    ap.add_argument("--pressure", action="store_true",
                    help="sample measure_pressure's face sums in every step (the reference "
                         "does once step > times * after)")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the learned-field phase (512 x 512 droplet plane)")
    ap.add_argument("--sur-particles", type=int, default=64_000_000)
    ap.add_argument("--sur-snapshots", type=int, default=64,
                    help="exact-solver snapshots the network trains on (one per step)")
    ap.add_argument("--sur-iters", type=int, default=2000, help="Adam steps of the training")
    ap.add_argument("--sur-steps", type=int, default=20, help="timed steps with and without the network")
    ap.add_argument("--no-graph", action="store_true",
                    help="time the steps one call at a time only (no HIP-graph replays)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one device copy of 1 GiB (a known-size kernel for "
                         "calibrating the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU decomposition is a design, not a measured path", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import mcpm
    import workloads as W

    if args.calibrate_copy:
        import torch
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 1 GiB read + 1 GiB written")
        del src, dst

    N = args.particles
    unbatched = None
    if args.workload == "replicas":
        common, reps, rarrs, metas = W.replicas(n_rep=args.replicas, seed0=0)
        counts = [a[0].size for a in rarrs]
        N = int(sum(counts))
        args.cells = int(sum(d["n_cells"] for d in reps))
        log(f"workload: {len(reps)} replicas, N={N} cells={args.cells} FP32")
        if args.unbatched_steps > 0:
            sims = []
            for d, a, m in zip(reps, rarrs, metas):
                s1 = mcpm.Simulation(d["params"])
                s1.upload(*a)
                s1.cell_update()
                s1.init_omega_max(m["n_gas"])
                sims.append(s1)
            for s1 in sims:
                s1.step(0)
            for s1 in sims:
                s1.sync()
            t0 = time.perf_counter()
            for k in range(args.unbatched_steps):
                for s1 in sims:
                    s1.step(1 + k)
            for s1 in sims:
                s1.sync()
            unbatched = 1000.0 * (time.perf_counter() - t0) / args.unbatched_steps
            log(f"unbatched: {unbatched:.2f} ms per step of all replicas, one context after another")
            del sims
        sim = mcpm.Simulation.batch(common, reps)
        sim.upload_batch(counts, *[np.concatenate([a[q] for a in rarrs]) for q in range(4)])
        del rarrs
        sim.cell_update()
        sim.init_omega_max_batch([m["n_gas"] for m in metas])
        p = common
    elif args.workload == "droplets":
        p, arrs, meta = W.droplets(N=64_000_000 if N == 20_000_000 else N)
        N = meta["N"]
        args.cells = p.n_cells * p.n_cells_x
        log(f"workload: droplets N={N} cells={p.n_cells_x}x{p.n_cells} dt={p.delta_t} FP32")
        sim = mcpm.Simulation(p)
        sim.upload_xy(*arrs)
        del arrs
        sim.cell_update()
        sim.init_omega_max(meta["n_gas"])
    elif args.workload == "dense":
        p, arrs, meta = W.dense_slab(n_density=args.density)
        if args.model == "fp":
            p.model = mcpm.MODEL_FP_DENSE
        N = meta["N"]
        args.cells = p.n_cells
        log(f"workload: dense slab N={N} cells={p.n_cells} n={args.density} model={args.model} "
            f"dt={p.delta_t:.5f} FP32")
        sim = mcpm.Simulation(p)
        sim.upload(*arrs)
        del arrs
        sim.cell_update()
        sim.init_omega_max(meta["n_gas"])
    else:
        p, arrs, meta = W.interface(N=N, n_cells=args.cells)
        log(f"workload: interface N={N} cells={args.cells} dt={p.delta_t} FP32")
        sim = mcpm.Simulation(p)
        sim.upload(*arrs)
        del arrs
        sim.cell_update()
        sim.init_omega_max(meta["n_gas"])
    if args.pressure:
        sim.measure_pressure(True)
    # ESMC interface and replica steps run as HIP-graph replays (no host wait
    # inside a step, bit for bit the same steps as one mcpm_step call each);
    # the first warmup step is a plain call (it sizes the capture)
    use_graph = (not args.no_graph and not args.pressure
                 and (args.workload in ("interface", "replicas")
                      or (args.workload == "dense" and args.model == "esmc")))
    step = 0
    t0 = time.time()
    for w in range(args.warmup):
        if use_graph and w > 0:
            if w == 1:
                sim.graph_capture()
            st = sim.graph_steps(step, 1)
        else:
            st = sim.step(step)
        step += 1
    if use_graph and args.warmup < 2:
        sim.graph_capture()
    sim.sync()
    log(f"warmup {args.warmup} steps in {time.time() - t0:.2f}s, last {st}")

    # the measured steps, no instrumentation inside
    sim.sync()
    t0 = time.perf_counter()
    if use_graph:
        gs = sim.graph_steps(step, args.steps)
        stats = [{k: v / args.steps for k, v in gs.items()}]
        step += args.steps
    else:
        stats = []
        for _ in range(args.steps):
            stats.append(sim.step(step))
            step += 1
    sim.sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = N * args.steps / elapsed
    ms_step = 1000.0 * elapsed / args.steps
    log(f"timed {args.steps} steps ({'graph replays' if use_graph else 'one call each'}): "
        f"{ms_step:.3f} ms/step")
    # one call per step against graph replays, in alternating blocks (the
    # step cost drifts up as omega_max grows, so blocks of each side alternate)
    mode_cmp = None
    if use_graph:
        blk = max(2, args.steps // 4)
        t_e = t_g = 0.0
        for _ in range(2):
            sim.sync()
            t0 = time.perf_counter()
            for _ in range(blk):
                sim.step(step)
                step += 1
            sim.sync()
            t_e += time.perf_counter() - t0
            t0 = time.perf_counter()
            sim.graph_steps(step, blk)
            step += blk
            sim.sync()
            t_g += time.perf_counter() - t0
        mode_cmp = {"steps_each": 2 * blk, "one_call_per_step_ms": 1000.0 * t_e / (2 * blk),
                    "graph_ms": 1000.0 * t_g / (2 * blk)}
        log(f"alternating blocks: {mode_cmp}")
    # the reference's cell_update takes the cell moments every step and its
    # post-processing samples them (cell_poly_velocity); the step above
    # leaves out what the collision model does not need: the same steps with
    # both added, for a second figure
    sampled_ms = None
    if args.workload in ("interface", "dense"):
        sim.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            sim.step(step)
            sim.compute_moments()
            sim.sample()
            step += 1
        sim.sync()
        sampled_ms = 1000.0 * (time.perf_counter() - t0) / args.steps
        log(f"with moments and sampling each step: {sampled_ms:.3f} ms/step")
    # as many steps again with HIP-event timers on the library's stream, for
    # the per-kernel figures (events around every launch cost a little time)
    sim.timing_enable(True)
    tstats = []
    for _ in range(args.steps):
        tstats.append(sim.step(step))
        step += 1
    sim.sync()
    # HIP events on the library's stream: the collision rounds per k_round
    # launch, the other classes per call; per-step figures = avg x calls / steps
    timers = {}
    for name, k in (("collide_rounds", mcpm.TIMER_COLLIDE), ("sort", mcpm.TIMER_SORT),
                    ("candidates", mcpm.TIMER_GEN), ("field", mcpm.TIMER_FIELD),
                    ("step", mcpm.TIMER_STEP), ("fp_velocity", mcpm.TIMER_FP_VELOCITY),
                    ("fp_closures", mcpm.TIMER_FP_CLOSURES)):
        ms, calls = sim.timing(k)
        timers[name] = {"avg_ms": ms, "calls": calls, "per_step_ms": ms * calls / args.steps}
    sim.timing_enable(False)
    timers = {k: v for k, v in timers.items() if v["calls"] > 0}
    # operation counts of the timed steps; round and revisit counts (and the
    # roofline's per-launch figures) of the same steps as the timers
    acc = float(np.mean([s["accepted"] for s in stats]))
    valid = float(np.mean([s["valid"] for s in stats]))
    rounds = float(np.mean([s["rounds"] for s in tstats]))
    revisits = float(np.mean([s["revisits"] for s in tstats]))
    t_valid = float(np.mean([s["valid"] for s in tstats]))
    t_acc = float(np.mean([s["accepted"] for s in tstats]))
    nbands = sim.get_tuning(mcpm.TUNE_BANDS)
    log(f"timed {args.steps} steps: {ms_step:.3f} ms/step, {value:.4g} particle-steps/s")
    log(f"phases: {json.dumps(timers)}; valid ops {valid:.0f}, accepted {acc:.0f}, rounds {rounds}")

    # Dominant kernel and its roofline.  Algorithmic bytes:
    #  sort (push pass 1 + scatter pass): read 16 B + read 16 B + write 16 B per
    #  particle (pass 1 reads the record, pass 2 reads it again and writes it);
    #  collision rounds (one k_check and one k_round launch per round): per valid operation
    #  the two particle records read once (2 x 16 B) and its 16-byte operation
    #  record, per accepted collision the two records written back (2 x 16 B),
    #  summed over the step's launches.
    sort_ms = timers["sort"]["per_step_ms"]
    coll_ms = timers.get("collide_rounds", {"per_step_ms": 0.0})["per_step_ms"]
    fpv_ms = timers.get("fp_velocity", {"per_step_ms": 0.0})["per_step_ms"]
    plane = args.workload == "droplets"
    rec = 32.0 if plane else 16.0       # FP32 record: (x, y, u1..3, pad) or (y, u1..3)
    passes = 2.0 if (plane and args.cells > 16384) else 1.0  # the plane's two-digit sort
    sort_bytes = 3.0 * rec * N * passes
    # collision rounds: each valid operation reads its 16-byte record once and
    # the two particle records once; an accepted collision writes both back
    coll_bytes = t_valid * (2.0 * rec + 16.0) + t_acc * 2.0 * rec
    step_alg = 44 if plane else 36      # 2 S_p + 4 with S_p = 20 (plane) or 16 (slab)
    if fpv_ms >= max(coll_ms, sort_ms):
        # k_fp_velocity: one compulsory read and write of each 16-byte record
        launches = 1.0
        dom, dom_ms, dom_bytes = "k_fp_velocity", fpv_ms, 2.0 * rec * N
    elif coll_ms >= sort_ms:
        # one reservation round = one k_check and one k_round launch
        launches = timers["collide_rounds"]["calls"] / args.steps
        dom, dom_ms, dom_bytes = "k_check + k_round", coll_ms, coll_bytes
    else:
        launches = 2.0 * passes
        dom, dom_ms, dom_bytes = "k_push (both passes)", sort_ms, sort_bytes
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    traffic = measured_traffic(dom, N, args.cells)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic,
                "launches_per_step": launches,
                "avg_launch_us": round(1000.0 * dom_ms / max(launches, 1.0), 2),
                "alg_bytes_per_launch": round(dom_bytes / max(launches, 1.0)),
                "step_alg_bytes_per_particle": step_alg,
                "step_frac": round(value * step_alg / 1e9 / HBM_PEAK_GBS, 4)}
    if dom == "k_check + k_round":
        roofline["random_access"] = random_access_roofline(
            N, int(rec), t_valid, t_acc, revisits, nbands, coll_ms, launches)

    cpu = None
    if not args.no_cpu_baseline and args.workload == "interface":
        cpu = cpu_baseline(args)
    surrogate = None
    if not args.no_surrogate and args.workload in ("interface", "droplets"):
        sim.close()
        del sim
        surrogate = surrogate_phase(args)
    if args.workload == "replicas":
        desc = (f"replicas: {args.replicas} independent periodic 400 sigma interface slabs in one "
                f"batched context, {N} particles, {args.cells} cells; per replica N log-uniform "
                "in [5e4, 5e5], n_liq in [0.45, 0.6], T in [0.6, 0.9], cells in [512, 2048], "
                "seeds 0..255; ESMC + screened-Poisson, dt=0.02")
    elif args.workload == "dense":
        desc = (f"dense slab: 20 sigma between diffuse walls at T=1.0 and 1.5, {p.n_cells} cells, "
                f"{N} particles, n={args.density}, T0=1.25, "
                + ("Enskog ESMC" if args.model == "esmc" else "dense-gas Fokker-Planck (FP_dense)")
                + ", dt=0.05 mean collision time, FP32 particles, FP64 moments")
    elif plane:
        desc = (f"droplets: periodic 128 x 128 sigma plane, {p.n_cells_x} x {p.n_cells} cells, "
                f"{N} particles, 20 discs n=0.55 in vapour 0.02, T=0.75, ESMC + "
                "screened-Poisson FEM, dt=0.01")
    else:
        desc = ("interface: periodic 400 sigma, 4096 cells, 20M particles, "
                "n=0.55/0.02, T=0.7, ESMC + screened-Poisson, dt=0.02")

    out = {
        "metric": "particle-steps/s on 1 MI355X (Enskog-Vlasov "
                  + ("plane step)" if plane else ("dense slab step)" if args.workload == "dense"
                                                  else "interface step)")),
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 particles / fp64 field and moments",
        "data": "synthetic (numpy Philox initial state)",
        "config": {"workload": desc,
                   "particles": N, "cells": args.cells, "parallelism": "1 GPU",
                   "pressure_sampling": bool(args.pressure),
                   "step_mode": ("HIP-graph replays (mcpm_graph_steps)" if use_graph
                                 else "one mcpm_step call per step"),
                   "graph_vs_one_call_per_step": mode_cmp,
                   "ms_per_step_with_moments_and_sampling": sampled_ms,
                   "unbatched_ms_per_step": unbatched,
                   "valid_ops_per_step": valid, "accepted_per_step": acc,
                   "reservation_rounds": rounds,
                   "phases_ms_per_step": {k: v["per_step_ms"] for k, v in timers.items()}},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "surrogate": surrogate,
    }
    print(json.dumps(out), flush=True)


def surrogate_phase(args):
    """The learned field on the 512 x 512 droplet plane (64M particles, 20
    discs at 0.55 in vapour 0.02, T = 0.75, periodic, screened-Poisson FEM).

    The network (width 64, four Fourier layers, 16 modes per axis, BF16 on
    MFMA) trains on the GPU on snapshots that the exact solver writes along
    the run (one per step), then is measured on later steps it never saw:
    force error of the Q1-interpolated field against the exact solve of the
    same density, the FEM residual of its field, the field time (deposit +
    network + residual check + core term, against deposit + conjugate
    gradients + core term), and the step time with the network standing in
    for the solve (exact solve every 50 steps) against the exact step."""
    import numpy as np

    import mcpm
    import workloads as W

    p, arrs, meta = W.droplets(N=args.sur_particles)
    N = meta["N"]
    n0, n1, m = p.n_cells_x, p.n_cells, 16
    log(f"surrogate: droplets N={N} {n0}x{n1}")
    sim = mcpm.Simulation(p)
    sim.upload_xy(*arrs)
    del arrs
    sim.cell_update()
    sim.init_omega_max(meta["n_gas"])
    step = 0
    for _ in range(3):
        sim.step(step)
        step += 1

    def timed_steps(k):
        nonlocal step
        sim.timing_enable(True)
        sim.sync()
        t0 = time.perf_counter()
        for _ in range(k):
            sim.step(step)
            step += 1
        sim.sync()
        ms = 1000.0 * (time.perf_counter() - t0) / k
        fms, calls = sim.timing(mcpm.TIMER_FIELD)
        sim.timing_enable(False)
        return ms, fms

    sim.poisson_solver()
    nd, pot = sim.nodes()
    # the best any m-mode spectral operator can do on this density (exact on
    # its modes, the best pointwise gain above them), before the run heats up
    floor_start = mcpm.spectral_force_floor(nd, p.dx, p.bst, p.sigma, m)
    net = mcpm.FieldSurrogate(n0, n1, m, mcpm.FieldSurrogate.init_params(m, seed=0))
    net.set_norm(nd.mean(), 1.0 / nd.std(), pot.mean(), 1.0 / pot.std())
    S = args.sur_snapshots
    net.reserve_snapshots(S)
    t0 = time.perf_counter()
    for _ in range(S):
        sim.step(step)
        step += 1
        net.add_snapshot(sim, 0.05)
    t_snap = time.perf_counter() - t0

    def held_out(k):
        # force error of the network's field against the exact solve of the
        # same density: of the whole field (the exact core term included on
        # both sides), and of the part the network stands in for (the FEM
        # solution, core term left out); the m-mode floor at the same state
        nonlocal step
        errs, errs_fem, res, floors = [], [], [], []
        for _ in range(k):
            sim.step(step)
            step += 1
            st = net.field(sim, accept_residual=-1.0)
            nd_k, pn = sim.nodes()
            sim.poisson_solver()
            _, pe = sim.nodes()
            fe, fn = mcpm.plane_force(pe, p.dx), mcpm.plane_force(pn, p.dx)
            fc = mcpm.plane_force(mcpm._periodic_nodes(
                mcpm.plane_core(nd_k[:-1, :-1], p.dx, p.bst, p.sigma)), p.dx)
            errs.append(float(np.linalg.norm(fn - fe) / np.linalg.norm(fe)))
            errs_fem.append(float(np.linalg.norm(fn - fe) / np.linalg.norm(fe + fc)))
            res.append(st["residual"])
            floors.append(mcpm.spectral_force_floor(nd_k, p.dx, p.bst, p.sigma, m)["floor"])
        return (float(np.mean(errs)), float(np.mean(res)), float(np.mean(errs_fem)),
                float(np.mean(floors)))

    err0, res0, errf0, _ = held_out(2)
    T = args.sur_iters
    t0 = time.perf_counter()
    l1 = net.train(T // 2, 1e-3, seed=1)
    l2 = net.train(T - T // 2, 3e-4, seed=2)
    t_train = time.perf_counter() - t0
    err1, res1, errf1, floor1 = held_out(4)
    log(f"surrogate: loss {l1[:10].mean():.4f} -> {l2[-20:].mean():.5f}; force error "
        f"{err0:.4f} -> {err1:.5f}; residual {res0:.3f} -> {res1:.4f}")

    # the field alone, network against exact, same deposit and core term
    for _ in range(2):
        net.field(sim, accept_residual=-1.0)
    sim.sync()
    t0 = time.perf_counter()
    for _ in range(10):
        net.field(sim, accept_residual=-1.0)
    sim.sync()
    f_net = 100.0 * (time.perf_counter() - t0)
    t0 = time.perf_counter()
    for _ in range(10):
        sim.poisson_solver()
    sim.sync()
    f_ex = 100.0 * (time.perf_counter() - t0)
    t0 = time.perf_counter()
    for _ in range(5):
        st_ws = net.field(sim, exact=True)
    sim.sync()
    f_warm = 200.0 * (time.perf_counter() - t0)

    # the step with and without the network, in alternating blocks (the
    # droplet system's collision work grows along the run)
    blocks, k = 4, max(1, args.sur_steps // 8)
    ex, sur = [], []
    for _ in range(blocks):
        sim.set_field_surrogate(None)
        ex.append(timed_steps(k))
        sim.set_field_surrogate(net, accept_residual=-1.0, every=50)
        sur.append(timed_steps(k))
    step_exact = float(np.mean([a for a, _ in ex]))
    field_exact = float(np.mean([b for _, b in ex]))
    step_sur = float(np.mean([a for a, _ in sur]))
    field_sur = float(np.mean([b for _, b in sur]))
    # with the residual gate at 1e-3 the exact solve runs whenever the
    # network's FEM residual is larger
    sim.set_field_surrogate(net, accept_residual=1e-3, every=50)
    sim.step(step)
    step += 1
    gated = sim.field_surrogate_stats()
    sim.set_field_surrogate(None)
    sim.close()
    P = n0 * n1
    flops = 4 * (2.0 * 2 * m * n0 * 64 * n1 + 2.0 * 64 * (64 + 2 * m) * P)  # MFMA work per forward
    return {
        "workload": f"droplets: periodic 128 x 128 sigma plane, {n0} x {n1} cells, {N} particles, "
                    "20 discs n=0.55 in vapour 0.02, T=0.75, ESMC + screened-Poisson FEM, dt=0.01",
        "network": {"layers": 4, "width": 64, "modes": m, "dtype": "bf16 MFMA, fp32 accumulate",
                    "params": int(mcpm.FieldSurrogate.param_count(m))},
        "training": {"snapshots": S, "snapshot_s": round(t_snap, 2), "iters": T,
                     "ms_per_iter": round(1000.0 * t_train / max(T, 1), 3),
                     "loss_first": float(l1[:10].mean()), "loss_last": float(l2[-20:].mean())},
        "force_rel_err": {"untrained": err0, "trained": err1, "held_out_steps": 4,
                          "fem_part_untrained": errf0, "fem_part_trained": errf1,
                          "floor_of_any_16_mode_operator": {"first_steps": floor_start["floor"],
                                                            "held_out_steps": floor1}},
        "fem_residual": {"untrained": res0, "trained": res1},
        "field_ms": {"exact": f_ex, "network": f_net, "exact_after_network": f_warm,
                     "speedup": f_ex / f_net,
                     "exact_solver": "FFT (periodic power-of-two plane)" if (
                         n0 & (n0 - 1)) == 0 and (n1 & (n1 - 1)) == 0 else "conjugate gradients"},
        "step_ms": {"blocks": f"{blocks} x {k} steps each, alternating",
                    "exact": step_exact, "network_every_50": step_sur,
                    "speedup": step_exact / step_sur,
                    "field_ms_per_step_exact": field_exact, "field_ms_per_step_network": field_sur},
        "residual_gate_1e-3": gated,
        "mfma_gflop_per_forward": round(flops / 1e9, 2),
    }


def random_access_roofline(n, rec, valid, acc, revisits, nbands, coll_ms, launches):
    """The reservation rounds against the device's random-access rates.

    Every access of an operation goes to an unpredictable particle, so the
    rounds are priced in accesses at the rates mcpm_probe_random_access
    measures on arrays of the same shape, summed over the access kinds (they
    share one memory system).  Per step, in k_check + k_round: 2 priority
    stores per valid operation outside band 0 (band 0 is stored by k_gen) and
    per revisit, 2 priority-word reads (k_check) and 2 mark-byte reads
    (k_round; one mark byte per 32 particles) per visit (operations +
    revisits), 2 record reads per operation
    (its one winning visit), 2 record writes per accepted collision, and the
    16-byte operation records streamed in (twice per visit, once per band
    admission) and out (each revisit) at the HBM peak.  The atomics and
    word re-reads of particles that several operations share, or whose mark
    byte another particle of the group set, are left out."""
    import mcpm
    rates = mcpm.probe_random_access(n, rec)
    visits = valid + revisits
    counts = {"scatter_word": 2.0 * valid * (1.0 - 1.0 / max(nbands, 1)) + 2.0 * revisits,
              "gather_word": 2.0 * visits,
              "gather_byte": 2.0 * visits,
              "gather_record": 2.0 * valid,
              "scatter_record": 2.0 * acc}
    floor_s = sum(counts[k] / rates[k] for k in counts)
    floor_s += 16.0 * (2.0 * visits + valid + revisits) / (HBM_PEAK_GBS * 1e9)
    floor_ms = 1000.0 * floor_s
    return {"rates_G_per_s": {k: round(v / 1e9, 2) for k, v in rates.items()},
            "accesses_per_step": {k: round(v) for k, v in counts.items()},
            "visits_per_step": round(visits),
            "floor_ms_per_step": round(floor_ms, 3),
            "floor_us_per_launch": round(1000.0 * floor_ms / max(launches, 1.0), 2),
            "frac": round(floor_ms / coll_ms, 4)}


def measured_traffic(kernel, particles, cells):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, copy-calibrated)
    from the committed rocprofv3 summary of this workload, or None."""
    path = os.path.join(ROOT, "profiles", "interface_20M_hbm_traffic.md")
    if particles != 20_000_000 or cells != 4096 or not os.path.exists(path):
        return None
    key = f"traffic_bytes_per_launch[{kernel}]:"
    for line in open(path):
        line = line.strip().strip("`")
        if line.startswith(key):
            return float(line.split(":", 1)[1])
    return None


def _host_cpu():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, avail


def cpu_baseline(args):
    """The CPU oracle built for speed (liboracle_fast.so: the same source at
    g++ -O3 -march=x86-64-v3) on the interface workload at the full mesh,
    timed on this host: (i) one thread on a 20M-particle sample, the serial
    execution of the reference; (ii) every host core this process may use
    (at most 16, the box's share), each thread stepping its own independent
    2M-particle sample (the oracle's collision order is sequential, so the
    cores share the work as independent systems, as batched replicas would),
    throughput summed over the threads."""
    import threading

    sys.path.insert(0, ORACLE)
    import oracle as orc
    import workloads as W

    model, avail = _host_cpu()
    cells = args.cells

    def make(n, seed):
        p, arrs, meta = W.interface(N=n, n_cells=cells, seed=seed)
        o = orc.OracleSim(p, fast=True)
        o.upload(*arrs)
        o.cell_update()
        o.init_omega_max(meta["n_gas"])
        o.step(0)  # one untimed step
        return o

    n1 = args.cpu_particles
    o = make(n1, 1)
    t0 = time.perf_counter()
    for s in range(1, 1 + args.cpu_steps):
        o.step(s)
    dt1 = time.perf_counter() - t0
    v1 = n1 * args.cpu_steps / dt1
    o.close()
    del o
    log(f"cpu oracle, 1 thread: {args.cpu_steps} steps of {n1} particles in {dt1:.2f}s -> "
        f"{v1:.4g} particle-steps/s")
    T = max(1, min(16, avail))
    nT = max(100_000, min(2_000_000, n1 // 10))
    sims = [make(nT, 100 + t) for t in range(T)]
    times = [0.0] * T

    def run(t):
        t0 = time.perf_counter()
        for s in range(1, 1 + args.cpu_steps):
            sims[t].step(s)
        times[t] = time.perf_counter() - t0

    th = [threading.Thread(target=run, args=(t,)) for t in range(T)]
    t0 = time.perf_counter()
    for x in th:
        x.start()
    for x in th:
        x.join()
    dtT = time.perf_counter() - t0
    vT = T * nT * args.cpu_steps / dtT
    for o in sims:
        o.close()
    log(f"cpu oracle, {T} threads: {args.cpu_steps} steps of {T} x {nT} particles in {dtT:.2f}s "
        f"-> {vT:.4g} particle-steps/s")
    return {"value": vT, "unit": "particle-steps/s", "cores": T, "kind": "port",
            "sample": f"interface workload, {cells} cells: {T} threads each stepping its own "
                      f"{nT}-particle system {args.cpu_steps} steps after 1 untimed step (C++ "
                      "oracle, liboracle_fast.so = g++ -O3 -march=x86-64-v3, threads over "
                      "independent systems)",
            "host_cpu": model, "host_cores_available": avail,
            "single_thread": {"value": v1, "cores": 1,
                              "sample": f"interface workload at {n1} particles, {cells} cells, "
                                        f"{args.cpu_steps} steps after 1 untimed step, one "
                                        "thread (the reference is serial)"}}


if __name__ == "__main__":
    main()

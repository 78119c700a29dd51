#!/usr/bin/env python3
"""Run a reference settings file through the library the way the reference's
time loop does (main.cpp:389-669):

  read the file (readSettingsFile)  -> mcpm_read_settings
  initial particles (initialization) -> mcpm_initialize
  omega_max (main.cpp:250-269)       -> mcpm_init_omega_max(n)
  per step: cell_poly_velocity once step > after, write_post_processing when
  step > after and step % every == 0 (or step == init_step), then the step
  itself (velocity model, long-range force, stream and boundary, re-binning).

    python run_settings.py injection.in --steps 200 --out PostProc

The particles stay on the device; only the PostProc files come back.  The
reference's thermostat call for problem injection (main.cpp:429-437) does
nothing for that problem (its body is commented out, Main_Functions.cpp:
1679-1690), so the loop has no counterpart for it.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import mcpm  # noqa: E402


def run(path: str, steps: int | None = None, out: str = "PostProc", seed: int = 1,
        precision: int = mcpm.FP64, log=print) -> dict:
    s = mcpm.read_settings(path)
    p = mcpm.Params.from_c(s.params)
    p.seed = seed
    p.precision = precision
    s.params = p.to_c()
    sim = mcpm.Simulation(p)
    sim.initialize(s)
    sim.init_omega_max(s.n)
    os.makedirs(out, exist_ok=True)
    n_steps = s.num_steps if steps is None else steps
    t0 = time.time()
    written = []
    acc = 0
    for step in range(n_steps):
        if step > s.after:
            sim.sample()
        if (step > s.after and step % s.every == 0) or step == s.init_step:
            sim.write_postproc(s, step, out)
            written.append(step)
        acc += sim.step(step)["accepted"]
        if step % max(1, s.every) == 0:
            log(f"step {step}: {time.time() - t0:.2f} s, accepted so far {acc}")
    sim.sync()
    return {"steps": n_steps, "written": written, "accepted": acc,
            "particles": s.n_particles, "seconds": time.time() - t0}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("settings")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--out", default="PostProc")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--fp32", action="store_true")
    args = ap.parse_args()
    r = run(args.settings, args.steps, args.out, args.seed,
            mcpm.FP32 if args.fp32 else mcpm.FP64)
    print(r)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""UVLM hot-path benchmark on one MI355X.

A step is one UVLM timestep (morph -> shed -> matrix -> RHS -> solve ->
loads -> advect, simple_simulator.cpp:240-286) of the workload, run by the
C++ SimpleSimulator on the gfx950 kernels.  Default workload: NACA4412
cambered wing, 20x80 rings, root-hinged flapping +/-30 deg at k = 0.3 with
5 deg pitch, reference "RK4" wake advection, steps 0..1000 (the wake grows to
80,000 rings).  `--warmup W` untimed steps (from step 0), then `--steps K`
timed steps.

Prints one JSON line.  `value` is Biot-Savart pairs per second evaluated by
the GPU over the timed steps (one pair = one segment evaluation, 56 FLOP);
`timesteps_per_s` and the reference-loop pair model rate are reported beside
it.  `cpu_baseline` times the oracle (a C restatement of the reference's
loop, OpenMP over targets like the reference) on a bounded sample of the
same sums at the final state, on the host's cores.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_PAIR = 56.0
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector (and matrix) peak, AMD published

PRECISIONS = {"f64exact": 0x0, "f64": 0x1, "f32c": 0x2}
SOLVES = {"f64": 0, "f32ir": 1, "bf16ir": 2}


def cpu_baseline(sim, case, seconds):
    """Oracle flow velocity ((-ff + bound) + wake, advect.h:16-24) at a sample
    of wake nodes of the final GPU state, timed on the host cores."""
    import oracle
    oracle.build()
    wing = sim.wing_nodes()
    g = sim.gamma()
    wn, wg = sim.wake()
    targets = wn.reshape(-1, 3)
    n_src = g.size + wg.size

    def run(nt):
        idx = np.linspace(0, targets.shape[0] - 1, nt).astype(int)
        x = targets[idx]
        t0 = time.perf_counter()
        oracle.induced_velocity(x, wing, g)
        if wg.size:
            oracle.induced_velocity(x, wn, wg)
        return time.perf_counter() - t0, 4.0 * nt * n_src

    dt, pairs = run(16)
    rate = pairs / max(dt, 1e-9)
    nt = int(min(targets.shape[0], max(16, seconds * rate / (4.0 * n_src))))
    dt, pairs = run(nt)
    return {"value": pairs / dt, "unit": "pairs/s", "cores": oracle.num_threads(),
            "kind": "port",
            "sample": (f"oracle flow velocity at {nt} of {targets.shape[0]} wake nodes "
                       f"against all {n_src} rings of the final state "
                       f"({pairs:.3e} pairs in {dt:.1f} s)")}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed UVLM timesteps")
    ap.add_argument("--warmup", type=int, default=1, help="untimed steps before")
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c5"])
    ap.add_argument("--precision", default="f64exact", choices=sorted(PRECISIONS))
    ap.add_argument("--solve", default="f64", choices=sorted(SOLVES))
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py runs on exactly one GPU: the UVLM step of one case is one "
              "GPU's work here (multi-GPU sharding by targets is a design note, "
              "not a measured configuration)", file=sys.stderr)
        sys.exit(2)

    from uvlm_amd import Simulation
    from uvlm_amd import workloads as W

    total = args.warmup + args.steps
    case = W.WORKLOADS[args.workload](steps=total - 1)
    case["precision"] = PRECISIONS[args.precision]
    case["solve_mode"] = SOLVES[args.solve]
    sim = Simulation(case, total - 1)

    for s in range(args.warmup):
        sim.step(s)
    sim.sync()
    e0, r0 = sim.pairs()
    sim.profile(True)
    sim.profile_read()
    t0 = time.perf_counter()
    for s in range(args.warmup, total):
        sim.step(s)
        if (s - args.warmup) % 100 == 99:
            print(f"# step {s}", file=sys.stderr, flush=True)
    sim.sync()
    T = time.perf_counter() - t0
    launches, kms, kpairs = sim.profile_read()
    sim.profile(False)
    e1, r1 = sim.pairs()

    pairs = e1 - e0
    achieved = FLOP_PER_PAIR * kpairs / (kms / 1e3) / 1e12 if kms > 0 else 0.0
    loads = sim.loads()
    out = {
        "metric": "Biot-Savart pairs/s & timesteps/s vs plain FP64 GPU, CL/CT rel err <=1e-4",
        "value": pairs / T,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * T / args.steps,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": {"f64exact": "f64", "f64": "f64", "f32c": "f32"}[args.precision],
        "data": "synthetic",
        "config": {"workload": case["name"], "wings": len(case["wings"]),
                   "rings_per_wing": [case["rows"], case["cols"]],
                   "advect": case["advect"], "dt": case["dt"],
                   "timed_steps": [args.warmup, total - 1],
                   "final_wake_rings": len(case["wings"]) * (total - 1) * case["cols"],
                   "precision": args.precision, "solve": args.solve},
        "timesteps_per_s": args.steps / T,
        "reference_model_pairs_per_s": (r1 - r0) / T,
        "final_CD_CL": loads[-1].tolist() if len(loads) else None,
        "roofline": {"bound": "valu", "kernel": "k_nbody (Biot-Savart N-body)",
                     "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "launches": launches, "kernel_ms": kms,
                     "kernel_share_of_step_time": kms / 1e3 / T},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sim, case, args.cpu_seconds)
        out["cpu_baseline"]["gpu_over_cpu_pairs_per_s"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

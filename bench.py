#!/usr/bin/env python3
"""UVLM hot-path benchmark on one MI355X.

A step is one UVLM timestep (morph -> shed -> matrix -> RHS -> solve ->
loads -> advect, simulator/simple_simulator.cpp:240-286) run by the C++
SimpleSimulator on the gfx950 kernels.  Default workload: NACA4412-cambered
wing, 20x80 rings, root-hinged flapping +/-30 deg at k = 0.3 with 5 deg
pitch, reference "RK4" wake advection, 1000 steps (the wake grows to 80,000
rings).  `--warmup W` untimed steps (from step 0), then `--steps K` timed
steps; the defaults time steps 1..1000 of the 1001-iteration run.  The
default method is FP64 throughout: segment-chain N-body in FP64 and the
kept-inverse FP64 solve (iterative refinement from the previous step's
circulations to max|dx| <= 1e-14 max|x|); `--solve f64` selects the blocked
LU instead.

Prints one JSON line.  `value` is Biot-Savart pairs per second in the
reference's counting: the pairs the reference loop evaluates for the timed
steps, 4 [N^2 + N Nw + L (N + Nw) + S V (N + Nw)] per step (one pair = one
BiotSavartLaw call, vortex.cpp:21-39), divided by the wall time of those
steps with inputs resident on the GPU.  `roofline` describes the dominant
kernel (the FP64 segment-chain N-body): achieved = 56 FLOP x segment pairs
it evaluated / its summed duration from HIP events on its stream.
`cpu_baseline` times the oracle (a C restatement of the reference's loop,
OpenMP over targets like the reference) on a bounded sample of the same
sums at the final state, on the host's cores.

Other modes: --workload c1|c3|c5 (single-case workloads), --workload c4
(256 batched cases), --precision, --solve, --farfield THETA (near/far split
of the wake sum, audited), --mlp-every K (in-run training of the far-field
correction network every K steps); --compare also runs the plain FP64
method (exact sums, LU) over the same steps and reports the speed-up and
the CL/CT differences; --calibrate-copy runs a plain device copy of known
size for the PMC traffic calibration.
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

METRIC = "Biot-Savart pairs/s & timesteps/s vs plain FP64 GPU, CL/CT rel err ≤1e-4"
FLOP_PER_PAIR = 56.0
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector peak (AMD published)
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector peak (AMD published)
# dominant kernel and its peak per precision
ROOFLINE_KERNEL = {"f64exact": ("Biot-Savart ring pairs (k_nbody)", PEAK_FP64_TFLOPS),
                   "f64": ("Biot-Savart N-body (k_chain_f64)", PEAK_FP64_TFLOPS),
                   "f32c": ("Biot-Savart N-body (k_chain_f32c)", PEAK_FP32_TFLOPS)}
PRECISIONS = {"f64exact": 0x0, "f64": 0x1, "f32c": 0x2}
SOLVES = {"f64": 0, "f32ir": 1, "bf16ir": 2, "fullpiv": 3, "inv": 4}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_chain_f64_naca4412_20x80_rk4_traffic.json")


def cpu_baseline(sim, seconds):
    """Oracle flow velocity of bound + wake rings (advect.h:16-24) at a sample
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

    nt = 16
    dt, pairs = run(nt)
    while dt < 1.0 and nt < targets.shape[0]:  # calibrate on >= 1 s of work
        nt = min(targets.shape[0], nt * 4)
        dt, pairs = run(nt)
    if dt < seconds and nt < targets.shape[0]:
        nt = int(min(targets.shape[0], nt * seconds / dt))
        dt, pairs = run(nt)
    return {"value": pairs / dt, "unit": "pairs/s", "cores": oracle.num_threads(),
            "kind": "port",
            "sample": (f"oracle sum of all {n_src} bound and wake rings at {nt} of the "
                       f"{targets.shape[0]} wake nodes of the final state "
                       f"({pairs:.3e} pairs in {dt:.1f} s)")}


def traffic_from_profile():
    """HBM bytes per launch of the N-body from the committed PMC summary."""
    try:
        with open(TRAFFIC_FILE) as f:
            d = json.load(f)
        return float(d["hbm_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def sync_all():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except ImportError:
        pass


def calibrate_copy(nbytes=1 << 28):
    """A plain device-to-device copy of known size (for FETCH/WRITE_SIZE)."""
    import torch
    x = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": 3}))


def plain_run(args):
    """The plain FP64 method (exact sums, LU) over the same steps: wall time
    of the timed steps and the CL/CD history."""
    from uvlm_amd import Simulation
    from uvlm_amd import workloads as W

    total = args.warmup + args.steps
    case = W.WORKLOADS[args.workload](steps=total - 1)
    case["precision"] = PRECISIONS["f64"]
    case["solve_mode"] = SOLVES["f64"]
    sim = Simulation(case, total - 1)
    for s in range(args.warmup):
        sim.step(s)
    sim.sync()
    t0 = time.perf_counter()
    for s in range(args.warmup, total):
        sim.step(s)
        if (s - args.warmup) % 100 == 99:
            print(f"# plain step {s}", file=sys.stderr, flush=True)
    sim.sync()
    T = time.perf_counter() - t0
    loads = sim.loads()
    sim.close()
    return T, loads


def load_differences(loads, ref, early=30):
    """CL/CT differences relative to each wing's largest |CL|, |CT| of the
    reference run: pointwise over steps 0..early (before the wake's
    sensitivity to rounding has grown) and of the means over the run."""
    n = min(len(loads), len(ref))
    out = {}
    for q, name in ((1, "CL"), (0, "CT")):
        a, b = loads[:n, :, q], ref[:n, :, q]
        scale = np.maximum(np.abs(b).max(axis=0), 1e-300)
        out[f"{name}_max_rel_err_steps_0_{early}"] = (
            np.abs(a[:early + 1] - b[:early + 1]).max(axis=0) / scale).tolist()
        out[f"{name}_mean_rel_err"] = (np.abs(a.mean(axis=0) - b.mean(axis=0)) / scale).tolist()
        out[f"{name}_max_rel_err_all_steps"] = (np.abs(a - b).max(axis=0) / scale).tolist()
    return out


def run_single(args):
    from uvlm_amd import Simulation
    from uvlm_amd import workloads as W
    from uvlm_amd.gpu import mlp_init

    total = args.warmup + args.steps
    case = W.WORKLOADS[args.workload](steps=total - 1)
    case["precision"] = PRECISIONS[args.precision]
    case["solve_mode"] = SOLVES[args.solve]
    sim = Simulation(case, total - 1)
    if args.farfield > 0:
        sim.set_farfield(args.farfield, args.audit, args.ff_tol)
        if args.ff_reach > 0:
            sim.set_farfield_focus(args.ff_reach, args.ff_theta_min)
    for s in range(args.warmup):
        sim.step(s)
    sim.sync()
    sync_all()
    e0, r0 = sim.pairs()
    sim.profile(True)
    sim.profile_read()
    w = mlp_init(5) if args.mlp_every > 0 else None
    learned = 0
    net_pairs = net_trips = 0  # counters survive retraining (which resets them)
    t0 = time.perf_counter()
    for s in range(args.warmup, total):
        sim.step(s)
        if (w is not None and args.farfield > 0 and s >= args.mlp_start
                and (s - args.mlp_start) % args.mlp_every == 0):
            if learned:
                p_, t_ = sim.farfield_mlp_stats()
                net_pairs += p_
                net_trips += t_
            w, _ = sim.farfield_learn(w, args.farfield, args.mlp_samples, args.mlp_iters,
                                      512, 1e-3, seed=s)
            learned += 1
        if (s - args.warmup) % 100 == 99:
            print(f"# step {s}", file=sys.stderr, flush=True)
    sim.sync()
    sync_all()
    T = time.perf_counter() - t0
    launches, kms, kpairs = sim.profile_read()
    sim.profile(False)
    e1, r1 = sim.pairs()
    loads = sim.loads()
    achieved = FLOP_PER_PAIR * kpairs / (kms / 1e3) / 1e12 if kms > 0 else 0.0
    traffic = traffic_from_profile() if (args.workload == "c2" and args.precision == "f64"
                                         and args.solve == "inv" and args.steps == 1000
                                         and args.warmup == 1) else None
    nw = len(case["wings"])
    out = {
        "metric": METRIC,
        "value": (r1 - r0) / T,
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
        "config": {"workload": f"{case['name']}_{total - 1}steps", "wings": nw,
                   "rings_per_wing": [case["rows"], case["cols"]],
                   "advect": case["advect"], "dt": case["dt"],
                   "timed_steps": [args.warmup, total - 1],
                   "final_wake_rings": nw * (total - 1) * case["cols"],
                   "precision": args.precision, "solve": args.solve,
                   "farfield_theta": args.farfield if args.farfield > 0 else None},
        "timesteps_per_s": args.steps / T,
        "gpu_evaluated_pairs_per_s": (e1 - e0) / T,
        "final_CD_CL": loads[-1].tolist() if len(loads) else None,
        "roofline": {"bound": "valu", "kernel": ROOFLINE_KERNEL[args.precision][0],
                     "achieved": achieved, "peak": ROOFLINE_KERNEL[args.precision][1],
                     "unit": "TFLOP/s",
                     "frac": achieved / ROOFLINE_KERNEL[args.precision][1], "traffic": traffic,
                     "launches": launches, "avg_launch_us": 1e3 * kms / max(launches, 1),
                     "flop_per_launch": FLOP_PER_PAIR * kpairs / max(launches, 1),
                     "kernel_share_of_step_time": kms / 1e3 / T},
    }
    if args.farfield > 0:
        err, audits, active = sim.farfield_stats()
        out["farfield"] = {"theta": args.farfield, "audit_every": args.audit,
                           "last_audit_rel_err": err, "audits": audits, "active": active}
        if args.ff_reach > 0:
            out["farfield"]["graded"] = {"reach": args.ff_reach, "theta_min": args.ff_theta_min}
        if w is not None:
            pairs, trips = sim.farfield_mlp_stats()
            pairs += net_pairs
            trips += net_trips
            out["farfield"]["network"] = {"trained": learned, "every": args.mlp_every,
                                          "samples": args.mlp_samples,
                                          "iters": args.mlp_iters, "corrected_pairs": pairs,
                                          "exact_fallbacks": trips}
    if args.solve == "inv":
        sw, rf, rb, _ = sim.inverse_stats()
        out["inverse_solve"] = {"sweeps": sw, "newton_schulz_refreshes": rf, "rebuilds": rb}
    if args.compare:
        Tp, lp = plain_run(args)
        out["vs_plain_fp64"] = {"plain_ms_per_step": 1e3 * Tp / args.steps,
                                "timesteps_per_s_ratio": Tp / T,
                                **load_differences(loads, lp)}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sim, args.cpu_seconds)
    return out


def run_batch(args):
    from uvlm_amd import workloads as W
    from uvlm_amd.gpu import Batch

    total = args.warmup + args.steps
    cases = W.batch_c4(steps=total - 1)
    b = Batch(cases, max_steps=total - 1, advect="RK2")
    b.run(args.warmup, graph=not args.no_graph)
    chosen, times = b.autotune()
    b.sync()
    p0 = b.pairs()
    sync_all()
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        k = min(50, args.steps - done)
        b.run(k, graph=not args.no_graph)
        done += k
        print(f"# step {args.warmup + done - 1}", file=sys.stderr, flush=True)
    b.sync()
    sync_all()
    T = time.perf_counter() - t0
    p1 = b.pairs()
    ref = 0.0
    for c in cases:
        ref += W.pair_model(c, steps=total - 1, stages=2) - W.pair_model(c, steps=args.warmup - 1,
                                                                          stages=2)
    return {
        "metric": METRIC, "value": ref / T, "unit": "pairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * T / args.steps,
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"batch_256_cases_rk2_{total - 1}steps", "cases": len(cases),
                   "timed_steps": [args.warmup, total - 1], "autotuned_variant": chosen,
                   "hipgraph": not args.no_graph},
        "case_steps_per_s": len(cases) * args.steps / T,
        "gpu_evaluated_pairs_per_s": (p1 - p0) / T,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed UVLM timesteps")
    ap.add_argument("--warmup", type=int, default=1, help="untimed steps before")
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--precision", default="f64", choices=sorted(PRECISIONS))
    ap.add_argument("--solve", default="inv", choices=sorted(SOLVES))
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--farfield", type=float, default=0.0,
                    help="near/far split of the wake sum with this theta (0: exact sums)")
    ap.add_argument("--audit", type=int, default=50, help="far field: exact audit every N steps")
    ap.add_argument("--ff-tol", type=float, default=1e-3, help="far field: audit tolerance")
    ap.add_argument("--ff-reach", type=float, default=0.0,
                    help="far field: graded criterion beyond this many wing radii (0: off)")
    ap.add_argument("--ff-theta-min", type=float, default=2.0,
                    help="far field: smallest theta of the graded criterion")
    ap.add_argument("--mlp-every", type=int, default=0,
                    help="far field: train the correction network every K steps (0: off)")
    ap.add_argument("--mlp-start", type=int, default=20, help="first training step")
    ap.add_argument("--mlp-samples", type=int, default=60000)
    ap.add_argument("--mlp-iters", type=int, default=1000)
    ap.add_argument("--compare", action="store_true",
                    help="also time the plain FP64 method over the same steps")
    ap.add_argument("--no-graph", action="store_true",
                    help="batch: launch each step's kernels directly instead of a hipGraph")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: the UVLM step of a case runs on a "
              "single MI355X (sharding over GPUs is a design note, not a measured "
              "configuration)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.warmup < 1:
        args.warmup = 1
    out = run_batch(args) if args.workload == "c4" else run_single(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

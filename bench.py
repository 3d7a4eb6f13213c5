#!/usr/bin/env python3
"""UVLM hot-path benchmark on one MI355X.

A step is one UVLM timestep (morph -> shed -> matrix -> RHS -> solve ->
loads -> advect, simulator/simple_simulator.cpp:240-286) run by the C++
SimpleSimulator on the gfx950 kernels.

Default workload: two tandem NACA 8300-camber wings, 16 x 64 rings each,
gap 1.5 c, flapping -(pi/4) cos(wt + n 90 deg) at k = 0.1, Euler wake
advection, 2000 steps (2001 iterations; the wake grows to 256,000 rings).
The whole run is done twice in this process:

* the fast form: near/far split of the wake sum (lattice-block moment
  expansions for far clusters, theta 4 near the wings graded beyond 1.5
  wing radii to 1.2, exact audits every 50 steps), a correction network
  for the far clusters trained in the run on the run's own wake (steps
  100, 700, 1300, 1900) and evaluated once per (target tile, cluster) on
  the BF16 matrix cores, FP32 compensated near sums with FP64 near-core
  re-evaluation, the wings' moment expansions for wake nodes beyond the
  first 24 rows, kept-inverse solve (FP32 inverse, FP64 residuals);
* the plain FP64 method: every sum exact (FP64 segment chains), blocked
  FP64 LU.

`--steps K --warmup W` choose the timed window inside the full-length run,
never the run length: the last K steps of the run, with the W steps before
them untimed (everything earlier is untimed too).  The window is bracketed
by device synchronisations on both sides.  `value` is timesteps/s of the
fast form over the window; `vs_plain_fp64` gives both runs' timesteps/s over
the whole run (steps 1..2000) and over the same window, and the CL/CT
differences.  `roofline` describes the fast form's dominant kernel, the
near-block sum `k_ff_near` (56 FLOP per (target, segment) pair, pairs
counted on the device, duration from HIP events on its stream over every
launch of the fast run).  `c2_naca4412` is a second block: the NACA4412
20 x 80 wing, RK4, 1000 steps, FP64 segment chains and the kept-inverse
solve, with the roofline of `k_chain_f64`.  `cpu_baseline` times the oracle
(a C restatement of the reference's loop, OpenMP over targets like the
reference) on a bounded sample of the plain method's sums at the final
state of the plain run, on the host's cores.

Other modes: --workload c1|c2|c3|c5 runs one configuration alone (same
window rules, `--precision`, `--solve`, `--farfield` choose the method),
--workload c4 the 256 batched cases; --no-plain, --no-c2, --no-cpu skip
legs; --net-compare also times the fast form without the correction
network (after the plain run); --spread adds a second FP64 run of
the tandem workload that differs
from the plain one only by rounding (kept-inverse solve instead of the LU)
and reports, per 100-step window, the fast run's CL/CT differences beside
the two FP64 runs' spread; --calibrate-copy runs a plain device copy of known size for the PMC
traffic calibration.
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

METRIC = "Biot-Savart pairs/s & timesteps/s vs plain FP64 GPU, CL/CT rel err ≤1e-4"
FLOP_PER_PAIR = 56.0
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector peak (AMD published)
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector peak (AMD published)
PRECISIONS = {"f64exact": 0x0, "f64": 0x1, "f32c": 0x2}
SOLVES = {"f64": 0, "f32ir": 1, "bf16ir": 2, "fullpiv": 3, "inv": 4}
PROFILES = os.path.join(ROOT, "profiles")
TRAFFIC = {  # committed PMC summaries (HBM bytes per launch) of the default runs
    "k_chain_f64": os.path.join(PROFILES, "k_chain_f64_naca4412_20x80_rk4_traffic.json"),
    "k_ff_near": os.path.join(PROFILES, "k_ff_near_tandem_2x16x64_traffic.json"),
}
# The fast form of the tandem workload.  The correction network is trained
# in the run from its own wake (first at step `first`, then warm-started
# every `every` steps) and corrects the far clusters from then on.
NET = {"first": 100, "every": 600, "samples": 12000, "iters_first": 80, "iters": 30,
       "batch": 4096, "lr": 2e-3}
FAST = {"precision": "f32c", "solve": "inv", "farfield": 4.0, "ff_reach": 1.5,
        "ff_theta_min": 1.2, "audit": 50, "ff_tol": 1e-3, "mlp": NET}
PLAIN = {"precision": "f64", "solve": "f64", "farfield": 0.0}


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def traffic_of(kernel):
    try:
        with open(TRAFFIC[kernel]) as f:
            return float(json.load(f)["hbm_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def calibrate_copy(nbytes=1 << 28):
    """A plain device-to-device copy of known size (for FETCH/WRITE_SIZE)."""
    import torch
    x = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": 3}))


def run_case(case, method, window, warmup, label):
    """Every step 0..S of the workload (S = case["steps"]).  Returns wall
    times of the whole run (steps 1..S) and of the window (the last
    `window` steps), the load history, and the kernels' event profiles
    over the whole run.  The simulation is returned open (for the CPU leg)."""
    from uvlm_amd import Simulation

    S = int(case["steps"])
    case = dict(case)
    case["precision"] = PRECISIONS[method["precision"]]
    case["solve_mode"] = SOLVES[method["solve"]]
    sim = Simulation(case, S)
    if method.get("farfield", 0) > 0:
        sim.set_farfield(method["farfield"], method.get("audit", 50), method.get("ff_tol", 1e-3))
        if method.get("ff_reach", 0) > 0:
            sim.set_farfield_focus(method["ff_reach"], method.get("ff_theta_min", 2.0))
    K = max(1, min(window, S))
    first = S - K + 1  # first timed step
    net = method.get("mlp")
    w, trainings, t_train = None, 0, 0.0
    if net:
        from uvlm_amd.gpu import mlp_init
        w = mlp_init(5)
    sim.profile(True)
    sim.profile_read()
    sim.profile_read_near()
    sim.step(0)
    sim.sync()
    t_run0 = time.perf_counter()
    t_win = None
    for s in range(1, S + 1):
        if s == first:
            sim.sync()
            tw0 = time.perf_counter()
        sim.step(s)
        if net and s >= net["first"] and (s - net["first"]) % net["every"] == 0:
            t0 = time.perf_counter()
            w, _ = sim.farfield_learn(w, method["farfield"], max_samples=net["samples"],
                                      iters=net["iters_first"] if trainings == 0 else net["iters"],
                                      batch=net["batch"], lr=net["lr"], seed=s)
            t_train += time.perf_counter() - t0
            trainings += 1
        if s % 200 == 0:
            log(f"{label} step {s}")
    sim.sync()
    t_end = time.perf_counter()
    t_win = t_end - tw0
    t_run = t_end - t_run0
    chain = sim.profile_read()
    near = sim.profile_read_near()
    sim.profile(False)
    out = {"sim": sim, "S": S, "K": K, "first": first, "warmup": min(warmup, first - 1),
           "t_run": t_run, "t_win": t_win, "loads": sim.loads(), "chain": chain,
           "near": near, "pairs": sim.pairs()}
    if method.get("farfield", 0) > 0:
        out["ff"] = sim.farfield_stats()
    if net:
        pairs, trips = sim.farfield_mlp_stats()
        out["net"] = {"trainings": trainings, "training_s": t_train, "pairs_corrected": pairs,
                      "pairs_tripped_to_exact": trips}
    if method["solve"] == "inv":
        out["inv"] = sim.inverse_stats()
    return out


def roofline(kernel, prof, t_run, peak, traffic_key=None):
    launches, kms, kpairs = prof
    achieved = FLOP_PER_PAIR * kpairs / (kms / 1e3) / 1e12 if kms > 0 else 0.0
    return {"bound": "valu", "kernel": kernel, "achieved": achieved, "peak": peak,
            "unit": "TFLOP/s", "frac": achieved / peak,
            "traffic": traffic_of(traffic_key) if traffic_key else None,
            "launches": launches, "avg_launch_us": 1e3 * kms / max(launches, 1),
            "flop_per_launch": FLOP_PER_PAIR * kpairs / max(launches, 1),
            "pairs_per_launch": kpairs / max(launches, 1),
            "kernel_share_of_run_time": kms / 1e3 / t_run}


def load_differences(loads, ref, early=30):
    """CL/CT differences relative to each wing's largest |CL|, |CT| of the
    reference run: pointwise over steps 0..early (before the wake's
    sensitivity to rounding has grown), of the run means, and pointwise
    over every step."""
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


def windowed_max(loads, ref, width=100):
    """Largest CL/CT difference relative to each wing's largest |CL|, |CT|
    of `ref`, per window of `width` steps: [quantity][window][wing]."""
    n = min(len(loads), len(ref))
    out = {}
    for q, name in ((1, "CL"), (0, "CT")):
        a, b = loads[:n, :, q], ref[:n, :, q]
        scale = np.maximum(np.abs(b).max(axis=0), 1e-300)
        d = np.abs(a - b) / scale
        out[name] = [d[k:k + width].max(axis=0).tolist() for k in range(0, n, width)]
    return out


def spread_report(fast, plain, plain2, width=100):
    """Pointwise differences of the fast run from the plain run, beside the
    spread of two FP64 runs that differ only by rounding (the plain method
    with the kept-inverse solve in place of the LU), over the same steps."""
    f, p, q = fast["loads"], plain["loads"], plain2["loads"]
    out = {"second_fp64_run": "FP64 segment chains, kept-inverse FP64 solve",
           "window_steps": width,
           "fast_vs_plain": windowed_max(f, p, width),
           "fp64_vs_fp64": windowed_max(q, p, width),
           "fast_vs_second_fp64": windowed_max(f, q, width),
           "fp64_vs_fp64_all_steps": load_differences(q, p)}
    ratio = {}
    for name in ("CL", "CT"):
        a = np.array(out["fast_vs_plain"][name])
        b = np.array(out["fp64_vs_fp64"][name])
        ratio[name] = (a / np.maximum(b, 1e-300)).max(axis=0).tolist()
    out["largest_window_ratio_fast_over_fp64_spread"] = ratio
    return out


def cpu_baseline(sim, case, seconds):
    """Oracle flow velocity of every bound and wake ring (advect.h:16-24) at
    a sample of the wake nodes of the final state, timed on the host cores
    for at least `seconds`; the plain method's timesteps/s on the CPU is
    extrapolated from that pair rate and the run's pair model."""
    import oracle
    from uvlm_amd import workloads as W
    oracle.build()
    wing, g = sim.wing_nodes(), sim.gamma()
    wn, wg = sim.wake()
    targets = wn.reshape(-1, 3)
    n_src = g.size + wg.size
    rng = np.random.default_rng(7)
    nt = 64
    done_pairs, done_t, rounds = 0.0, 0.0, 0
    while done_t < seconds:
        x = targets[rng.integers(0, targets.shape[0], nt)]
        t0 = time.perf_counter()
        oracle.induced_velocity(x, wing, g)
        if wg.size:
            oracle.induced_velocity(x, wn, wg)
        dt = time.perf_counter() - t0
        done_t += dt
        done_pairs += 4.0 * nt * n_src
        rounds += 1
        if dt < 0.5:
            nt *= 2
    rate = done_pairs / done_t
    S = int(case["steps"])
    cpu_run_s = (W.pair_model(case, steps=S) - W.pair_model(case, steps=0)) / rate
    return {"value": S / cpu_run_s, "unit": "timesteps/s", "cores": oracle.num_threads(),
            "kind": "port", "pairs_per_s": rate,
            "sample": (f"oracle sum of all {n_src} bound and wake rings at random wake "
                       f"nodes of the final state ({done_pairs:.3e} pairs in {done_t:.1f} s, "
                       f"{rounds} rounds); timesteps/s of the plain method over steps 1..{S} "
                       f"extrapolated from that rate and the run's pair model "
                       f"({cpu_run_s:.0f} s of CPU time)")}


def bench_tandem(args):
    from uvlm_amd import workloads as W

    case = W.tandem_c5()
    S = int(case["steps"])
    fast_m = dict(FAST)
    if args.fast_theta_min:
        fast_m["ff_theta_min"] = args.fast_theta_min
    if args.no_net:
        fast_m.pop("mlp")
    log("fast form")
    fast = run_case(case, fast_m, args.steps, args.warmup, "fast")
    fast_sim = fast.pop("sim")
    K, first = fast["K"], fast["first"]
    ref_pairs_win = W.pair_model(case, steps=S) - W.pair_model(case, steps=first - 1)
    out = {
        "metric": METRIC, "value": K / fast["t_win"], "unit": "timesteps/s", "n_gpus": 1,
        "steps": K, "warmup": fast["warmup"], "ms_per_step": 1e3 * fast["t_win"] / K,
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"{case['name']}_{S}steps", "wings": 2,
                   "rings_per_wing": [case["rows"], case["cols"]], "advect": case["advect"],
                   "dt": case["dt"], "run_steps": S, "timed_steps": [first, S],
                   "final_wake_rings": 2 * S * case["cols"],
                   "method": {"sums": "near/far split of the wake: lattice-block moment "
                                      "expansions (orders 0-3) for far clusters with a "
                                      "learned per-tile correction, exact near blocks in "
                                      "FP32 with FP64 accumulation and FP64 near-core "
                                      "re-evaluation; bound sheets exact at the wings and "
                                      "in the first 24 wake rows, by each wing's moment "
                                      "expansion beyond",
                              "theta": fast_m["farfield"],
                              "graded": {"reach": fast_m["ff_reach"],
                                         "theta_min": fast_m["ff_theta_min"]},
                              "audit_every": fast_m["audit"],
                              "network": (dict(fast_m["mlp"], arch="32-128-128-128-30 SiLU, "
                                               "BF16 MFMA inference, trained in the run")
                                          if fast_m.get("mlp") else None),
                              "solve": "kept-inverse (FP32 inverse, FP64 residuals), "
                                       "iterative refinement to 1e-14",
                              "loads": "FP64"}},
        "timesteps_per_s_run": S / fast["t_run"],
        "work_rate_reference_counting": {
            "value": ref_pairs_win / fast["t_win"], "unit": "pairs/s",
            "note": ("pairs the reference loop would evaluate in the window "
                     "(4 [N^2 + N Nw + L (N + Nw) + V (N + Nw)] per step) / window time: "
                     "a work-equivalent rate, not a hardware rate")},
        "roofline": roofline("k_ff_near<F32C> (near-block sums of the far-field split)",
                             fast["near"], fast["t_run"], PEAK_FP32_TFLOPS, "k_ff_near"),
    }
    out["roofline"]["note"] = ("averaged over every launch of the fast run (steps 0..2000), "
                               "as rocprofv3 sees it; FP32 segment law, peak = FP32 vector")
    if "ff" in fast:
        err, audits, active = fast["ff"]
        out["farfield"] = {"last_audit_rel_err": err, "audits": audits, "active": bool(active)}
    if "net" in fast:
        out["network"] = fast["net"]
    if "inv" in fast:
        sw, rf, rb, _ = fast["inv"]
        out["inverse_solve"] = {"sweeps": sw, "newton_schulz_refreshes": rf, "rebuilds": rb}
    del fast_sim
    plain_sim = None
    if not args.no_plain:
        log("plain FP64 method")
        plain = run_case(case, PLAIN, args.steps, args.warmup, "plain")
        plain_sim = plain.pop("sim")
        out["vs_plain_fp64"] = {
            "plain_ms_per_step_run": 1e3 * plain["t_run"] / S,
            "fast_ms_per_step_run": 1e3 * fast["t_run"] / S,
            "timesteps_per_s_ratio_run": plain["t_run"] / fast["t_run"],
            "plain_ms_per_step_window": 1e3 * plain["t_win"] / K,
            "timesteps_per_s_ratio_window": plain["t_win"] / fast["t_win"],
            **load_differences(fast["loads"], plain["loads"])}
        out["vs_plain_fp64"]["plain_chain_roofline"] = roofline(
            "k_chain_f64 (plain run)", plain["chain"], plain["t_run"], PEAK_FP64_TFLOPS)
        if args.spread:
            log("second FP64 run (rounding spread)")
            r2 = run_case(case, dict(PLAIN, solve="inv"), args.steps, args.warmup, "fp64-b")
            r2.pop("sim").close()
            out["rounding_spread"] = spread_report(fast, plain, r2)
    if fast_m.get("mlp") and args.net_compare:
        log("fast form without the network")
        m2 = {k: v for k, v in fast_m.items() if k != "mlp"}
        r2 = run_case(case, m2, args.steps, args.warmup, "fast-no-net")
        r2.pop("sim").close()
        cmp = {"ms_per_step_run": 1e3 * r2["t_run"] / S,
               "with_network_ms_per_step_run": 1e3 * fast["t_run"] / S}
        if plain_sim is not None:
            cmp["timesteps_per_s_ratio_run_vs_plain"] = plain["t_run"] / r2["t_run"]
            cmp.update(load_differences(r2["loads"], plain["loads"]))
        out["fast_without_network"] = cmp
    if not args.no_c2:
        log("NACA4412 block")
        c2 = W.naca4412_c2()
        r = run_case(c2, {"precision": "f64", "solve": "inv"}, 1000, 0, "c2")
        sim2 = r.pop("sim")
        S2 = r["S"]
        ref2 = W.pair_model(c2, steps=S2) - W.pair_model(c2, steps=0)
        out["c2_naca4412"] = {
            "workload": f"{c2['name']}_{S2}steps", "ms_per_step": 1e3 * r["t_run"] / S2,
            "timesteps_per_s": S2 / r["t_run"],
            "work_rate_reference_counting_pairs_per_s": ref2 / r["t_run"],
            "roofline": roofline("k_chain_f64", r["chain"], r["t_run"], PEAK_FP64_TFLOPS,
                                 "k_chain_f64")}
        sim2.close()
    if not args.no_cpu:
        log("CPU baseline")
        sim = plain_sim
        if sim is None:
            raise SystemExit("the CPU baseline samples the plain run's final state")
        out["cpu_baseline"] = cpu_baseline(sim, case, args.cpu_seconds)
    if plain_sim is not None:
        plain_sim.close()
    return out


def bench_single(args):
    from uvlm_amd import workloads as W
    case = W.WORKLOADS[args.workload]()
    m = {"precision": args.precision, "solve": args.solve, "farfield": args.farfield,
         "ff_reach": args.ff_reach, "ff_theta_min": args.ff_theta_min, "audit": args.audit,
         "ff_tol": args.ff_tol}
    r = run_case(case, m, args.steps, args.warmup, args.workload)
    sim = r.pop("sim")
    S, K, first = r["S"], r["K"], r["first"]
    peak = PEAK_FP32_TFLOPS if args.precision == "f32c" else PEAK_FP64_TFLOPS
    chain_name = {"f64exact": "k_nbody", "f64": "k_chain_f64", "f32c": "k_chain_f32c"}
    out = {"metric": METRIC, "value": K / r["t_win"], "unit": "timesteps/s", "n_gpus": 1,
           "steps": K, "warmup": r["warmup"], "ms_per_step": 1e3 * r["t_win"] / K,
           "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
           "dtype": {"f64exact": "f64", "f64": "f64", "f32c": "f32+f64"}[args.precision],
           "data": "synthetic",
           "config": {"workload": f"{case['name']}_{S}steps", "timed_steps": [first, S],
                      "precision": args.precision, "solve": args.solve,
                      "farfield_theta": args.farfield or None},
           "timesteps_per_s_run": S / r["t_run"],
           "chain_roofline": roofline(chain_name[args.precision], r["chain"], r["t_run"], peak,
                                      "k_chain_f64" if (args.workload == "c2"
                                                        and args.precision == "f64") else None)}
    if args.farfield > 0:
        out["near_roofline"] = roofline("k_ff_near", r["near"], r["t_run"], peak)
    sim.close()
    return out


def bench_batch(args):
    from uvlm_amd import workloads as W
    from uvlm_amd.gpu import Batch

    S = 300
    cases = W.batch_c4(steps=S)
    K = max(1, min(args.steps, S))
    first = S - K + 1
    b = Batch(cases, max_steps=S, advect="RK2")
    b.run(1, graph=not args.no_graph)  # step 0
    chosen, _ = b.autotune()
    b.sync()
    t0 = time.perf_counter()

    def advance(lo, hi):  # steps lo..hi, in chunks of at most 50
        s = lo
        while s <= hi:
            k = min(50, hi - s + 1)
            b.run(k, graph=not args.no_graph)
            s += k
            log(f"batch step {s - 1}")

    advance(1, first - 1)
    b.sync()
    tw0 = time.perf_counter()
    advance(first, S)
    b.sync()
    t_end = time.perf_counter()
    t_win = t_end - tw0
    ref = sum(W.pair_model(c, steps=S, stages=2) - W.pair_model(c, steps=first - 1, stages=2)
              for c in cases)
    return {"metric": METRIC, "value": len(cases) * K / t_win, "unit": "case-timesteps/s",
            "n_gpus": 1, "steps": K, "warmup": min(args.warmup, first - 1),
            "ms_per_step": 1e3 * t_win / K, "higher_is_better": True, "scaling": "strong",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"batch_256_cases_rk2_{S}steps", "cases": len(cases),
                       "timed_steps": [first, S], "autotuned_variant": chosen,
                       "hipgraph": not args.no_graph},
            "timesteps_per_s_run": S / (t_end - t0),
            "work_rate_reference_counting_pairs_per_s": ref / t_win}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000,
                    help="timed window: the last K steps of the full-length run")
    ap.add_argument("--warmup", type=int, default=0,
                    help="untimed steps just before the window")
    ap.add_argument("--workload", default="c5", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--single", action="store_true",
                    help="c5: run one method alone (as the other workloads) instead of the "
                         "fast-vs-plain comparison")
    ap.add_argument("--precision", default="f64", choices=sorted(PRECISIONS))
    ap.add_argument("--solve", default="inv", choices=sorted(SOLVES))
    ap.add_argument("--farfield", type=float, default=0.0)
    ap.add_argument("--audit", type=int, default=50)
    ap.add_argument("--ff-tol", type=float, default=1e-3)
    ap.add_argument("--ff-reach", type=float, default=0.0)
    ap.add_argument("--ff-theta-min", type=float, default=2.0)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--fast-theta-min", type=float, default=0.0,
                    help="tandem fast form: the graded criterion's floor (default %g)"
                    % FAST["ff_theta_min"])
    ap.add_argument("--no-net", action="store_true",
                    help="tandem fast form without the correction network")
    ap.add_argument("--net-compare", action="store_true",
                    help="also time the fast form without the correction network "
                         "(fast_without_network)")
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--spread", action="store_true",
                    help="c5: also run a second FP64 run that differs from the plain one "
                         "only by rounding, and report both runs' differences per window")
    ap.add_argument("--calibrate-copy", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: the UVLM step of a case runs on a "
              "single MI355X (sharding over GPUs is a design note, not a measured "
              "configuration)", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.workload == "c4":
        out = bench_batch(args)
    elif args.workload == "c5" and not args.single:
        if args.no_plain:
            args.no_cpu = True
        out = bench_tandem(args)
    else:
        out = bench_single(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

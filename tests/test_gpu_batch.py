"""Batched lockstep engine (uvlm_batch_*) vs the oracle run case by case.

Every case of a batch is an independent single-wing SimpleSimulator
(simulator/simple_simulator.cpp:288-313); the batch must reproduce each
case's circulations, wake and CD/CL.  The batch sums with FP64 segment
chains (FMA, refined rsqrt), evaluates the kinematics with the device's
sin/cos and solves with LU (partial pivoting's choices), the oracle with the
reference's order, glibc and full pivoting, so agreement is to 1e-9 relative
(the largest magnitude of each quantity), not bitwise.
"""
import math

import numpy as np
import pytest

from uvlm_amd import UVLMError
from uvlm_amd import _native as N
from uvlm_amd import workloads as W
from uvlm_amd.gpu import Batch

pytestmark = pytest.mark.gpu
TOL = 1e-9


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale


def mixed_cases(advect):
    """Uneven lattices (1 to 5 LU panels, up to 3 update tiles per side),
    different motions, time steps and attack angles."""
    base = [
        dict(rows=3, cols=6, span=4.0, plug=(0.1, 1.0, 0.0), flap=(0, 0, 0), dt=0.125,
             alpha=W.ALPHA5, generator="rect", digit=0),
        dict(rows=6, cols=12, span=6.0, plug=(0, 0, 0), flap=(math.pi / 6, 0.6, 0.0),
             dt=2 * math.pi / 0.6 / 40, alpha=W.ALPHA5, generator="naca4", digit=44),
        dict(rows=2, cols=4, span=2.0, plug=(0.2, 2.0, 0.3), flap=(0.1, 2.0, 1.0), dt=0.1,
             alpha=0.0, generator="rect", digit=0),
        dict(rows=10, cols=14, span=5.0, plug=(0.05, 0.8, 0.0), flap=(0, 0, 0), dt=0.15,
             alpha=3 * math.pi / 180, generator="naca4", digit=83),
    ]
    cases = []
    for k, b in enumerate(base):
        w = W._wing(generator=b["generator"], digit=b["digit"], span=b["span"],
                    plug=b["plug"], flap=b["flap"], alpha=b["alpha"])
        cases.append({"name": f"batch_mixed_{k}", "wings": [w], "rows": b["rows"],
                      "cols": b["cols"], "forward_flight": (-1.0, 0.0, 0.0), "dt": b["dt"],
                      "advect": advect, "ignore_steps": 2, "steps": 0})
    return cases


def oracle_run(orc, case, steps):
    o = orc.Sim(case)
    loads = []
    for s in range(steps):
        o.step(s)
        loads.append(o.loads()[0].copy())
    return o, np.array(loads)


def check_case(b, k, o, o_loads):
    name = b.cases[k]["name"]
    assert rel_err(b.gamma(k), o.gamma()[0]) < TOL, name
    assert rel_err(b.wing_nodes(k), o.wing_nodes()[0]) < TOL, name
    gn, gg = b.wake(k)
    on, og = o.wake()
    assert gn.shape == on[0].shape and gg.shape == og[0].shape, name
    assert rel_err(gn, on[0]) < TOL, name
    assert rel_err(gg, og[0]) < TOL, name
    gl = b.loads(k)
    assert gl.shape == o_loads.shape, name
    assert rel_err(gl[:, 0], o_loads[:, 0]) < TOL, name  # CD (CT = -CD)
    assert rel_err(gl[:, 1], o_loads[:, 1]) < TOL, name  # CL


@pytest.mark.parametrize("advect", ["euler", "RK2", "RK4"])
def test_batch_matches_oracle_per_case(orc, advect):
    steps = 12
    cases = mixed_cases(advect)
    print(f"[batch {advect}] {len(cases)} cases, {steps} steps", flush=True)
    b = Batch(cases, max_steps=steps, advect=advect, ignore_steps=2)
    b.run(steps, graph=True)
    b.sync()
    assert b.steps == steps
    for k, case in enumerate(cases):
        o, ol = oracle_run(orc, case, steps)
        check_case(b, k, o, ol)


def test_graph_replay_equals_eager(orc):
    """The captured step replays the eager launch sequence exactly."""
    cases = mixed_cases("RK2")
    a = Batch(cases, max_steps=10, advect="RK2", ignore_steps=2)
    g = Batch(cases, max_steps=10, advect="RK2", ignore_steps=2)
    for chunk in (1, 3, 4):  # eager work and syncs between replays
        a.run(chunk, graph=False)
        g.run(chunk, graph=True)
        a.sync()
        g.sync()
        for k in range(len(cases)):
            np.testing.assert_array_equal(a.gamma(k), g.gamma(k))
            np.testing.assert_array_equal(a.wake(k)[0], g.wake(k)[0])
            np.testing.assert_array_equal(a.loads(k), g.loads(k))


def test_autotune_keeps_results(orc):
    """Every launch shape computes the same sums: a batch retuned midway
    agrees with the oracle."""
    cases = mixed_cases("euler")
    b = Batch(cases, max_steps=10, advect="euler")
    b.run(4)
    chosen, times = b.autotune()
    assert 0 <= chosen < 4 and np.all(times[:4] > 0)
    b.run(6)
    b.sync()
    for k, case in enumerate(cases):
        check_case(b, k, *oracle_run(orc, case, 10))


def test_batch_c4_sample_first_steps(orc):
    """A slice of the 256-case workload (its first 6 cases, real sizes)."""
    cases = W.batch_c4(cases=6, steps=5)
    b = Batch(cases, max_steps=5, advect="RK2")
    b.run(5)
    b.sync()
    for k, case in enumerate(cases):
        check_case(b, k, *oracle_run(orc, case, 5))


def test_capacity_and_arguments(ctx):
    cases = mixed_cases("euler")[:1]
    b = Batch(cases, max_steps=2, advect="euler", ctx=ctx)
    b.run(3)
    with pytest.raises(UVLMError) as e:
        b.run(1)
    assert e.value.code == N.ERR_CAPACITY
    bad = mixed_cases("euler")[:1]
    bad[0]["cols"] = 5
    with pytest.raises(UVLMError):
        Batch(bad, max_steps=2, ctx=ctx)
    with pytest.raises(UVLMError) as e:
        Batch(cases, max_steps=2, advect="AB2", ctx=ctx)
    assert e.value.code == N.ERR_ARG


def test_pivoted_fallback_agrees(orc):
    """The partially pivoted fallback of the batched solve (normally taken
    only by cases whose pivoting checks fail) gives the same runs."""
    cases = mixed_cases("RK2")
    b = Batch(cases, max_steps=8, advect="RK2", ignore_steps=2)
    b.force_pivoted(True)
    b.run(8)
    b.sync()
    for k, case in enumerate(cases):
        check_case(b, k, *oracle_run(orc, case, 8))


# --- the 256-case workload at its full size ---------------------------------
C4_STEPS = 300


@pytest.fixture(scope="module")
def c4_full():
    cases = W.batch_c4(steps=C4_STEPS)
    print(f"[batch c4] {len(cases)} cases, {C4_STEPS} steps, RK2, graph", flush=True)
    b = Batch(cases, max_steps=C4_STEPS, advect="RK2")
    b.run(C4_STEPS + 1, graph=True)  # steps 0..300
    b.sync()
    yield cases, b


def c4_sample(cases, k=8):
    """k cases spread over the lattice sizes (smallest to largest N)."""
    order = sorted(range(len(cases)), key=lambda i: (cases[i]["rows"] * cases[i]["cols"], i))
    return [order[round(q * (len(order) - 1) / (k - 1))] for q in range(k)]


def test_c4_full_invariants(c4_full):
    """Every case of the 256: 301 load rows, all finite; 300 wake circulation
    rows; the trailing-edge wake row moved at -ff from the final trailing
    edge (advect.h:41-45; RK2 uses Euler there), bitwise."""
    cases, b = c4_full
    assert b.steps == C4_STEPS + 1
    for k, c in enumerate(cases):
        loads = b.loads(k)
        assert loads.shape == (C4_STEPS + 1, 2) and np.all(np.isfinite(loads)), k
        nodes, gam = b.wake(k)
        assert nodes.shape == (C4_STEPS + 1, c["cols"] + 1, 3), k
        assert gam.shape == (C4_STEPS, c["cols"]), k
        assert np.all(np.isfinite(nodes)) and np.all(np.isfinite(gam)), k
        te = b.wing_nodes(k)[-1]
        ff = np.asarray(c["forward_flight"], dtype=np.float64)
        np.testing.assert_array_equal(nodes[0], te + (-ff) * c["dt"])


def test_c4_full_early_window_vs_oracle(orc, c4_full):
    """8 cases of different sizes: CD and CL of steps 0..40 of the full batch
    agree with the oracle to 1e-9 (before the roll-up amplifies rounding:
    two oracle runs that differ only in pivoting part by up to 1e-5 in
    the loads of some cases by step 300)."""
    cases, b = c4_full
    early = 40
    for k in c4_sample(cases):
        c = cases[k]
        print(f"[batch c4] oracle case {k}: {c['rows']}x{c['cols']}", flush=True)
        _, ol = oracle_run(orc, c, early + 1)
        gl = b.loads(k)[: early + 1]
        assert rel_err(gl[:, 0], ol[:, 0]) < TOL, k
        assert rel_err(gl[:, 1], ol[:, 1]) < TOL, k


def test_c4_full_run_of_a_stable_case_vs_oracle(orc, c4_full):
    """Case 245 (5 x 8 rings) is one whose oracle runs with full and with
    partial pivoting stay within 1e-14 of each other for all 300 steps: its
    whole state at step 300 agrees with the oracle to 1e-9."""
    cases, b = c4_full
    k = 245
    o, ol = oracle_run(orc, cases[k], C4_STEPS + 1)
    check_case(b, k, o, ol)

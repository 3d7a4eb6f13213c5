"""Near/far split of the wake sum (uvlm_farfield_velocity, the simulation's
far-field advection) against the oracle's direct sums.

Pins: (1) with theta so large that every block is near, the split is the
exact sum over a partition of the rings (the blocks' sub-lattices are closed
loops) -- agreement to rounding; (2) the moment expansion of one block
converges at its order: the relative error falls like (r/|d|)^3 (first
omitted moment) when the distance doubles; (3) on a rolled-up tandem wake the velocity error stays within a
bound that falls with theta; (4) simulations with the split and audits.
"""
import math

import numpy as np
import pytest

from uvlm_amd import Simulation
from uvlm_amd import workloads as W

pytestmark = pytest.mark.gpu


def wavy_lattice(rows, cols, seed=0, num=1):
    rng = np.random.default_rng(seed)
    i, j = np.meshgrid(np.arange(rows + 1), np.arange(cols + 1), indexing="ij")
    sheets = []
    for n in range(num):
        nodes = np.stack([0.3 * i + 0.05 * np.sin(0.4 * j) + 2.0 * n,
                          0.1 * j - 0.05 * cols, 0.2 * np.sin(0.3 * i + 0.2 * j)], -1)
        sheets.append(nodes + rng.normal(scale=0.003, size=nodes.shape))
    gamma = 0.05 + 0.02 * rng.normal(size=(num, rows, cols))
    return np.stack(sheets), gamma


def morton_keys(x):
    """30-bit Morton keys of points in their bounding box, 10 bits per axis
    (x in bit 0, y in bit 1, z in bit 2 of each triple), IEEE FP64 like the
    device's key kernel."""
    lo, hi = x.min(0), x.max(0)
    w = hi - lo
    keys = np.zeros(x.shape[0], dtype=np.uint32)
    for d in range(3):
        t = (x[:, d] - lo[d]) / w[d] if w[d] > 0 else np.zeros(x.shape[0])
        c = np.floor(np.clip(t, 0.0, 1.0) * 1023.0).astype(np.uint32)
        for b in range(10):
            keys |= ((c >> np.uint32(b)) & np.uint32(1)) << np.uint32(3 * b + d)
    return keys


@pytest.mark.parametrize("n", [1, 63, 64, 1025, 9000, 16384, 16385, 300000, 1200000])
def test_target_order_is_a_stable_sort_of_the_morton_keys(ctx, n):
    """The split's tile order comes from its own LSD radix sort: the keys are
    the Morton keys of a direct restatement, the order is numpy's stable
    argsort of them (equal keys in index order), padded with -1; sizes
    cover the one-workgroup sort (up to 16384 keys, partial and full
    wavefronts) and the multi-workgroup one (one partial chunk, several
    rows, rows of several chunks)."""
    rng = np.random.default_rng(n)
    x = rng.normal(size=(n, 3)) * np.array([20.0, 3.0, 0.5])
    if n > 100:
        x[::5] = x[7]                              # repeated points: equal keys
        x[1::11, 2] = np.round(x[1::11, 2], 1)     # shared high digits
    keys, order = ctx.farfield_target_order(x)
    np.testing.assert_array_equal(keys, morton_keys(x))
    np.testing.assert_array_equal(order[:n], np.argsort(keys, kind="stable"))
    assert (order[n:] == -1).all() and order.size == (n + 63) // 64 * 64


def test_all_near_is_the_exact_sum(ctx, orc):
    nodes, gamma = wavy_lattice(37, 40, seed=1, num=2)
    rng = np.random.default_rng(2)
    x = np.concatenate([nodes.reshape(-1, 3)[::7], rng.normal(size=(200, 3)) * 3])
    ref = orc.induced_velocity(x, nodes, gamma)
    got = ctx.farfield_velocity(x, nodes, gamma, theta=1e12)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def test_expansion_converges_at_its_order(ctx, orc):
    """One level-0 block (4 x 16 rings) seen from growing distances: every
    target beyond theta = 2 uses the expansion, whose relative error falls
    ~8x per doubling of the distance (first omitted term ~ (r/|d|)^3)."""
    nodes, gamma = wavy_lattice(4, 16, seed=3)
    c = nodes.reshape(-1, 3).mean(0)
    rad = np.linalg.norm(nodes.reshape(-1, 3) - c, axis=1).max()
    rng = np.random.default_rng(4)
    dirs = rng.normal(size=(64, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    errs = []
    for ratio in (4.0, 8.0, 16.0):
        x = c + ratio * rad * dirs
        ref = orc.induced_velocity(x, nodes, gamma)
        # one target per call: a tile's own radius must not hide the distance
        got = np.concatenate([ctx.farfield_velocity(x[k:k + 1], nodes, gamma, theta=2.0)
                              for k in range(len(x))])
        errs.append(np.max(np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)))
    assert errs[0] < 2e-2
    assert errs[1] < errs[0] / 6 and errs[2] < errs[1] / 6


@pytest.mark.parametrize("theta,bound", [(3.0, 3e-4), (5.0, 5e-5)])
def test_rolled_up_tandem_wake(ctx, orc, theta, bound):
    """Wake nodes of a tandem state (2 sheets, 60 rows x 32 columns after a
    flapping run on the oracle): error of the split velocity relative to the
    largest induced velocity."""
    case = W.tandem_c5(steps=60, rows=4, cols=32)
    sim = orc.Sim(case)
    for s in range(61):
        sim.step(s)
    nodes, gamma = sim.wake()
    x = nodes.reshape(-1, 3)
    ref = orc.induced_velocity(x, nodes, gamma)
    got = ctx.farfield_velocity(x, nodes, gamma, theta=theta)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"theta {theta}: max error {err:.2e} of the largest induced velocity")
    assert err < bound


def test_simulation_with_farfield_and_audits(orc):
    """The tandem workload (reduced lattice) with the split: audits run and
    report the stage difference, CL/CT follow the exact run."""
    steps = 80
    case = W.tandem_c5(steps=steps, rows=4, cols=16)
    a = Simulation(case, steps)
    b = Simulation(case, steps)
    b.set_farfield(theta=4.0, audit_every=20, tol=1e-3)
    for s in range(steps + 1):
        a.step(s)
        b.step(s)
    a.sync()
    b.sync()
    err, audits, active = b.farfield_stats()
    # steps 1..80 advect a wake (step 0 has none): audits at 1, 21, 41, 61
    assert audits == 4 and active and 0 <= err < 1e-3
    la, lb = a.loads(), b.loads()
    scale = np.abs(la).max(axis=0)
    assert np.all(np.abs(la - lb).max(axis=0) / scale < 1e-3)


def test_audit_switches_the_split_off(orc):
    """A tolerance no expansion can meet: the first audit switches the split
    off and the run continues exactly."""
    steps = 30
    case = W.tandem_c5(steps=steps, rows=4, cols=16)
    b = Simulation(case, steps)
    b.set_farfield(theta=2.0, audit_every=10, tol=1e-15)
    for s in range(steps + 1):
        b.step(s)
    b.sync()
    err, audits, active = b.farfield_stats()
    assert not active and audits >= 1 and err > 1e-15


def test_strided_audit_of_a_large_wake(orc):
    """Wakes of 32768 nodes or more are audited on every 8th node: the audit
    still reports the stage difference, and a tolerance no expansion can
    meet still switches the split off."""
    case = W.tandem_c5(steps=273)  # 2 x 16 x 64: 130 wake nodes per step
    a = Simulation(case, 273)
    a.set_farfield(theta=4.0, audit_every=30, tol=1e-3)
    b = Simulation(case, 273)
    b.set_farfield(theta=4.0, audit_every=1000, tol=1e-3)
    for s in range(272):
        print(f"[strided audit] step {s}", flush=True) if s % 50 == 0 else None
        a.step(s)
        b.step(s)
    a.sync()
    err, audits, active = a.farfield_stats()
    # audits at steps 1, 31, ..., 271; the last one's wake has 2 * 272 * 65 nodes
    assert audits == 10 and active and 0 < err < 1e-3
    b.set_farfield(theta=2.0, audit_every=10, tol=1e-15)
    prev, _ = b.wake()  # node rows 1.. of the state the next step audits
    b.step(272)  # audited: 2 * 273 * 65 wake nodes, every 8th summed exactly
    b.sync()
    err, audits, active = b.farfield_stats()
    assert audits == 2 and not active and err > 1e-15
    # The reported difference is the strided maximum over that state:
    # nodes off + 8k (off = audits before this one, mod 8), trailing-edge
    # rows skipped, against the oracle's exact stage velocity there; the
    # split's velocities are the Euler update's (post - pre) / dt.
    R, C1, dt = case["rows"], case["cols"] + 1, case["dt"]
    ff = np.array(case["forward_flight"])
    wing, gam = b.wing_nodes(), b.gamma()
    post, wgam = b.wake()
    te = wing[:, R]
    np.testing.assert_array_equal(post[:, 0], te + dt * -ff)
    pre = np.concatenate([te[:, None], prev], axis=1)
    vel = ((post - pre) / dt).reshape(-1, 3)
    V, per = pre.shape[0] * pre.shape[1] * C1, pre.shape[1] * C1
    idx = np.arange(1 % 8, V, 8)
    idx = idx[idx % per >= C1]
    x = pre.reshape(-1, 3)[idx]
    exact = -ff + orc.induced_velocity(x, wing, gam) + orc.induced_velocity(x, pre, wgam)
    want = np.abs(vel[idx] - exact).max() / np.abs(exact).max()
    assert 1e-9 < want < 1e-3
    np.testing.assert_allclose(err, want, rtol=1e-6)


def test_simulation_learns_the_correction(orc):
    """In-run learning on the tandem workload (reduced lattice): the
    network trained on the run's own wake corrects the far pairs of the
    following steps; loads stay within the split's tolerance of the exact
    run and the audits pass."""
    from oracle import mlp as M
    steps = 80
    case = W.tandem_c5(steps=steps, rows=4, cols=32)
    a = Simulation(case, steps)
    b = Simulation(case, steps)
    b.set_farfield(theta=2.5, audit_every=10, tol=1e-3)
    w = M.init(5)
    for s in range(steps + 1):
        a.step(s)
        b.step(s)
        if s in (50, 65):
            w, loss = b.farfield_learn(w, theta=2.5, max_samples=20000, iters=500, seed=s)
            assert np.isfinite(loss)
    a.sync()
    b.sync()
    pairs, trips = b.farfield_mlp_stats()
    err, audits, active = b.farfield_stats()
    assert pairs > 0 and active and err < 1e-3
    la, lb = a.loads(), b.loads()
    scale = np.abs(la).max(axis=0)
    assert np.all(np.abs(la - lb).max(axis=0) / scale < 1e-3)


def test_full_size_fast_form_with_the_network():
    """The benchmark's fast form at the tandem workload's full size (2 x
    16 x 64 rings): split with the graded criterion, F32C near sums,
    kept-inverse solve, the correction network trained in the run at step
    100 (same settings as the benchmark): CL/CT within 1e-4 of the plain
    FP64 run over steps 0..30 (later the wake's roll-up amplifies any
    rounding difference, so the 160-step run means are held to 1e-3), the
    audits pass and the network corrects far pairs without tripping."""
    from uvlm_amd import _native as N
    from uvlm_amd.gpu import mlp_init
    steps = 160
    base = W.tandem_c5(steps=steps)
    plain = Simulation(dict(base, precision=N.PREC_F64, solve_mode=N.SOLVE_F64), steps)
    fast = Simulation(dict(base, precision=N.PREC_F32C, solve_mode=N.SOLVE_F64_INV), steps)
    fast.set_farfield(theta=4.0, audit_every=50, tol=1e-3)
    fast.set_farfield_focus(reach=1.5, theta_min=1.2)
    w = mlp_init(5)
    for s in range(steps + 1):
        if s % 40 == 0:
            print(f"[full-size network] step {s}", flush=True)
        plain.step(s)
        fast.step(s)
        if s == 100:
            w, loss = fast.farfield_learn(w, 4.0, max_samples=12000, iters=80, batch=4096,
                                          lr=2e-3, seed=s)
            assert np.isfinite(loss)
    plain.sync()
    fast.sync()
    pairs, trips = fast.farfield_mlp_stats()
    err, audits, active = fast.farfield_stats()
    assert pairs > 0 and trips == 0
    assert active and audits == 4 and err < 1e-3
    lp, lf = plain.loads(), fast.loads()
    assert np.isfinite(lf).all() and lf.shape == lp.shape == (steps + 1, 2, 2)
    scale = np.abs(lp).max(axis=0)
    assert (np.abs(lf[:31] - lp[:31]).max(axis=0) / scale).max() <= 1e-4
    assert (np.abs(lf.mean(axis=0) - lp.mean(axis=0)) / scale).max() <= 1e-3


def test_graded_criterion_away_from_the_wings(orc):
    """The graded criterion (theta falls towards theta_min beyond reach wing
    radii): in effect on the wake away from the wings, loads within the
    split's tolerance of the exact run, the weighted audits pass; a reach
    beyond the whole wake gives the plain split bitwise; theta_min <= 1 is
    refused."""
    steps = 80
    case = W.tandem_c5(steps=steps, rows=4, cols=16)
    exact = Simulation(case, steps)
    plain, graded, wide = (Simulation(case, steps) for _ in range(3))
    for sim in (plain, graded, wide):
        sim.set_farfield(theta=4.0, audit_every=20, tol=1e-3)
    graded.set_farfield_focus(reach=0.5, theta_min=1.5)
    wide.set_farfield_focus(reach=1e6, theta_min=1.5)
    with pytest.raises(Exception):
        graded.set_farfield_focus(reach=1.0, theta_min=1.0)
    sims = (exact, plain, graded, wide)
    for s in range(steps + 1):
        for sim in sims:
            sim.step(s)
    for sim in sims:
        sim.sync()
    le, lp, lg, lw = (sim.loads() for sim in sims)
    np.testing.assert_array_equal(lw, lp)
    assert np.abs(lg - lp).max() > 0.0
    scale = np.abs(le).max(axis=0)
    assert np.all(np.abs(lg - le).max(axis=0) / scale < 1e-3)
    err, audits, active = graded.farfield_stats()
    assert audits == 4 and active and 0 <= err < 1e-3


def test_f32c_near_sums_in_the_split(orc):
    """The split with F32C near sums (the near blocks' records walked in FP32
    from FP64 differences, FP64 re-evaluation near the cut-off): loads
    within 1e-4 of the same split in FP64 over steps before the tandem wake
    amplifies rounding, within the split's tolerance of the exact run, and
    the audits pass."""
    from uvlm_amd import _native as N
    steps = 30
    case = W.tandem_c5(steps=steps, rows=4, cols=16)
    exact = Simulation(case, steps)
    f64 = Simulation(case, steps)
    c32 = dict(case)
    c32["precision"] = N.PREC_F32C
    f32 = Simulation(c32, steps)
    for sim in (f64, f32):
        sim.set_farfield(theta=3.0, audit_every=10, tol=1e-3)
    sims = (exact, f64, f32)
    for s in range(steps + 1):
        for sim in sims:
            sim.step(s)
    for sim in sims:
        sim.sync()
    le, l64, l32 = (sim.loads() for sim in sims)
    scale = np.abs(le).max(axis=0)
    d32 = np.abs(l32 - l64).max(axis=0) / scale
    print(f"F32C vs FP64 split: {d32.ravel()}")
    assert np.all(d32 < 1e-4) and np.abs(l32 - l64).max() > 0.0
    assert np.all(np.abs(l32 - le).max(axis=0) / scale < 1e-3)
    err, audits, active = f32.farfield_stats()
    assert audits == 3 and active and 0 <= err < 1e-3

"""Invariants of the oracle's SimpleSimulator loop (simple_simulator.cpp
:171-313, advect.cpp) that pin it where the reference has no test."""
import math
import os

import numpy as np
import pytest

from uvlm_amd import workloads as W


def small_plate(steps=12, advect="euler", **kw):
    c = W.plate_c1(steps=steps, rows=3, cols=6)
    c["advect"] = advect
    c.update(kw)
    return c


def test_shedding_and_trailing_edge_row(orc):
    """Shed (:171-188): at step s the wake has s+1 node rows and s Gamma
    rows; Gamma row 0 is the TE row of the previous solve; after Euler
    advection node row 0 is the TE row moved by -ff*dt (advect.h:41-45)."""
    case = small_plate()
    sim = orc.Sim(case)
    ff = np.array(case["forward_flight"])
    prev_te = None
    for s in range(6):
        sim.step(s)
        nodes, gam = sim.wake()
        assert nodes.shape[1] == s + 1 and gam.shape[1] == s
        te = sim.wing_nodes()[:, -1]
        np.testing.assert_array_equal(nodes[:, 0], te + (-ff) * case["dt"])
        if s > 0:
            np.testing.assert_array_equal(gam[:, 0], prev_te)
        prev_te = sim.gamma()[:, -1].copy()


def test_no_penetration(orc):
    """A Gamma = rhs: the normal velocity at every collocation point vanishes
    (simple_simulator.cpp:103-124)."""
    case = small_plate(steps=8)
    sim = orc.Sim(case)
    ff = np.array(case["forward_flight"])
    for s in range(5):
        sim.step(s)
    wake_nodes, wake_gam = sim.wake()
    g_prev = sim.gamma()
    s = 5
    sim.step(s)
    wing = sim.wing_nodes()
    # rebuild the pre-advection wake of step s: prepend TE nodes / TE Gamma
    pre_nodes = np.concatenate([wing[:, -1:], wake_nodes], axis=1)
    pre_gam = np.concatenate([g_prev[:, -1:], wake_gam], axis=1)
    cp, nrm, _ = orc.lattice_geometry(wing)
    init = orc.build_wing(orc.GEN_RECT, 0, 1.0, 4.0, 3, 6)
    cp0, _, _ = orc.lattice_geometry(init[None])
    g = sim.gamma()
    ub = orc.induced_velocity(cp, wing, g)
    uw = orc.induced_velocity(cp, pre_nodes, pre_gam)
    vm = np.array([orc.morph_velocity(case["wings"][0], x, s * case["dt"]) for x in cp0])
    vn = np.einsum("kd,kd->k", ub + uw - vm - ff, nrm)
    assert np.max(np.abs(vn)) < 1e-12


def test_flat_plate_symmetry(orc):
    case = small_plate(steps=10)
    sim = orc.Sim(case)
    for s in range(11):
        sim.step(s)
    g = sim.gamma()
    np.testing.assert_allclose(g, g[:, :, ::-1], rtol=1e-11, atol=1e-14)
    nodes, _ = sim.wake()
    mirrored = nodes[:, :, ::-1] * np.array([1, -1, 1])
    np.testing.assert_allclose(nodes, mirrored, rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("advect", ["euler", "RK2", "RK4", "AB2", "RK4classic"])
def test_zero_circulation_translates_wake(orc, advect):
    """alpha = 0 and no motion: rhs = ff.n = 0, Gamma = 0, every scheme
    reduces to translation by -ff*dt per step."""
    case = small_plate(steps=8, advect=advect)
    case["wings"][0]["alpha"] = 0.0
    case["wings"][0]["plug"] = (0.0, 0.0, 0.0)
    sim = orc.Sim(case)
    for s in range(8):
        sim.step(s)
    assert np.all(sim.gamma() == 0)
    nodes, gam = sim.wake()
    assert np.all(gam == 0)
    te = sim.wing_nodes()[:, -1]
    dt = case["dt"]
    for i in range(nodes.shape[1]):
        np.testing.assert_allclose(nodes[:, i], te + np.array([1.0, 0, 0]) * dt * (i + 1),
                                   rtol=0, atol=1e-14)


def test_rk_schemes_use_euler_near_the_trailing_edge(orc):
    """RK2/RK4 with ignore_steps >= rows reduce to Euler (advect.cpp:74-78,
    170-177)."""
    ref = orc.Sim(small_plate(steps=6, advect="euler"))
    for advect in ("RK2", "RK4"):
        sim = orc.Sim(small_plate(steps=6, advect=advect, ignore_steps=100))
        r2 = orc.Sim(small_plate(steps=6, advect="euler"))
        for s in range(6):
            sim.step(s)
            r2.step(s)
        np.testing.assert_array_equal(sim.wake()[0], r2.wake()[0])
    del ref


def test_rk4_reference_formula_differs_from_classic(orc):
    a = orc.Sim(small_plate(steps=12, advect="RK4", ignore_steps=1))
    b = orc.Sim(small_plate(steps=12, advect="RK4classic", ignore_steps=1))
    for s in range(12):
        a.step(s)
        b.step(s)
    da = a.wake()[0]
    db = b.wake()[0]
    diff = np.max(np.abs(da - db))
    assert 0 < diff < 1e-2


def test_steady_lift_matches_kutta_joukowski(orc):
    """No motion: CL from the Joukowski loads approaches the Kutta-Joukowski
    lift of the trailing-edge strip circulations, rho U sum_j Gamma_TE,j dy_j
    (the steady formula of steady_problem_main.cpp:197-209)."""
    case = W.plate_c1(steps=60, rows=4, cols=8)
    case["wings"][0]["plug"] = (0.0, 0.0, 0.0)
    sim = orc.Sim(case)
    for s in range(61):
        sim.step(s)
    cd, cl = sim.loads()[0]
    g = sim.gamma()[0]
    nodes = sim.wing_nodes()[0]
    dy = np.diff(nodes[-1, :, 1])
    L = 1.0 * 1.0 * np.sum(g[-1] * dy)
    cl_kj = L / (0.5 * 1.0 * 4.0)
    assert 0.2 < cl < 0.45
    assert abs(cl - cl_kj * math.cos(5 * math.pi / 180)) / cl_kj < 0.05


def test_erase_wake_after(orc):
    case = small_plate(steps=10, erase_wake_after=4)
    sim = orc.Sim(case)
    for s in range(8):
        sim.step(s)
    nodes, gam = sim.wake()
    # rows grow by one per step and shrink by one from step 4 on
    assert nodes.shape[1] == 4 and gam.shape[1] == 3


def test_pair_count_model(orc):
    case = small_plate(steps=5, advect="RK2")
    sim = orc.Sim(case)
    for s in range(6):
        sim.step(s)
    assert sim.pairs == pytest.approx(W.pair_model(case, steps=5, stages=2))


def test_oracle_wake_sensitivity(orc):
    """The plate workload's wake roll-up amplifies rounding: the oracle run
    with partial instead of full pivoting (solutions equal to ~1e-15 each
    step) stays within 1e-13 of the full-pivot history up to step 150 and
    departs by more than 1e-10 by step 200.  Long-run parity at 1e-10 is
    therefore a bitwise property (same operation order), not a tolerance."""
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "plate_c1_200.npz"))
    snaps = list(gold["snap_steps"])
    sim = orc.Sim(W.plate_c1(steps=200), lu=orc.LU_PARTIALPIV)
    errs = {}
    for s in range(201):
        sim.step(s)
        if s in snaps:
            ref = gold["gammas"][snaps.index(s)]
            errs[s] = np.abs(sim.gamma() - ref).max() / np.abs(ref).max()
    assert max(errs[s] for s in snaps if s <= 150) < 1e-13
    assert errs[200] > 1e-10

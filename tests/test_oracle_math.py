"""Pins of the oracle by what the mathematics fixes (no reference build
exists: Eigen, glog, gflags and the multiple_sheet submodule are absent)."""
import math

import numpy as np
import pytest


def quad_segment(a, b, x, n=400):
    """Biot-Savart integral (1/4pi) int dl x r / |r|^3 of the unit-strength
    segment a->b, Gauss-Legendre in long double."""
    s, w = np.polynomial.legendre.leggauss(n)
    s = (s + 1) / 2
    w = w / 2
    a, b, x = (np.asarray(v, dtype=np.longdouble) for v in (a, b, x))
    r0 = b - a
    u = np.zeros(3, dtype=np.longdouble)
    for si, wi in zip(s, w):
        r = x - (a + si * r0)
        u += wi * np.cross(r0, r) / np.linalg.norm(r) ** 3
    return np.asarray(u / (4 * np.pi), dtype=np.float64)


def test_segment_matches_quadrature(orc):
    rng = np.random.default_rng(1)
    for _ in range(40):
        a, b = rng.normal(size=3), rng.normal(size=3)
        x = rng.normal(size=3) * 2
        # keep the target away from the segment so the quadrature converges
        t = np.clip(np.dot(x - a, b - a) / np.dot(b - a, b - a), 0, 1)
        if np.linalg.norm(x - (a + t * (b - a))) < 0.3:
            continue
        ref = quad_segment(a, b, x)
        got = orc.biot_savart(a, b, x)
        np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-12)


def test_ring_far_field_is_a_dipole(orc):
    """A closed ring of strength G, area A, normal n (right-hand rule of the
    node order p0->p1->p2->p3) looks like a dipole m = G A n far away."""
    nodes = np.array([[[0, 0, 0], [0, 1.0, 0]], [[0.5, 0, 0], [0.5, 1.0, 0]]])  # 1x1 lattice
    G = 1.7
    p0, p1, p2 = nodes[0, 0], nodes[0, 1], nodes[1, 1]
    n = np.cross(p1 - p0, p2 - p1)
    n /= np.linalg.norm(n)
    m = G * 0.5 * n
    c = nodes.reshape(-1, 3).mean(axis=0)
    for d in (50.0, 200.0):
        for dirv in ([1, 0, 0], [0.3, 0.5, 0.8], [0, 0, 1]):
            rhat = np.asarray(dirv, float) / np.linalg.norm(dirv)
            x = c + d * rhat
            u = orc.induced_velocity([x], nodes[None], np.array([[[G]]]))[0]
            dip = (3 * np.dot(m, rhat) * rhat - m) / (4 * math.pi * d ** 3)
            np.testing.assert_allclose(u, dip, rtol=5 / d ** 2 * 10, atol=1e-9 / d ** 3)


def test_linearity_and_superposition(orc):
    rng = np.random.default_rng(2)
    nodes = rng.normal(size=(2, 3, 4, 3))
    g = rng.normal(size=(2, 2, 3))
    x = rng.normal(size=(7, 3)) * 3
    u = orc.induced_velocity(x, nodes, g)
    # scaling by a power of two is exact
    np.testing.assert_array_equal(orc.induced_velocity(x, nodes, 2 * g), 2 * u)
    # superposition over sheets (different association -> rounding only)
    u0 = orc.induced_velocity(x, nodes[:1], g[:1])
    u1 = orc.induced_velocity(x, nodes[1:], g[1:])
    np.testing.assert_allclose(u0 + u1, u, rtol=1e-13, atol=1e-15)


def test_voring_per_segment_vs_legacy_ring(orc):
    """VORING (p0->p1->p2->p3, x Gamma per segment) and the legacy ring
    (node[(k+1)%4] -> node[k]) circulate in opposite directions."""
    rng = np.random.default_rng(3)
    nodes = rng.normal(size=(1, 2, 2, 3))
    ring = np.array([nodes[0, 0, 0], nodes[0, 0, 1], nodes[0, 1, 1], nodes[0, 1, 0]])
    x = rng.normal(size=(5, 3)) * 2
    a = orc.induced_velocity(x, nodes, np.array([[[1.3]]]))
    b = orc.ring_induced_velocity(x, ring[None], [1.3])
    np.testing.assert_allclose(a, -b, rtol=1e-12, atol=1e-15)


def cutoff_threshold():
    """Smallest x >= 0 with fl(fl(sqrt(x))^2) >= 1e-10 (bisection on bits)."""
    lo = np.float64(0).view(np.uint64)
    hi = np.float64(1e-9).view(np.uint64)
    while hi - lo > 1:
        mid = lo + (hi - lo) // 2
        x = mid.view(np.float64)
        s = np.sqrt(x)
        if s * s < 1e-10:
            lo = mid
        else:
            hi = mid
    return hi.view(np.float64)


def test_cutoff_threshold_equivalence():
    """The GPU replaces `sqrt(dsq)^2 < 1e-10` (vortex.cpp:31) by `dsq < T`;
    both agree on every double near T."""
    T = cutoff_threshold()
    bits = np.arange(-20000, 20001, dtype=np.int64) + np.int64(np.float64(T).view(np.int64))
    x = bits.view(np.float64)
    s = np.sqrt(x)
    ref = (s * s) < 1e-10
    np.testing.assert_array_equal(ref, x < T)
    # and far from it
    for v in (0.0, 1e-30, 5e-11, 2e-10, 1.0):
        s = np.sqrt(v)
        assert ((s * s) < 1e-10) == (v < T)


@pytest.mark.parametrize("n", [1, 2, 7, 33, 128])
def test_fullpiv_lu_solves(orc, n):
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n)) + n * np.eye(n) * 0.5
    b = rng.normal(size=n)
    x, rank = orc.fullpiv_solve(A, b)
    assert rank == n
    np.testing.assert_allclose(A @ x, b, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)
    xp = orc.partialpiv_solve(A, b)
    np.testing.assert_allclose(xp, x, rtol=1e-10, atol=1e-12)


def test_fullpiv_needs_column_pivoting(orc):
    # zero leading column block: partial pivoting on columns alone still works,
    # but a permuted matrix checks the column-permutation bookkeeping
    A = np.array([[0.0, 2.0, 1.0], [0.0, 1.0, 3.0], [4.0, 0.0, 1.0]])
    b = np.array([1.0, 2.0, 3.0])
    x, rank = orc.fullpiv_solve(A, b)
    assert rank == 3
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-14)


def test_fullpiv_rank_deficient(orc):
    A = np.array([[1.0, 2.0], [2.0, 4.0]])
    x, rank = orc.fullpiv_solve(A, np.array([1.0, 2.0]))
    assert rank == 1
    with pytest.raises(np.linalg.LinAlgError):
        orc.partialpiv_solve(np.zeros((2, 2)), np.ones(2))


def test_naca_camber_closed_form(orc):
    # NACA 4412: m = 0.04, p = 0.4 -> max camber m at x = p c, zero at ends
    assert orc.naca4digit(0.0, 1.0, 44) == 0.0
    assert abs(orc.naca4digit(0.4, 1.0, 44) - 0.04) < 1e-15
    assert abs(orc.naca4digit(1.0, 1.0, 44)) < 1e-15
    # continuity of the slope at x = p c
    h = 1e-6
    l = (orc.naca4digit(0.4, 1, 44) - orc.naca4digit(0.4 - h, 1, 44)) / h
    r = (orc.naca4digit(0.4 + h, 1, 44) - orc.naca4digit(0.4, 1, 44)) / h
    assert abs(l) < 1e-6 and abs(r) < 1e-6

"""Known-answer tests of the reference (test/*.cpp) applied to the oracle.

Each test cites the reference test it ports; tolerances are the reference's
(EPS = 1e-10, EXPECT_DOUBLE_EQ = 4 ulps).
"""
import math
import os

import numpy as np
import pytest

EPS = 1e-10


def dbl_eq(a, b):
    return abs(a - b) <= 4 * np.spacing(max(abs(a), abs(b), np.finfo(float).tiny))


def square_ring(l, x0=None, y0=None):
    """testutil.h:23-45 GetSquareRing"""
    if x0 is None:
        return np.array([[0, 0, 0], [l, 0, 0], [l, l, 0], [0, l, 0]], float)
    x0 -= l / 2
    y0 -= l / 2
    return np.array([[x0, y0, 0], [x0 + l, y0, 0], [x0 + l, y0 + l, 0], [x0, y0 + l, 0]], float)


# vortex_test.cpp:34-49
def test_infinite_line(orc):
    a, b = [0, 0, -1e10], [0, 0, 1e10]
    for r, expect in ((1, 1 / 2 / math.pi), (2, 1 / 4 / math.pi), (8, 1 / 16 / math.pi)):
        u = orc.biot_savart(a, b, [r, 0, 0])
        np.testing.assert_allclose(u, [0, expect, 0], atol=EPS)


# vortex_test.cpp:51-59
def test_finite_length(orc):
    u = orc.biot_savart([0, 0, -1], [0, 0, 1], [1, 0, 0])
    assert abs(u[0]) < EPS and abs(u[2]) < EPS
    assert abs(u[1] - math.sqrt(2) / 4 / math.pi) < EPS


# vortex_test.cpp:61-79
def test_midpoint_and_endpoint_cutoff(orc):
    a = np.array([math.sqrt(2), 2, math.pi])
    b = np.array([100, 2121, 12434.0])
    assert np.linalg.norm(orc.biot_savart(a, b, (a + b) / 2)) < EPS
    a = np.array([math.pi, math.sqrt(2), -1])
    b = np.array([0, 0, 1.0])
    assert np.linalg.norm(orc.biot_savart(a, b, a)) < EPS
    assert np.linalg.norm(orc.biot_savart(a, b, b)) < EPS


# vortex_test.cpp:104-130 (ring with Gamma=4 pi, side 2: -4 sqrt2 z at centre)
def test_square_ring(orc):
    ring = np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], float)
    u = orc.ring_induced_velocity([[0, 0, 0]], ring, [4 * math.pi])[0]
    assert u[0] == 0 and u[1] == 0 and dbl_eq(u[2], -4 * math.sqrt(2))
    u = orc.ring_induced_velocity([[0, 0, 0]], ring, [4 * math.pi], chordwise=True)[0]
    assert u[0] == 0 and u[1] == 0 and dbl_eq(u[2], -2 * math.sqrt(2))


# vortex_test.cpp:132-199, 201-232
def test_ring_geometry(orc):
    g = orc.ring_geometry(square_ring(20000))
    assert np.array_equal(g["normal"], [0, 0, 1])
    assert np.array_equal(g["tangent"], [1, 0, 0])
    lean = np.array([[0, 0, 0], [1, 0, 1], [1, 1, 1], [0, 1, 0]], float)
    n = orc.ring_geometry(lean)["normal"]
    assert dbl_eq(n[0], -math.sqrt(0.5)) and n[1] == 0 and dbl_eq(n[2], math.sqrt(0.5))
    lean = np.array([[0, 0, 0], [1, 0, 0.5], [1, 1, 0.5], [0, 1, 0]], float)
    t = orc.ring_geometry(lean)["tangent"]
    assert dbl_eq(t[0], 2 / math.sqrt(5)) and t[1] == 0 and dbl_eq(t[2], 1 / math.sqrt(5))
    imp = orc.ring_impulse(square_ring(2), 10)
    assert np.array_equal(imp, [0, 0, -40]) or np.allclose(imp, [0, 0, -40], rtol=1e-15)
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 2.1, 0], [0, 2, 0]], float)
    assert abs(orc.ring_geometry(v)["area"] - 2.05) < EPS


# shed_test.cpp:21-50: 20 identical rings superpose
def test_many_rings(orc):
    ring = np.array([[1, 1, 0], [-1, 1, 0], [-1, -1, 0], [1, -1, 0]], float)
    rings = np.repeat(ring[None], 20, axis=0)
    g = np.full(20, 4 * math.pi)
    u = orc.ring_induced_velocity([[0, 0, 0]], rings, g)[0]
    assert u[0] == 0 and u[1] == 0 and dbl_eq(u[2], -4 * math.sqrt(2) * 20)
    u = orc.ring_induced_velocity([[0, 0, 0]], rings, g, chordwise=True)[0]
    assert dbl_eq(u[2], -4 * math.sqrt(2) / 2 * 20)


def _legacy_solve(orc, rings, freestream, vmorph_n=None):
    """linear.inl:14-45,113-130: A_ij = ring_j(centroid_i; 1).n_i,
    rhs = Vmorph.n - U.n (the wake term is discarded by the reference)."""
    n = len(rings)
    geo = [orc.ring_geometry(r) for r in rings]
    A = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            u = orc.ring_induced_velocity([geo[i]["centroid"]], rings[j], [1.0])[0]
            A[i, j] = np.dot(u, geo[i]["normal"])
    rhs = np.array([(0.0 if vmorph_n is None else vmorph_n[i]) -
                    np.dot(freestream, geo[i]["normal"]) for i in range(n)])
    x, rank = orc.fullpiv_solve(A, rhs)
    assert rank == n
    return x


# linear_test.cpp:61-81: single tilted panel, Gamma = sqrt2 pi a U sin(alpha) / 4
def test_single_panel_closed_form(orc):
    alpha, U, a = math.pi / 6, 2.0, 0.1
    ring = np.array([[0, 0, 0], [a * math.cos(alpha), 0, -a * math.sin(alpha)],
                     [a * math.cos(alpha), a, -a * math.sin(alpha)], [0, a, 0]])
    g = _legacy_solve(orc, [ring], np.array([U, 0, 0]))
    expected = math.sqrt(2) * math.pi * a * U * math.sin(alpha) / 4
    assert abs(g[0] - expected) < EPS


# linear_test.cpp:83-116: two panels, no penetration at the centroids
@pytest.mark.parametrize("plunge", [0.0, 0.1])
def test_two_panels_no_penetration(orc, plunge):
    rings = [square_ring(2, -1, 0), square_ring(2, 1, 0)]
    # Morphing plug(t) = 0.1 sin t: surface velocity at t=0 is (0,0,0.1)
    vmn = [plunge, plunge]
    g = _legacy_solve(orc, rings, np.array([0, 0, 1.0]), vmn)
    flat = np.stack(rings)
    for x in ([-1, 0, 0], [1, 0, 0]):
        v = orc.ring_induced_velocity([x], flat, g)[0]
        np.testing.assert_allclose(v, [0, 0, -1 + plunge], atol=EPS)


# morphing_test.cpp / linear_test.cpp:27-49: flap sin(t), normal velocity
def test_morphing_flap_velocity(orc):
    spec = {"chord": 1.0, "span": 4.0, "flap": (1.0, 1.0, -math.pi / 2)}  # cos(t - pi/2) = sin t
    # flap about x at the root: a point at y=1 moves with z-velocity y * dphi/dt = 1
    v = orc.morph_velocity(spec, [0, 1, 0], 0.0)
    assert abs(v[2] - 1) < 1e-9
    v = orc.morph_velocity(spec, [0, -2, 0], 0.0)
    assert abs(v[2] - 2) < 1e-9  # mirrored side flaps up too


# morphing_test.cpp: plug motion moves every node by plug(t) in z
def test_morphing_plug(orc):
    spec = {"chord": 1.0, "span": 4.0, "plug": (0.3, 2.0, 0.0)}
    for t in (0.0, 0.7, 1.9):
        x = orc.morph_perform(spec, [0.2, -1.5, 0.0], t)
        np.testing.assert_allclose(x, [0.2, -1.5, 0.3 * math.cos(2 * t)], atol=1e-15)
        v = orc.morph_velocity(spec, [0.2, -1.5, 0.0], t)
        np.testing.assert_allclose(v, [0, 0, -0.6 * math.sin(2 * t)], atol=1e-6)


# wing_test.cpp:30-58 (WholeWing mirror of a 2x2 half wing, chord 1, span 2)
def test_whole_wing_mirror(orc):
    p = orc.build_wing(orc.GEN_RECT, 0, 1.0, 4.0, 2, 4).reshape(-1, 3)
    assert p.shape == (15, 3)
    for k, e in ((0, [0, -2, 0]), (1, [0, -1, 0]), (2, [0, 0, 0]), (3, [0, 1, 0]),
                 (4, [0, 2, 0]), (5, [0.5, -2, 0]), (10, [1, -2, 0]), (14, [1, 2, 0])):
        np.testing.assert_allclose(p[k], e, atol=EPS)
    q = orc.build_wing(orc.GEN_RECT, 0, 1.0, 4.0, 2, 4, origin=(1, 2, 3)).reshape(-1, 3)
    np.testing.assert_allclose(q - p, np.tile([1, 2, 3], (15, 1)), atol=EPS)


def test_odd_columns_rejected(orc):
    with pytest.raises(ValueError):
        orc.build_wing(orc.GEN_RECT, 0, 1.0, 4.0, 2, 3)


# naca4digit_test.cpp:55-79 against the reference's own data file
def test_naca62xx_fixture(orc):
    path = os.path.join(os.path.dirname(__file__), "golden", "NACA62XX.dat")
    data = np.loadtxt(path)
    err = np.mean([(y - orc.naca4digit(x, 100, 62)) ** 2 for x, y in data])
    assert err < 1e-6


# rect_test.cpp / util_test: linspace
def test_linspace(orc):
    np.testing.assert_array_equal(orc.linspace(0, 1, 5), [0, 0.25, 0.5, 0.75, 1.0])

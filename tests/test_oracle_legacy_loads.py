"""Pins of the oracle's Lamb vector (calc_impulse.h:14-30) and legacy loads
(katz_plotkin.cpp:16-68, joukowski.h:63-99).

The Lamb vector is held to the reference's own known answers
(test/calc_impulse_test.cpp:6-52).  The loads are held to closed forms on a
flat plate with uneven spacing (legacy ring nodes (i,j), (i+1,j),
(i+1,j+1), (i,j+1): normal +z, tangent +x, CalcC = dx_i, CalcB = dy_j):

* Katz-Plotkin, steady: the chordwise differences of G telescope to the
  trailing-edge row, so F = rho U cos^2(a) sum_j G_TE,j dy_j P n, and the
  drag term vanishes (Um . P n = 0 for any geometry when the wake and
  chordwise-bound velocities are zero, as in the legacy loop).
* Katz-Plotkin, unsteady: dG/dt averaged with the ring ahead, the leading
  row halved: F_un = rho (d/dt) [dx_0/2 + sum_{i>0} dx_i] sum_j dy_j
  (u sin a + P n cos a) for a uniform change d.
* Pin = -(Um . n) F . n for a uniform Um.
* Container Joukowski: with zero source circulation every closed ring's
  sum of dl vanishes except the trailing-edge row, whose skipped edge
  leaves F = rho (U - Vm) x sum_j G_TE,j dy_j e_y (Kutta-Joukowski), and
  the unsteady term is sum rho n dG/dt dx dy.
"""
import numpy as np

SPEC0 = {"chord": 1.0, "span": 2.0, "alpha": 0.0}  # no motion


def square_ring(l):
    return np.array([[0, 0, 0], [l, 0, 0], [l, l, 0], [0, l, 0]], dtype=np.float64)


def test_lamb_vector_reference_cases(orc):
    """calc_impulse_test.cpp:6-52 (no other vortex)."""
    vx = np.array([[1, 0, 0], [2, 0, 0], [3, 0, 0], [4, 0, 0]], dtype=np.float64)
    vy = vx[:, [1, 0, 2]]
    np.testing.assert_array_equal(orc.lamb_vector(square_ring(2), 1.0, vx, np.zeros(3)),
                                  [0, 0, 4])
    np.testing.assert_array_equal(orc.lamb_vector(square_ring(2), 2.0, vx, np.zeros(3)),
                                  [0, 0, 8])
    np.testing.assert_array_equal(orc.lamb_vector(square_ring(2), 1.0, vy, np.zeros(3)),
                                  [0, 0, -4])


def test_lamb_vector_uniform_velocity_vanishes(orc):
    """A closed ring in a uniform velocity: sum of (u x dl) = u x 0."""
    v = np.tile([0.3, -0.7, 0.2], (4, 1))
    r = orc.lamb_vector(square_ring(1.5), 2.5, v, np.array([0.1, 0.0, 0.4]))
    assert np.abs(r).max() < 1e-15


def plate(xs, ys):
    R, C = len(xs) - 1, len(ys) - 1
    rings = np.zeros((R, C, 4, 3))
    for i in range(R):
        for j in range(C):
            rings[i, j, 0] = [xs[i], ys[j], 0]
            rings[i, j, 1] = [xs[i + 1], ys[j], 0]
            rings[i, j, 2] = [xs[i + 1], ys[j + 1], 0]
            rings[i, j, 3] = [xs[i], ys[j + 1], 0]
    return rings.reshape(-1, 4, 3)


XS = np.array([0.0, 0.1, 0.35, 0.7, 1.0])
YS = np.array([-1.0, -0.4, 0.1, 0.5, 1.0])


def test_katz_plotkin_steady_closed_form(orc):
    R, C = len(XS) - 1, len(YS) - 1
    a = 0.17
    U = np.array([np.cos(a), 0.0, np.sin(a)])
    rng = np.random.default_rng(1)
    g = rng.normal(size=R * C)
    out = orc.katz_plotkin_load(plate(XS, YS), g, g, R, C, SPEC0, U, 1.3, 0.0, 0.1)
    u = U / np.linalg.norm(U)
    Pn = np.array([0, 0, 1.0]) - u * u[2]
    gte = g.reshape(R, C)[-1]
    want = 1.3 * np.cos(a) ** 2 * np.sum(gte * np.diff(YS)) * Pn
    np.testing.assert_allclose(out[:3], want, rtol=1e-13, atol=1e-15)
    # Pin = -(Um . n) (F . n), Pout = 0
    np.testing.assert_allclose(out[3], -np.sin(a) * out[2], rtol=1e-13)
    assert out[4] == 0.0


def test_katz_plotkin_unsteady_row_average(orc):
    R, C = len(XS) - 1, len(YS) - 1
    a = 0.11
    U = np.array([np.cos(a), 0.0, np.sin(a)])
    g = np.random.default_rng(2).normal(size=R * C)
    d, dt, rho = 0.3, 0.05, 0.9
    with_d = orc.katz_plotkin_load(plate(XS, YS), g, g - d, R, C, SPEC0, U, rho, 0.0, dt)
    base = orc.katz_plotkin_load(plate(XS, YS), g, g, R, C, SPEC0, U, rho, 0.0, dt)
    u = U / np.linalg.norm(U)
    Pn = np.array([0, 0, 1.0]) - u * u[2]
    dx = np.diff(XS)
    W = (dx[0] / 2 + dx[1:].sum()) * np.diff(YS).sum()
    want = rho * d / dt * W * (u * np.sin(a) + Pn * np.cos(a))
    np.testing.assert_allclose((with_d - base)[:3], want, rtol=1e-12, atol=1e-14)


def test_container_joukowski_closed_forms(orc):
    R, C = len(XS) - 1, len(YS) - 1
    rings = plate(XS, YS)
    g = np.random.default_rng(3).normal(size=R * C)
    U = np.array([1.2, 0.1, -0.05])
    src, src_g = rings, np.zeros(R * C)  # the sources carry no circulation
    rho, dt = 1.1, 0.07
    st = orc.joukowski_container_load(rings, rings, g, g, R, C, src, src_g, SPEC0, U, rho,
                                      0.0, dt)
    gte = g.reshape(R, C)[-1]
    want = rho * np.cross(U, np.array([0.0, np.sum(gte * np.diff(YS)), 0.0]))
    np.testing.assert_allclose(st[:3], want, rtol=1e-13, atol=1e-15)
    assert st[3] == 0.0  # no surface motion
    np.testing.assert_allclose(st[4], -st[0] * np.linalg.norm(U), rtol=1e-15)
    # unsteady: n rho dG/dt CalcB CalcC per ring
    delta = np.random.default_rng(4).normal(size=R * C)
    un = orc.joukowski_container_load(rings, rings, g, g - delta, R, C, src, src_g, SPEC0, U,
                                      rho, 0.0, dt)
    area = (np.diff(XS)[:, None] * np.diff(YS)[None, :]).ravel()
    np.testing.assert_allclose((un - st)[:3], [0, 0, rho * np.sum(delta * area) / dt],
                               rtol=1e-12, atol=1e-14)


def test_container_joukowski_plunge(orc):
    """A plunging plate: the surface velocity Vm is uniform, so the steady
    part is rho (U - Vm) x (TE circulation), and Pin = F . Vm."""
    R, C = len(XS) - 1, len(YS) - 1
    rings = plate(XS, YS)
    g = np.random.default_rng(5).normal(size=R * C)
    spec = dict(SPEC0, plug=(0.2, 1.3, 0.4))
    t = 0.6
    vm = orc.morph_velocity(spec, np.zeros(3), t)
    U = np.array([1.0, 0.0, 0.0])
    out = orc.joukowski_container_load(rings, rings, g, g, R, C, rings, np.zeros(R * C), spec,
                                       U, 1.0, t, 0.1)
    gte = g.reshape(R, C)[-1]
    want = np.cross(U - vm, np.array([0.0, np.sum(gte * np.diff(YS)), 0.0]))
    np.testing.assert_allclose(out[:3], want, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(out[3], out[:3] @ vm, rtol=1e-9, atol=1e-12)

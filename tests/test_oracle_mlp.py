"""The network oracle (oracle/mlp.py) pinned by mathematics: the gradient
against central differences, Adam's first step against its closed form,
the moment expansion against the direct Biot-Savart sum (convergence at
its order), and the features of a unit ring against hand values."""
import numpy as np
import pytest

from oracle import mlp as M


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(0)
    w = M.init(3).astype(np.float64)
    w[:-1] += 0.05 * rng.normal(size=w.size - 1)  # non-zero biases too
    X = rng.normal(size=(7, M.IN))
    V = rng.normal(size=(7, M.OUT))
    _, g = M.loss_and_grad(w, X, V)
    idx = rng.choice(w.size - 1, size=60, replace=False)
    h = 1e-6
    for k in idx:
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fd = (M.loss_and_grad(wp, X, V)[0] - M.loss_and_grad(wm, X, V)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), k
    assert g[-1] == 0.0


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(1)
    w, g = rng.normal(size=50), rng.normal(size=50)
    w1, m, v = M.adam_step(w, g, np.zeros(50), np.zeros(50), 1, 1e-3)
    # t = 1: m_hat = g, v_hat = g^2
    np.testing.assert_allclose(w1, w - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)


def test_unit_ring_features():
    """A unit square ring of the lattice in the xy-plane (node row i along x,
    column j along y: the ring runs (0,0) -> (0,1) -> (1,1) -> (1,0), clockwise
    seen from +z): Q0 = 0 (a closed loop), Q1[i][j] = Gamma int x_j dx_i."""
    nodes = np.array([[[0.0, 0.0, 0.0], [0.0, 1.0, 0.0]], [[1.0, 0.0, 0.0], [1.0, 1.0, 0.0]]])
    c, r, Q = M.block_moments(nodes, np.array([[2.0]]))
    np.testing.assert_allclose(c, [0.5, 0.5, 0.0])
    assert r == pytest.approx(np.sqrt(0.5))
    np.testing.assert_allclose(Q[:3], 0.0, atol=1e-15)
    Q1 = Q[3:12].reshape(3, 3)
    want = np.zeros((3, 3))
    want[1, 0] = -2.0  # int x dy around the loop
    want[0, 1] = 2.0   # int y dx
    np.testing.assert_allclose(Q1, want, atol=1e-15)
    f = M.features([0.5, 0.5, 10.0], c, r, Q)
    np.testing.assert_allclose(f[:3], [0, 0, 1], atol=1e-15)
    assert f[3] == pytest.approx(np.sqrt(0.5) / 10.0)
    s = M.scale_of(Q, r)
    assert s > np.sqrt(8.0) / r  # |Q1| / r plus the higher moments' terms
    np.testing.assert_allclose(f[7:16], (want / (s * r)).reshape(-1), atol=1e-15)
    assert np.abs(f[4:31]).max() <= 1.0


def test_open_cluster_expansion_order(orc):
    """64 random open segments: the expansion (Q0..Q3) against the direct
    sum of BiotSavartLaw over the segments; the error falls ~16x per
    doubling of the distance (first omitted order 4 against the Q0 term)."""
    rng = np.random.default_rng(5)
    A = rng.normal(scale=0.5, size=(64, 3))
    E = rng.normal(scale=0.3, size=(64, 3))
    G = rng.normal(size=64)
    c = (A + 0.5 * E).mean(0)
    r = max(np.linalg.norm(A - c, axis=1).max(), np.linalg.norm(A + E - c, axis=1).max())
    Q = M.segment_moments(A, E, G, c)
    dirs = rng.normal(size=(16, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    errs = []
    for ratio in (4.0, 8.0, 16.0):
        x = c + ratio * r * dirs
        ref = np.zeros_like(x)
        for k in range(len(x)):
            for q in range(64):
                ref[k] += G[q] * orc.biot_savart(A[q], A[q] + E[q], x[k])
        got = np.array([M.expansion(xx, c, Q) for xx in x])
        errs.append(np.max(np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)))
    assert errs[0] < 2e-2
    assert errs[1] < errs[0] / 10 and errs[2] < errs[1] / 10


def test_moments_translate_consistently():
    """Moments about c2 equal those about c1 moved by the parent translation
    (Q1' = Q1 + Q0 t, ... with t = c1 - c2)."""
    rng = np.random.default_rng(6)
    A, E, G = rng.normal(size=(20, 3)), rng.normal(size=(20, 3)), rng.normal(size=20)
    c1, c2 = rng.normal(size=3), rng.normal(size=3)
    q = M.segment_moments(A, E, G, c1)
    t = c1 - c2
    Q0, Q1 = q[:3], q[3:12].reshape(3, 3)
    Q2, Q3 = q[12:39].reshape(3, 3, 3), q[39:].reshape(3, 3, 3, 3)
    e = np.einsum
    T1 = Q1 + e("i,j->ij", Q0, t)
    T2 = Q2 + e("ij,k->ijk", Q1, t) + e("ik,j->ijk", Q1, t) + e("i,j,k->ijk", Q0, t, t)
    T3 = (Q3 + e("ijk,l->ijkl", Q2, t) + e("ijl,k->ijkl", Q2, t) + e("ikl,j->ijkl", Q2, t)
          + e("ij,k,l->ijkl", Q1, t, t) + e("ik,j,l->ijkl", Q1, t, t)
          + e("il,j,k->ijkl", Q1, t, t) + e("i,j,k,l->ijkl", Q0, t, t, t))
    moved = np.concatenate([Q0, T1.reshape(-1), T2.reshape(-1), T3.reshape(-1)])
    np.testing.assert_allclose(moved, M.segment_moments(A, E, G, c2), rtol=1e-12, atol=1e-12)


def test_expansion_converges_to_direct_sum(orc):
    """Block moments + expansion against the oracle's direct sum: the error
    falls ~8x per doubling of the distance (first omitted order 4)."""
    rng = np.random.default_rng(2)
    i, j = np.meshgrid(np.arange(5), np.arange(17), indexing="ij")
    nodes = np.stack([0.3 * i, 0.1 * j, 0.05 * np.sin(0.5 * i + 0.3 * j)], -1)
    nodes = nodes + rng.normal(scale=0.01, size=nodes.shape)
    gamma = 0.1 + 0.03 * rng.normal(size=(4, 16))
    c, r, Q = M.block_moments(nodes, gamma)
    dirs = rng.normal(size=(20, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    errs = []
    for ratio in (4.0, 8.0, 16.0):
        x = c + ratio * r * dirs
        ref = orc.induced_velocity(x, nodes[None], gamma[None])
        got = np.array([M.expansion(xx, c, Q) for xx in x])
        errs.append(np.max(np.linalg.norm(got - ref, axis=1) / np.linalg.norm(ref, axis=1)))
    assert errs[0] < 2e-2
    assert errs[1] < errs[0] / 6 and errs[2] < errs[1] / 6


def test_ring_and_segment_partitions_agree():
    """The rings of a lattice and its segments (Gamma differences) carry the
    same moments: the ring tree's sources are the segment form's field."""
    rng = np.random.default_rng(7)
    nodes = rng.normal(size=(2, 4, 6, 3))
    gamma = rng.normal(size=(2, 3, 5))
    c = rng.normal(size=3)
    qs = M.segment_moments(*M.lattice_segments(nodes, gamma), c)
    qr = M.segment_moments(*M.ring_segments(*M.lattice_rings(nodes, gamma)), c)
    np.testing.assert_allclose(qr, qs, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(qr[:3], 0.0, atol=1e-12)  # closed loops


def test_tile_labels_are_the_taylor_polynomial_of_the_remainder(orc):
    """The training target of a (tile, cluster) pair is the remainder's
    Taylor polynomial about the tile centre: the polynomial it defines
    (tile_correction) matches the exact remainder with an error of third
    order in the offset (~8x smaller per halving), which a wrong sign, a
    missing 1/2 or a transposed index of the differences would break."""
    rng = np.random.default_rng(4)
    i, j = np.meshgrid(np.arange(5), np.arange(17), indexing="ij")
    nodes = np.stack([0.3 * i, 0.1 * j, 0.05 * np.sin(0.5 * i + 0.3 * j)], -1)
    gamma = 0.1 + 0.03 * rng.normal(size=(4, 16))
    c, r, Q = M.block_moments(nodes, gamma)
    tc = c + 3.0 * r * np.array([0.6, -0.3, 0.74]) / np.linalg.norm([0.6, -0.3, 0.74])

    def exact(x):
        return orc.induced_velocity(np.asarray(x), nodes[None], gamma[None])

    y = M.tile_labels(tc, c, r, Q, exact)
    D = np.linalg.norm(tc - c)
    u = np.array([0.3, 0.8, -0.52])
    u /= np.linalg.norm(u)
    errs, rems = [], []
    for eps in (0.16, 0.08, 0.04):  # below 0.04 the labels' own O(h^2) error shows
        x = tc + eps * D * u
        rem = exact(x[None])[0] - M.expansion(x, c, Q)
        errs.append(np.linalg.norm(rem - M.tile_correction(x, tc, c, r, Q, y)))
        rems.append(np.linalg.norm(rem))
    assert errs[0] < 0.1 * rems[0]
    assert errs[0] / errs[1] > 6.0 and errs[1] / errs[2] > 6.0
    assert M.tile_pair_distance(tc, 0.2 * D, c, r, 0) == D
    assert M.tile_pair_distance(tc, 0.4 * D, c, r, 0) == 0.0
    assert M.tile_pair_distance(tc, 0.2 * D, c, r, 3) == 0.0

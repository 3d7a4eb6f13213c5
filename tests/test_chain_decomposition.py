"""The fast modes sum a lattice as chains of merged segments (an interior
segment carries the difference of the circulations of its two rings).  This
checks that decomposition on the CPU with the oracle's segment law: chains
reproduce the ring-by-ring sum (velocity.h:14-24) to rounding."""
import numpy as np


def chains_of(nodes, gamma):
    """Spanwise: (i,j-1)->(i,j) with G(i,j-1) - G(i-1,j-1); chordwise:
    (i-1,j)->(i,j) with G(i-1,j-1) - G(i-1,j); G = 0 outside."""
    num, R1, C1, _ = nodes.shape
    R, C = R1 - 1, C1 - 1

    def G(n, i, j):
        return gamma[n, i, j] if 0 <= i < R and 0 <= j < C else 0.0

    segs = []
    for n in range(num):
        for i in range(R1):
            for j in range(1, C1):
                segs.append((nodes[n, i, j - 1], nodes[n, i, j], G(n, i, j - 1) - G(n, i - 1, j - 1)))
        for j in range(C1):
            for i in range(1, R1):
                segs.append((nodes[n, i - 1, j], nodes[n, i, j], G(n, i - 1, j - 1) - G(n, i - 1, j)))
    return segs


def test_chain_sum_equals_ring_sum(orc):
    rng = np.random.default_rng(4)
    nodes = rng.normal(scale=0.05, size=(2, 4, 6, 3))
    nodes[..., 0] += np.arange(4)[None, :, None] * 0.3
    nodes[..., 1] += np.arange(6)[None, None, :] * 0.25
    nodes[1, ..., 0] += 2.0
    gamma = rng.normal(size=(2, 3, 5))
    x = rng.normal(size=(20, 3)) * 1.5
    ref = orc.induced_velocity(x, nodes, gamma)
    segs = chains_of(nodes, gamma)
    assert len(segs) == 2 * (4 * 5 + 3 * 6)
    got = np.zeros_like(ref)
    for a, b, g in segs:
        for k in range(x.shape[0]):
            got[k] += g * orc.biot_savart(a, b, x[k])
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-13 * np.abs(ref).max())


def test_chain_sum_at_lattice_nodes(orc):
    """Targets on the lattice itself (wake nodes): the cut-off zeroes the
    segments that touch or are collinear with the target in both forms."""
    rng = np.random.default_rng(6)
    nodes = rng.normal(scale=0.02, size=(1, 5, 7, 3))
    nodes[..., 0] += np.arange(5)[None, :, None] * 0.2
    nodes[..., 1] += np.arange(7)[None, None, :] * 0.2
    gamma = rng.normal(size=(1, 4, 6))
    x = nodes.reshape(-1, 3)
    ref = orc.induced_velocity(x, nodes, gamma)
    got = np.zeros_like(ref)
    for a, b, g in chains_of(nodes, gamma):
        for k in range(x.shape[0]):
            got[k] += g * orc.biot_savart(a, b, x[k])
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())

"""Numpy restatement of the far-field split's block moments and expansion
and of the correction network: vorticity moments, the moment expansion,
network features, forward pass, loss gradient and Adam step, in float64.

TEST INFRASTRUCTURE ONLY.  This is the checker for the GPU far-field and
network kernels: only ``tests/`` import it.  Neither has a counterpart in
the reference (xeno14/UVLM sums every wake ring exactly,
src/uvlm/advect.h:16-46); their definitions are this module's:

* rings of a lattice in VORING order (vortex.cpp:145-160); segments of a
  lattice (another partition of the same field): per sheet,
  spanwise (i, j) -> (i, j+1) by node row, then chordwise (i, j) -> (i+1, j)
  by node column, strength = the difference of the ring circulations either
  side (zero outside);
* moments of a set of segments about c: Q_n[i, J] = sum Gamma e_i
  int_0^1 (a + t e - c)^(x)n dt, n = 0..3 (3 + 9 + 27 + 81 values);
* expansion: u_a = 1/4pi sum_n (-1)^n / n! eps_abc Q_n[c, J] T_{n+1}[b, J];
* features of a (target x, cluster) pair, cluster centre c, radius r:
  d = x - c, D = |d|, dh = d / D, rho = r / D,
  s = |Q0| + |Q1| / r + |Q2| / r^2 + |Q3| / r^3 (Frobenius norms),
  f = [dh (3), rho, Q0 / s (3), Q1 / (s r) (9), Q2 . dh / (s r^2) (9),
       Q2 : dh dh / (s r^2) (3), Q3 : dh dh dh / (s r^3) (3), 1];
* network: h1 = silu(W1 f + b1), h2 = silu(W2 h1 + b2), h3 = silu(W3 h2 + b3),
  y = W4 h3 + b4 (32 -> 128 -> 128 -> 128 -> 30);
* one evaluation per (target tile, cluster), features at the tile centre
  c_T (x = c_T above); a pair is corrected when the cluster is at most
  level 2, D = |c_T - c| > r and the tile radius is at most 0.3 D;
* target: the Taylor coefficients about c_T of the expansion's remainder
  R(x) = (u_exact - u_expansion)(x) / scale, scale = s rho^4 / (4 pi D^2),
  in delta = (x - c_T) / (0.15 D): value R (3), gradient G[i][j] (9),
  Hessian H[i][jk] for jk = 00, 01, 02, 11, 12, 22 (18), by central
  differences of step h = 0.02 D over 0, +-e_j, +-e_j +-e_k;
* correction at a target x of the tile: scale (R + G delta + 1/2 delta H delta);
* loss: mean over the batch of 1/2 |y - y*|^2.

Parameter vector (FP32 on the device): W1 [128][32], b1, W2 [128][128], b2,
W3 [128][128], b3, W4 [30][128], b4 [30], two pad words.
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT = 32, 128, 30
MAX_LEVEL, TILE_RATIO, FD_STEP, TAYLOR_UNIT = 2, 0.3, 0.02, 0.15
HESS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
SHAPES = [("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (HID, HID)), ("b3", (HID,)), ("W4", (OUT, HID)), ("b4", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES) + 2  # + pad


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = w[o:o + n].reshape(shp)
        o += n
    return out


def pack(p) -> np.ndarray:
    w = np.concatenate([np.asarray(p[name], dtype=np.float64).reshape(-1) for name, _ in SHAPES])
    return np.concatenate([w, [0.0, 0.0]])


def init(seed: int = 0) -> np.ndarray:
    """He-style uniform initialisation, output layer small."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shp in SHAPES:
        if name.startswith("W"):
            fan_in = shp[1]
            a = np.sqrt(3.0 / fan_in) * (0.1 if name == "W4" else 1.0)
            p[name] = rng.uniform(-a, a, size=shp)
        else:
            p[name] = np.zeros(shp)
    return pack(p).astype(np.float32)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(w, X):
    """y [n][3] for features X [n][32]; also the pre-activations."""
    p = unpack(w)
    X = np.asarray(X, dtype=np.float64)
    z1 = X @ p["W1"].T + p["b1"]
    h1 = silu(z1)
    z2 = h1 @ p["W2"].T + p["b2"]
    h2 = silu(z2)
    z3 = h2 @ p["W3"].T + p["b3"]
    h3 = silu(z3)
    y = h3 @ p["W4"].T + p["b4"]
    return y, (X, z1, h1, z2, h2, z3, h3)


def loss_and_grad(w, X, V):
    """Loss and its gradient in the parameter layout (pad word 0)."""
    p = unpack(w)
    y, (X, z1, h1, z2, h2, z3, h3) = forward(w, X)
    n = X.shape[0]
    dy = (y - V) / n
    loss = 0.5 * np.sum((y - V) ** 2) / n
    g = {"W4": dy.T @ h3, "b4": dy.sum(0)}
    dz3 = (dy @ p["W4"]) * dsilu(z3)
    g["W3"], g["b3"] = dz3.T @ h2, dz3.sum(0)
    dz2 = (dz3 @ p["W3"]) * dsilu(z2)
    g["W2"], g["b2"] = dz2.T @ h1, dz2.sum(0)
    dz1 = (dz2 @ p["W2"]) * dsilu(z1)
    g["W1"], g["b1"] = dz1.T @ X, dz1.sum(0)
    return loss, pack(g)


def adam_step(w, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t counted from 1); returns new (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


def lattice_segments(nodes, gamma):
    """Segments (a [n][3], e [n][3], g [n]) of lattices nodes [num][R+1][C+1][3],
    gamma [num][R][C], in the order above."""
    nodes = np.asarray(nodes, dtype=np.float64)
    gamma = np.asarray(gamma, dtype=np.float64)
    A, E, Gs = [], [], []
    for n in range(gamma.shape[0]):
        P, g = nodes[n], gamma[n]
        R, C = g.shape
        G = np.zeros((R + 2, C + 2))
        G[1:-1, 1:-1] = g
        for i in range(R + 1):
            for j in range(C):
                A.append(P[i, j]); E.append(P[i, j + 1] - P[i, j])
                Gs.append(G[i + 1, j + 1] - G[i, j + 1])
        for j in range(C + 1):
            for i in range(R):
                A.append(P[i, j]); E.append(P[i + 1, j] - P[i, j])
                Gs.append(G[i + 1, j] - G[i + 1, j + 1])
    return np.array(A), np.array(E), np.array(Gs)


def segment_moments(A, E, G, c):
    """Moments [120] (Q0, Q1, Q2, Q3 flattened) of segments about c."""
    Q = [np.zeros(3), np.zeros((3, 3)), np.zeros((3, 3, 3)), np.zeros((3, 3, 3, 3))]
    for a, e, g in zip(A, E, G):
        if g == 0.0:
            continue
        Ar = a - c
        m1 = Ar + e / 2
        m2 = np.outer(Ar, Ar) + (np.outer(Ar, e) + np.outer(e, Ar)) / 2 + np.outer(e, e) / 3
        ein = lambda *v: np.einsum("j,k,l->jkl", *v)
        m3 = (ein(Ar, Ar, Ar) + (ein(Ar, Ar, e) + ein(Ar, e, Ar) + ein(e, Ar, Ar)) / 2
              + (ein(Ar, e, e) + ein(e, Ar, e) + ein(e, e, Ar)) / 3 + ein(e, e, e) / 4)
        Q[0] += g * e
        Q[1] += g * np.outer(e, m1)
        Q[2] += g * np.einsum("i,jk->ijk", e, m2)
        Q[3] += g * np.einsum("i,jkl->ijkl", e, m3)
    return np.concatenate([q.reshape(-1) for q in Q])


def lattice_rings(nodes, gamma):
    """Rings (nodes P [n][4][3] in VORING order p0 = (i,j), p1 = (i,j+1),
    p2 = (i+1,j+1), p3 = (i+1,j), circulations G [n]) of lattices nodes
    [num][R+1][C+1][3], gamma [num][R][C], ordered (n, i, j)."""
    nodes = np.asarray(nodes, dtype=np.float64)
    gamma = np.asarray(gamma, dtype=np.float64)
    P = np.stack([nodes[:, :-1, :-1], nodes[:, :-1, 1:], nodes[:, 1:, 1:], nodes[:, 1:, :-1]], -2)
    return P.reshape(-1, 4, 3), gamma.reshape(-1)


def ring_segments(P, G):
    """The rings' segments p0 -> p1 -> p2 -> p3 -> p0 as (A, E, G)."""
    A = P.reshape(-1, 3)
    E = (np.roll(P, -1, axis=1) - P).reshape(-1, 3)
    return A, E, np.repeat(G, 4)


def scale_of(Q, r):
    """s = |Q0| + |Q1| / r + |Q2| / r^2 + |Q3| / r^3 of a cluster of radius r."""
    Q = np.asarray(Q, dtype=np.float64)
    return (np.linalg.norm(Q[:3]) + np.linalg.norm(Q[3:12]) / r
            + np.linalg.norm(Q[12:39]) / r ** 2 + np.linalg.norm(Q[39:120]) / r ** 3)


def features(x, c, r, Q):
    """Features [32] of target x and a cluster (centre c, radius r, moments
    Q [120]); None when the cluster carries no circulation."""
    Q = np.asarray(Q, dtype=np.float64)
    Q0, Q1 = Q[:3], Q[3:12].reshape(3, 3)
    Q2, Q3 = Q[12:39].reshape(3, 3, 3), Q[39:120].reshape(3, 3, 3, 3)
    d = np.asarray(x, dtype=np.float64) - np.asarray(c, dtype=np.float64)
    D = np.linalg.norm(d)
    dh = d / D
    s = scale_of(Q, r)
    if s == 0.0:
        return None
    f = np.zeros(IN)
    f[0:3] = dh
    f[3] = r / D
    f[4:7] = Q0 / s
    f[7:16] = Q1.reshape(-1) / (s * r)
    q2d = Q2 @ dh
    f[16:25] = q2d.reshape(-1) / (s * r * r)
    f[25:28] = q2d @ dh / (s * r * r)
    f[28:31] = ((Q3 @ dh) @ dh) @ dh / (s * r ** 3)
    f[31] = 1.0
    return f


def target_scale(x, c, r, Q):
    """s rho^4 / (4 pi D^2): the correction per unit network output."""
    d = np.asarray(x, dtype=np.float64) - np.asarray(c, dtype=np.float64)
    D = np.linalg.norm(d)
    return scale_of(Q, r) * (r / D) ** 4 / (4.0 * np.pi * D ** 2)


def expansion(x, c, Q):
    """Velocity of the moment expansion (orders 0..3) of a cluster at x."""
    Q = np.asarray(Q, dtype=np.float64)
    Q0, Q1 = Q[:3], Q[3:12].reshape(3, 3)
    Q2, Q3 = Q[12:39].reshape(3, 3, 3), Q[39:120].reshape(3, 3, 3, 3)
    d = np.asarray(x, dtype=np.float64) - np.asarray(c, dtype=np.float64)
    D = np.sqrt(d @ d)
    I = np.eye(3)
    T1 = -d / D ** 3
    T2 = 3 * np.einsum("b,j->bj", d, d) / D ** 5 - I / D ** 3
    T3 = (-15 * np.einsum("b,j,k->bjk", d, d, d) / D ** 7
          + 3 * (np.einsum("jb,k->bjk", I, d) + np.einsum("kb,j->bjk", I, d)
                 + np.einsum("jk,b->bjk", I, d)) / D ** 5)
    dddd = np.einsum("b,j,k,l->bjkl", d, d, d, d)
    six = (np.einsum("jb,k,l->bjkl", I, d, d) + np.einsum("kb,j,l->bjkl", I, d, d)
           + np.einsum("lb,j,k->bjkl", I, d, d) + np.einsum("jk,b,l->bjkl", I, d, d)
           + np.einsum("jl,b,k->bjkl", I, d, d) + np.einsum("kl,b,j->bjkl", I, d, d))
    three = (np.einsum("jb,kl->bjkl", I, I) + np.einsum("kb,jl->bjkl", I, I)
             + np.einsum("lb,jk->bjkl", I, I))
    T4 = 105 * dddd / D ** 9 - 15 * six / D ** 7 + 3 * three / D ** 5
    S = (np.einsum("c,b->cb", Q0, T1) - np.einsum("cj,bj->cb", Q1, T2)
         + 0.5 * np.einsum("cjk,bjk->cb", Q2, T3) - np.einsum("cjkl,bjkl->cb", Q3, T4) / 6.0)
    u = np.array([S[2, 1] - S[1, 2], S[0, 2] - S[2, 0], S[1, 0] - S[0, 1]])
    return u / (4.0 * np.pi)


def block_moments(nodes, gamma):
    """Centre (node mean), radius and moments [120] of one lattice block:
    nodes [R+1][C+1][3], ring circulations gamma [R][C] (zero outside)."""
    nodes = np.asarray(nodes, dtype=np.float64)
    c = nodes.reshape(-1, 3).mean(0)
    r = np.linalg.norm(nodes.reshape(-1, 3) - c, axis=1).max()
    A, E, G = lattice_segments(nodes[None], np.asarray(gamma)[None])
    return c, r, segment_moments(A, E, G, c)


def tile_pair_distance(tc, tr, c, r, level):
    """D = |c_T - c| when the network corrects the pair (tile centre tc,
    radius tr; cluster centre c, radius r, level), else 0."""
    D = float(np.linalg.norm(np.asarray(tc, dtype=np.float64) - np.asarray(c, dtype=np.float64)))
    return D if (level <= MAX_LEVEL and D > r and tr <= TILE_RATIO * D) else 0.0


def tile_labels(tc, c, r, Q, exact):
    """Training target [30] of the pair (tile centre tc, cluster c, r, Q):
    the remainder's Taylor coefficients by the central differences above;
    exact(points [m][3]) -> the cluster's exact velocity [m][3]."""
    tc = np.asarray(tc, dtype=np.float64)
    D = float(np.linalg.norm(tc - c))
    scale = target_scale(tc, c, r, Q)
    h = FD_STEP

    def R(o):
        x = tc + D * h * np.asarray(o, dtype=np.float64)
        return (exact(x[None])[0] - expansion(x, c, Q)) / scale

    E = np.eye(3)
    y = np.zeros(OUT)
    R0 = R(np.zeros(3))
    y[0:3] = R0
    Rp = [R(E[j]) for j in range(3)]
    Rm = [R(-E[j]) for j in range(3)]
    for i in range(3):
        for j in range(3):
            y[3 + 3 * i + j] = TAYLOR_UNIT * (Rp[j][i] - Rm[j][i]) / (2 * h)
    for m, (j, k) in enumerate(HESS):
        if j == k:
            hv = (Rp[j] - 2 * R0 + Rm[j]) / h ** 2
        else:
            hv = (R(E[j] + E[k]) - R(E[j] - E[k]) - R(-E[j] + E[k]) + R(-E[j] - E[k])) / (4 * h * h)
        for i in range(3):
            y[12 + 6 * i + m] = TAYLOR_UNIT ** 2 * hv[i]
    return y


def tile_correction(x, tc, c, r, Q, y):
    """Velocity correction at x from outputs y [30] of the pair (tile
    centre tc, cluster c, r, Q): scale (R + G d + 1/2 d H d),
    d = (x - tc) / (0.15 D)."""
    tc = np.asarray(tc, dtype=np.float64)
    D = float(np.linalg.norm(tc - c))
    d = (np.asarray(x, dtype=np.float64) - tc) / (TAYLOR_UNIT * D)
    y = np.asarray(y, dtype=np.float64)
    H = np.zeros((3, 3, 3))
    for m, (j, k) in enumerate(HESS):
        H[:, j, k] = H[:, k, j] = y[12 + 6 * np.arange(3) + m]
    v = y[0:3] + y[3:12].reshape(3, 3) @ d + 0.5 * np.einsum("ijk,j,k->i", H, d, d)
    return target_scale(tc, c, r, Q) * v

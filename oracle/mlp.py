"""Numpy restatement of the far-field correction network: input features,
forward pass, loss gradient and Adam step, in float64.

TEST INFRASTRUCTURE ONLY.  This is the checker for the GPU network kernels
(feature extraction, FP32 training forward/backward, BF16 inference): only
``tests/`` import it.  The network has no counterpart in the reference
(xeno14/UVLM sums every wake ring exactly, src/uvlm/advect.h:16-46); its
definition is this module's:

* features of a (target x, block) pair, the block with centre c, radius r and
  vorticity moments Q1 [3][3], Q2 [3][3][3], Q3 [3][3][3][3] (Q_n[i, J] = sum
  over its segments of Gamma e_i int_0^1 (a + t e - c)^(x)n dt):
  d = x - c, D = |d|, dh = d / D, rho = r / D,
  s = ||Q1||_F + ||Q2||_F / r + ||Q3||_F / r^2 (every feature is at most 1),
  f = [dh (3), rho, Q1 / s (9), Q2 . dh / (s r) (9), Q3 : dh dh / (s r^2) (9), 1];
* network: h1 = silu(W1 f + b1), h2 = silu(W2 h1 + b2), h3 = silu(W3 h2 + b3),
  y = W4 h3 + b4 (32 -> 128 -> 128 -> 128 -> 3);
* target: the expansion's remainder in units of its leading omitted term,
  y* = (u_exact - u_expansion) 4 pi D^3 / (s rho^3); the correction added to
  the expansion is y s rho^3 / (4 pi D^3);
* loss: mean over the batch of 1/2 |y - y*|^2.

Parameter vector (FP32 on the device): W1 [128][32], b1, W2 [128][128], b2,
W3 [128][128], b3, W4 [3][128], b4 [3], one pad word.
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT = 32, 128, 3
SHAPES = [("W1", (HID, IN)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (HID, HID)), ("b3", (HID,)), ("W4", (OUT, HID)), ("b4", (OUT,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES) + 1  # + pad


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
    return np.concatenate([w, [0.0]])


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


def features(x, c, r, Q):
    """Features [32] of target x and a block (centre c, radius r, moments Q
    [117]); None when the block carries no circulation."""
    Q = np.asarray(Q, dtype=np.float64)
    Q1 = Q[:9].reshape(3, 3)
    Q2 = Q[9:36].reshape(3, 3, 3)
    Q3 = Q[36:117].reshape(3, 3, 3, 3)
    d = np.asarray(x, dtype=np.float64) - np.asarray(c, dtype=np.float64)
    D = np.linalg.norm(d)
    dh = d / D
    s = scale_of(Q, r)
    if s == 0.0:
        return None
    f = np.zeros(IN)
    f[0:3] = dh
    f[3] = r / D
    f[4:13] = (Q1 / s).reshape(-1)
    f[13:22] = (Q2 @ dh).reshape(-1) / (s * r)
    f[22:31] = ((Q3 @ dh) @ dh).reshape(-1) / (s * r * r)
    f[31] = 1.0
    return f


def scale_of(Q, r):
    """s = ||Q1|| + ||Q2|| / r + ||Q3|| / r^2 of a block of radius r."""
    Q = np.asarray(Q, dtype=np.float64)
    return (np.linalg.norm(Q[:9]) + np.linalg.norm(Q[9:36]) / r
            + np.linalg.norm(Q[36:117]) / (r * r))


def target_scale(x, c, r, Q):
    """s rho^3 / (4 pi D^3): the correction per unit network output."""
    d = np.asarray(x, dtype=np.float64) - np.asarray(c, dtype=np.float64)
    D = np.linalg.norm(d)
    rho = r / D
    return scale_of(Q, r) * rho ** 3 / (4.0 * np.pi * D ** 3)


def expansion(x, c, Q):
    """Velocity of the moment expansion (orders 1..3) of a block at x."""
    Q = np.asarray(Q, dtype=np.float64)
    Q1 = Q[:9].reshape(3, 3)
    Q2 = Q[9:36].reshape(3, 3, 3)
    Q3 = Q[36:117].reshape(3, 3, 3, 3)
    d = np.asarray(x, dtype=np.float64) - np.asarray(c, dtype=np.float64)
    r2 = d @ d
    D = np.sqrt(r2)
    I = np.eye(3)
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
    S = (-np.einsum("cj,bj->cb", Q1, T2) + 0.5 * np.einsum("cjk,bjk->cb", Q2, T3)
         - np.einsum("cjkl,bjkl->cb", Q3, T4) / 6.0)
    u = np.array([S[2, 1] - S[1, 2], S[0, 2] - S[2, 0], S[1, 0] - S[0, 1]])
    return u / (4.0 * np.pi)


def block_moments(nodes, gamma):
    """Centre (node mean), radius and moments [117] of one block: a lattice
    of nodes [R+1][C+1][3] with ring circulations gamma [R][C] (zero
    outside), integrated segment by segment."""
    nodes = np.asarray(nodes, dtype=np.float64)
    R, C = gamma.shape
    c = nodes.reshape(-1, 3).mean(0)
    r = np.linalg.norm(nodes.reshape(-1, 3) - c, axis=1).max()
    G = np.zeros((R + 2, C + 2))
    G[1:-1, 1:-1] = gamma
    Q1, Q2, Q3 = np.zeros((3, 3)), np.zeros((3, 3, 3)), np.zeros((3, 3, 3, 3))

    def seg(a, b, g):
        nonlocal Q1, Q2, Q3
        if g == 0.0:
            return
        A, E = a - c, b - a
        # int_0^1 (A + tE)^(x)n dt, n = 1, 2, 3
        m1 = A + E / 2
        m2 = (np.outer(A, A) + (np.outer(A, E) + np.outer(E, A)) / 2 + np.outer(E, E) / 3)
        AAA = np.einsum("j,k,l->jkl", A, A, A)
        AAE = (np.einsum("j,k,l->jkl", A, A, E) + np.einsum("j,k,l->jkl", A, E, A)
               + np.einsum("j,k,l->jkl", E, A, A))
        AEE = (np.einsum("j,k,l->jkl", A, E, E) + np.einsum("j,k,l->jkl", E, A, E)
               + np.einsum("j,k,l->jkl", E, E, A))
        EEE = np.einsum("j,k,l->jkl", E, E, E)
        m3 = AAA + AAE / 2 + AEE / 3 + EEE / 4
        Q1 += g * np.outer(E, m1)
        Q2 += g * np.einsum("i,jk->ijk", E, m2)
        Q3 += g * np.einsum("i,jkl->ijkl", E, m3)

    for i in range(R + 1):
        for j in range(C):
            seg(nodes[i, j], nodes[i, j + 1], G[i + 1, j + 1] - G[i, j + 1])
    for j in range(C + 1):
        for i in range(R):
            seg(nodes[i, j], nodes[i + 1, j], G[i + 1, j] - G[i + 1, j + 1])
    return c, r, np.concatenate([Q1.reshape(-1), Q2.reshape(-1), Q3.reshape(-1)])

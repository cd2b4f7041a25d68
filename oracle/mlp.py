"""ORACLE — test infrastructure only. FP64 numpy restatement of the warm-start network and
its features (the network has no counterpart in the reference; it is pinned here by central
finite differences and closed-form feature values).

Parameter layout: W1[64][12] b1[64] W2[64][64] b2[64] w3[64] b3.
"""
from __future__ import annotations

import numpy as np

IN, HID = 12, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + HID * IN
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID
NUM_PARAMS = OFF_B3 + 1


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(HID, IN), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(HID, HID),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def softplus(o):
    return np.where(o > 20, o, np.log1p(np.exp(np.minimum(o, 20))))


def forward(p, F):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    F = np.asarray(F, dtype=np.float64)
    h1 = np.maximum(F @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    o = h2 @ w3 + b3
    return softplus(o), (F, h1, h2, o)


def loss(p, F, t, w):
    s, _ = forward(p, F)
    return float(np.sum(np.asarray(w, np.float64) * np.abs(s - np.asarray(t, np.float64))))


def grad(p, F, t, w):
    """Analytic gradient of sum_m w_m |s_m - t_m| (sign(0) = 0)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    s, (F, h1, h2, o) = forward(p, F)
    g = np.sign(s - t) * w / (1 + np.exp(-o))           # dL/do
    dh2 = (g[:, None] * w3[None, :]) * (h2 > 0)
    dh1 = (dh2 @ W2) * (h1 > 0)
    out = np.zeros(NUM_PARAMS)
    out[OFF_W1:OFF_B1] = (dh1.T @ F).ravel()
    out[OFF_B1:OFF_W2] = dh1.sum(0)
    out[OFF_W2:OFF_B2] = (dh2.T @ h1).ravel()
    out[OFF_B2:OFF_W3] = dh2.sum(0)
    out[OFF_W3:OFF_B3] = h2.T @ g
    out[OFF_B3] = g.sum()
    return out


def finite_difference(p, F, t, w, idx, h=1e-6):
    p = np.asarray(p, np.float64).copy()
    out = []
    for i in idx:
        a = p.copy(); a[i] += h
        b = p.copy(); b[i] -= h
        out.append((loss(a, F, t, w) - loss(b, F, t, w)) / (2 * h))
    return np.array(out)


def deg_bucket(d):
    d = np.minimum(np.asarray(d, np.float64), 2 ** 24)
    return np.minimum(192, np.rint(8 * np.log2(1 + d)))


def features(indeg, outdeg, v, alpha, i, k):
    """Features of pairs (i, k) given v = [A^j P for j = 0..4] (dense [N][K] arrays)."""
    n = v[0].shape[0]
    vv = np.stack([vj[i, k] for vj in v], axis=1).astype(np.float64)
    hop = np.full(len(i), 5)
    for j in range(4, -1, -1):
        hop = np.where(vv[:, j] > 0, j, hop)
    F = np.zeros((len(i), IN))
    F[:, 0] = deg_bucket(indeg[i]) / 8
    F[:, 1] = deg_bucket(outdeg[i]) / 8
    F[np.arange(len(i)), 2 + hop] = 1
    c, a = 1 - alpha, alpha
    q1 = c * vv[:, 0]
    q2 = q1 + c * a * vv[:, 1]
    q3 = q2 + c * a * a * vv[:, 2]
    F[:, 8] = np.log1p(n * q1) / 8
    F[:, 9] = np.log1p(n * q2) / 8
    F[:, 10] = np.log1p(n * q3) / 8
    F[:, 11] = 1
    return F, hop

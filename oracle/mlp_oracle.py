"""FP64 numpy restatement of the learned cut-off network — TEST
INFRASTRUCTURE ONLY (checker for the GPU forward/backward kernels of
phantom-grape_amd/csrc/mlp_kernels.h; no counterpart in the reference).

Network: 1-16-16-1, tanh, input x = 2v - 1, parameters p[321] laid out as
w1[16], b1[16], W2[16][16] (row = second-layer unit), b2[16], w3[16], b3.
Loss: sum_s wt_s (y_s - t_s)^2 / n.

Pinned by central finite differences of the loss and of the outputs (the
Jacobian), tests/test_oracle.py, and by the closed-form cut-off shape it is
trained to reproduce.
"""
from __future__ import annotations

import numpy as np

H = 16
NP = 321


def unpack(p):
    p = np.asarray(p, np.float64)
    return p[0:16], p[16:32], p[32:288].reshape(16, 16), p[288:304], p[304:320], p[320]


def forward(p, v):
    w1, b1, W2, b2, w3, b3 = unpack(p)
    x = 2.0 * np.asarray(v, np.float64) - 1.0
    h1 = np.tanh(np.outer(x, w1) + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ w3 + b3


def loss_and_grad(p, v, t, wt):
    w1, b1, W2, b2, w3, b3 = unpack(p)
    v = np.asarray(v, np.float64)
    t = np.asarray(t, np.float64)
    wt = np.asarray(wt, np.float64)
    n = len(v)
    x = 2.0 * v - 1.0
    h1 = np.tanh(np.outer(x, w1) + b1)          # (n, 16)
    h2 = np.tanh(h1 @ W2.T + b2)                # (n, 16)
    y = h2 @ w3 + b3
    r = y - t
    loss = float(np.sum(wt * r * r) / n)
    dy = 2.0 * wt * r / n                       # (n,)
    g = np.zeros(NP)
    g[320] = dy.sum()
    g[304:320] = dy @ h2
    da2 = (dy[:, None] * w3[None, :]) * (1.0 - h2 * h2)  # (n, 16)
    g[288:304] = da2.sum(0)
    g[32:288] = (da2.T @ h1).reshape(-1)
    dh1 = da2 @ W2                              # (n, 16)
    da1 = dh1 * (1.0 - h1 * h1)
    g[16:32] = da1.sum(0)
    g[0:16] = da1.T @ x
    return loss, g


def jacobian(p, v):
    """Outputs y[n] and Jacobian J[n][321] = dy/dp (FP64, backpropagation)."""
    w1, b1, W2, b2, w3, b3 = unpack(p)
    v = np.asarray(v, np.float64)
    n = len(v)
    x = 2.0 * v - 1.0
    h1 = np.tanh(np.outer(x, w1) + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    y = h2 @ w3 + b3
    J = np.zeros((n, NP))
    J[:, 320] = 1.0
    J[:, 304:320] = h2
    da2 = w3[None, :] * (1.0 - h2 * h2)
    J[:, 288:304] = da2
    J[:, 32:288] = (da2[:, :, None] * h1[:, None, :]).reshape(n, 256)
    da1 = (da2 @ W2) * (1.0 - h1 * h1)
    J[:, 16:32] = da1
    J[:, 0:16] = da1 * x[:, None]
    return y, J


def normal_eq(p, v, t, wt):
    """J^T J, J^T r, r^T r of the weighted residuals r = sqrt(wt) (y - t),
    J = dr/dp (what the GPU fit kernels compute)."""
    y, J = jacobian(p, v)
    sw = np.sqrt(np.asarray(wt, np.float64))
    Jw = J * sw[:, None]
    r = sw * (y - np.asarray(t, np.float64))
    return Jw.T @ Jw, Jw.T @ r, float(r @ r)


def s2_unit(u, reff):
    """Target of the S2 form: H(u) = R(u, 1), the S2 shape at softening 1
    (reff = the oracle's restatement of pg5_table.c:8-22)."""
    return np.array([reff(1.0, float(x)) for x in np.atleast_1d(u)])


def target(f, r, rc, r0):
    """Training target of the network: v = sqrt(r/rc), Y = f(r) (r^3 + r0^3)."""
    r = np.asarray(r, np.float64)
    return np.sqrt(r / rc), f(r) * (r ** 3 + r0 ** 3)

"""ORACLE — test infrastructure only. FP64 restatement of the learned coarse
correction network (no reference counterpart): message passing over a coarse
AMG level's graph, forward and backward, on the CPU with numpy.

Graph: the level's CSR operator A (symmetric). Neighbour weights
S_ij = |a_ij| / a_ii for the off-diagonal entries of row i.

Features per node (F = 4): the restricted residual over its RMS, the
diagonal and the off-diagonal l1 sum over the mean diagonal, and 1.

Network (hidden width H, L message-passing layers):
    h0 = relu(x W0 + b0)
    m_l = S h_{l-1};  h_l = relu(h_{l-1} Ws_l + m_l Wn_l + b_l),  l = 1..L
    y = h_L wo + bo
The predicted correction is y * RMS(r) / mean(diag). Loss: mean squared error
of y against the exact correction in the same units.

Parameters are one flat vector, in order: W0 (F x H, row major), b0 (H), then
for each layer Ws (H x H), Wn (H x H), b (H), then wo (H), bo (1).
"""
from __future__ import annotations

import numpy as np

F = 4


def parameter_count(hidden: int, layers: int) -> int:
    return F * hidden + hidden + layers * (2 * hidden * hidden + hidden) + hidden + 1


def unpack(theta, hidden, layers):
    H = hidden
    k = 0

    def take(shape):
        nonlocal k
        size = int(np.prod(shape))
        out = theta[k:k + size].reshape(shape)
        k += size
        return out
    W0, b0 = take((F, H)), take((H,))
    Ws, Wn, b = [], [], []
    for _ in range(layers):
        Ws.append(take((H, H)))
        Wn.append(take((H, H)))
        b.append(take((H,)))
    wo, bo = take((H,)), take((1,))
    assert k == theta.size
    return W0, b0, Ws, Wn, b, wo, bo


def neighbour_matrix(row_ptr, col, val):
    """Dense S (small graphs only): S_ij = |a_ij| / a_ii off the diagonal."""
    n = len(row_ptr) - 1
    S = np.zeros((n, n))
    for i in range(n):
        diag = next(val[j] for j in range(row_ptr[i], row_ptr[i + 1]) if col[j] == i)
        for j in range(row_ptr[i], row_ptr[i + 1]):
            if col[j] != i:
                S[i, col[j]] = abs(val[j]) / diag
    return S


def features(row_ptr, col, val, r):
    """x [n, 4] and the output scale RMS(r) / mean(diag)."""
    n = len(row_ptr) - 1
    diag = np.zeros(n)
    off = np.zeros(n)
    for i in range(n):
        for j in range(row_ptr[i], row_ptr[i + 1]):
            if col[j] == i:
                diag[i] = val[j]
            else:
                off[i] += abs(val[j])
    rms = np.sqrt(np.mean(np.asarray(r, np.float64) ** 2))
    mean_diag = diag.mean()
    x = np.stack([np.asarray(r, np.float64) / rms, diag / mean_diag, off / mean_diag, np.ones(n)], axis=1)
    return x, rms / mean_diag


def forward(theta, S, x, hidden, layers):
    W0, b0, Ws, Wn, b, wo, bo = unpack(theta, hidden, layers)
    h = [np.maximum(x @ W0 + b0, 0.0)]
    m = []
    for l in range(layers):
        m.append(S @ h[-1])
        h.append(np.maximum(h[-1] @ Ws[l] + m[-1] @ Wn[l] + b[l], 0.0))
    y = h[-1] @ wo + bo[0]
    return y, (h, m)


def loss_and_gradient(theta, S, x, target, hidden, layers):
    """Mean squared error and its gradient with respect to theta."""
    W0, b0, Ws, Wn, b, wo, bo = unpack(theta, hidden, layers)
    y, (h, m) = forward(theta, S, x, hidden, layers)
    n = x.shape[0]
    diff = y - target
    loss = float(np.mean(diff ** 2))
    dy = 2.0 * diff / n
    g_wo = h[-1].T @ dy
    g_bo = np.array([dy.sum()])
    dh = np.outer(dy, wo)
    g_Ws, g_Wn, g_b = [None] * layers, [None] * layers, [None] * layers
    for l in reversed(range(layers)):
        dz = dh * (h[l + 1] > 0)
        g_Ws[l] = h[l].T @ dz
        g_Wn[l] = m[l].T @ dz
        g_b[l] = dz.sum(axis=0)
        dm = dz @ Wn[l].T
        dh = dz @ Ws[l].T + S.T @ dm
    dz0 = dh * (h[0] > 0)
    g_W0 = x.T @ dz0
    g_b0 = dz0.sum(axis=0)
    parts = [g_W0.ravel(), g_b0]
    for l in range(layers):
        parts += [g_Ws[l].ravel(), g_Wn[l].ravel(), g_b[l]]
    parts += [g_wo, g_bo]
    return loss, np.concatenate(parts)


def initial_parameters(hidden: int, layers: int, seed: int = 0) -> np.ndarray:
    """He-style random weights, zero biases (the product's initialisation)."""
    rng = np.random.default_rng(seed)
    H = hidden
    parts = [rng.standard_normal((F, H)).ravel() * np.sqrt(2.0 / F), np.zeros(H)]
    for _ in range(layers):
        parts += [rng.standard_normal((H, H)).ravel() * np.sqrt(1.0 / H),
                  rng.standard_normal((H, H)).ravel() * np.sqrt(1.0 / H), np.zeros(H)]
    parts += [rng.standard_normal(H) * np.sqrt(1.0 / H), np.zeros(1)]
    return np.concatenate(parts)

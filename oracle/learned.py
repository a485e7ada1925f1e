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

With bf16=True, forward and loss_and_gradient round to bfloat16 (round to
nearest even, via float32) at the points where the device kernels store
bf16: the hidden states h_l and aggregates m_l after they are computed, the
layer weights Ws and Wn as multiplied, and the backpropagated dz_l as
multiplied into the layer and weight gradients; everything else stays FP64.
The device result then differs from it only by FP32 accumulation.

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


def neighbour_operator(row_ptr, col, val):
    """S as a scipy.sparse CSR matrix (large graphs), same entries as
    neighbour_matrix."""
    import scipy.sparse as sp
    row_ptr = np.asarray(row_ptr)
    col = np.asarray(col)
    val = np.asarray(val, np.float64)
    n = len(row_ptr) - 1
    rows = np.repeat(np.arange(n), np.diff(row_ptr))
    diag = np.zeros(n)
    on = col == rows
    diag[rows[on]] = val[on]
    off = ~on
    return sp.csr_matrix((np.abs(val[off]) / diag[rows[off]], (rows[off], col[off])), shape=(n, n))


def features(row_ptr, col, val, r):
    """x [n, 4] and the output scale RMS(r) / mean(diag)."""
    n = len(row_ptr) - 1
    rows = np.repeat(np.arange(n), np.diff(np.asarray(row_ptr)))
    col = np.asarray(col)
    val = np.asarray(val, np.float64)
    on = col == rows
    diag = np.zeros(n)
    diag[rows[on]] = val[on]
    off = np.bincount(rows[~on], weights=np.abs(val[~on]), minlength=n)
    rms = np.sqrt(np.mean(np.asarray(r, np.float64) ** 2))
    mean_diag = diag.mean()
    x = np.stack([np.asarray(r, np.float64) / rms, diag / mean_diag, off / mean_diag, np.ones(n)], axis=1)
    return x, rms / mean_diag


def bf16(a):
    """Round to bfloat16 (nearest even) through float32; returned as float64."""
    f = np.asarray(a, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(theta, S, x, hidden, layers, bf16_storage=False):
    W0, b0, Ws, Wn, b, wo, bo = unpack(theta, hidden, layers)
    q = bf16 if bf16_storage else (lambda v: v)
    h = [q(np.maximum(x @ W0 + b0, 0.0))]
    m = []
    for l in range(layers):
        m.append(q(S @ h[-1]))
        h.append(q(np.maximum(h[-1] @ q(Ws[l]) + m[-1] @ q(Wn[l]) + b[l], 0.0)))
    y = h[-1] @ wo + bo[0]
    return y, (h, m)


def loss_and_gradient(theta, S, x, target, hidden, layers, bf16_storage=False):
    """Mean squared error and its gradient with respect to theta."""
    W0, b0, Ws, Wn, b, wo, bo = unpack(theta, hidden, layers)
    q = bf16 if bf16_storage else (lambda v: v)
    y, (h, m) = forward(theta, S, x, hidden, layers, bf16_storage)
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
        g_b[l] = dz.sum(axis=0)
        dz = q(dz)
        g_Ws[l] = h[l].T @ dz
        g_Wn[l] = m[l].T @ dz
        dm = dz @ q(Wn[l]).T
        dh = dz @ q(Ws[l]).T + S.T @ dm
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

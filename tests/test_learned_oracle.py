"""The FP64 restatement of the learned coarse correction network
(oracle/learned.py) against finite differences, and its building blocks
against closed forms."""
import numpy as np
import pytest

import oracle
from oracle import learned


def small_graph(n=30, seed=0):
    """A symmetric M-matrix on a random sparse pattern (diagonally dominant)."""
    rng = np.random.default_rng(seed)
    A = np.zeros((n, n))
    for i in range(n):
        for j in rng.choice(n, 3, replace=False):
            if j != i:
                w = rng.uniform(0.5, 2.0)
                A[i, j] = A[j, i] = -w
    for i in range(n):
        A[i, i] = -A[i].sum() + rng.uniform(0.1, 1.0)
    row_ptr, col, val = [0], [], []
    for i in range(n):
        for j in range(n):
            if A[i, j] != 0:
                col.append(j)
                val.append(A[i, j])
        row_ptr.append(len(col))
    return np.array(row_ptr), np.array(col), np.array(val), A


def test_neighbour_matrix_and_features_closed_form():
    row_ptr, col, val, A = small_graph()
    S = learned.neighbour_matrix(row_ptr, col, val)
    d = np.diag(A)
    expected = np.abs(A - np.diag(d)) / d[:, None]
    assert np.allclose(S, expected)
    assert np.array_equal(learned.neighbour_operator(row_ptr, col, val).toarray(), S)
    r = np.random.default_rng(1).standard_normal(len(d))
    x, scale = learned.features(row_ptr, col, val, r)
    assert np.allclose(x[:, 0] * np.sqrt(np.mean(r ** 2)), r)
    assert np.allclose(x[:, 1] * d.mean(), d)
    assert np.allclose(x[:, 2] * d.mean(), np.abs(A - np.diag(d)).sum(axis=1))
    assert np.all(x[:, 3] == 1) and scale == pytest.approx(np.sqrt(np.mean(r ** 2)) / d.mean())


@pytest.mark.parametrize("hidden,layers", [(8, 3), (5, 1), (6, 2)])
def test_gradient_matches_finite_differences(hidden, layers):
    """Central differences (h = 1e-6) of the FP64 loss agree with the
    backward pass for every parameter to 1e-6 relative: a dropped transpose,
    a missing S^T term or a wrong relu mask fails."""
    row_ptr, col, val, A = small_graph(24, seed=hidden + layers)
    S = learned.neighbour_matrix(row_ptr, col, val)
    rng = np.random.default_rng(7)
    x = rng.standard_normal((24, learned.F))
    t = rng.standard_normal(24)
    theta = learned.initial_parameters(hidden, layers, seed=3) + 0.05 * rng.standard_normal(
        learned.parameter_count(hidden, layers))
    loss, grad = learned.loss_and_gradient(theta, S, x, t, hidden, layers)
    assert grad.size == learned.parameter_count(hidden, layers)
    eps = 1e-6
    numeric = np.empty_like(theta)
    for k in range(theta.size):
        e = np.zeros_like(theta)
        e[k] = eps
        numeric[k] = (learned.loss_and_gradient(theta + e, S, x, t, hidden, layers)[0]
                      - learned.loss_and_gradient(theta - e, S, x, t, hidden, layers)[0]) / (2 * eps)
    assert np.abs(numeric - grad).max() <= 1e-6 * max(1.0, np.abs(grad).max())


def test_forward_is_linear_in_the_output_layer():
    """y depends affinely on (wo, bo): doubling wo and bo doubles y."""
    row_ptr, col, val, A = small_graph(20, seed=4)
    S = learned.neighbour_matrix(row_ptr, col, val)
    x = np.random.default_rng(2).standard_normal((20, learned.F))
    theta = learned.initial_parameters(6, 2, seed=1)
    y1, _ = learned.forward(theta, S, x, 6, 2)
    theta2 = theta.copy()
    theta2[-7:] *= 2.0  # wo (6) and bo (1)
    y2, _ = learned.forward(theta2, S, x, 6, 2)
    assert np.allclose(y2, 2 * y1)


def test_library_parameter_count_matches_oracle():
    """The C ABI's parameter layout size (no GPU needed) is the oracle's."""
    import kutta_amd
    for layers in (1, 2, 3, 4):
        assert kutta_amd.LearnedNetwork.parameter_count(layers) == learned.parameter_count(64, layers)
    assert kutta_amd.LearnedNetwork.parameter_count(0) == -1
    assert kutta_amd.LearnedNetwork.parameter_count(5) == -1


def test_bf16_rounding_closed_forms():
    """Round to nearest even at 8 significant bits, through float32."""
    one = 1.0
    ulp = 2.0 ** -7  # bf16 spacing just above 1
    assert learned.bf16(one + ulp / 2) == one            # tie: to even (1)
    assert learned.bf16(one + 3 * ulp / 2) == one + 2 * ulp  # tie: to even (1 + 2 ulp)
    assert learned.bf16(one + 0.51 * ulp) == one + ulp
    assert learned.bf16(-(one + 0.49 * ulp)) == -one
    assert learned.bf16(3.0) == 3.0 and learned.bf16(0.0) == 0.0
    x = np.random.default_rng(2).standard_normal(1000)
    r = learned.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8)
    assert np.array_equal(learned.bf16(r), r)  # idempotent


def test_bf16_storage_mode_is_a_small_perturbation():
    """The bf16-storage variant stays within bf16 error of the FP64 network
    (whose gradient the finite-difference tests pin)."""
    row_ptr, col, val, A = small_graph(40, seed=4)
    H, L = 16, 2
    S = learned.neighbour_matrix(row_ptr, col, val)
    r = np.random.default_rng(5).standard_normal(len(row_ptr) - 1)
    x, _ = learned.features(row_ptr, col, val, r)
    theta = learned.initial_parameters(H, L, seed=6)
    y64, _ = learned.forward(theta, S, x, H, L)
    y16, _ = learned.forward(theta, S, x, H, L, bf16_storage=True)
    assert np.linalg.norm(y16 - y64) / np.linalg.norm(y64) < 3e-2
    t = np.random.default_rng(7).standard_normal(len(r))
    l64, g64 = learned.loss_and_gradient(theta, S, x, t, H, L)
    l16, g16 = learned.loss_and_gradient(theta, S, x, t, H, L, bf16_storage=True)
    assert abs(l16 - l64) / l64 < 3e-2
    assert np.linalg.norm(g16 - g64) / np.linalg.norm(g64) < 0.2

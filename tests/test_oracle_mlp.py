"""Pins for the oracle MLP of the learned ALS warm start: backprop against
central finite differences of the loss, and the parameter layout against a
direct evaluation."""
import numpy as np

from oracle import peps_oracle as O


def test_mlp_backprop_matches_finite_differences():
    rng = np.random.default_rng(0)
    n_in, w, n_out, B = 3, 5, 2, 4
    n = w * n_in + w + w * w + w + n_out * w + n_out
    p = rng.standard_normal(n) * 0.7
    X = rng.standard_normal((n_in, B))
    T = rng.standard_normal((n_out, B))
    loss, g = O.mlp_loss_grads(p, X, T, n_in, w, n_out)
    h = 1e-6
    fd = np.zeros(n)
    for i in range(n):
        e = np.zeros(n)
        e[i] = h
        lp = O.mlp_loss_grads(p + e, X, T, n_in, w, n_out)[0]
        lm = O.mlp_loss_grads(p - e, X, T, n_in, w, n_out)[0]
        fd[i] = (lp - lm) / (2 * h)
    assert np.abs(fd - g).max() < 1e-7 * max(1.0, np.abs(g).max())
    Y = O.mlp_forward(p, X, n_in, w, n_out)
    assert abs(loss - 0.5 * np.sum((Y - T) ** 2) / B) < 1e-14


def test_mlp_parameter_layout():
    """W1[i, j] is p[i + j * width] (column-major) and b1 follows W1."""
    n_in, w, n_out = 2, 3, 1
    n = w * n_in + w + w * w + w + n_out * w + n_out
    p = np.zeros(n)
    p[1 + 1 * w] = 2.0          # W1[1, 1]
    p[w * n_in + 1] = 0.5       # b1[1]
    off = w * n_in + w
    p[off + 1 + 1 * w] = 1.0    # W2[1, 1]
    off += w * w + w
    p[off + 1] = 3.0            # W3[0, 1]
    X = np.array([[0.0], [1.0]])
    # H1[1] = 2*1 + 0.5 = 2.5; H2[1] = 2.5; Y = 3 * 2.5
    assert abs(O.mlp_forward(p, X, n_in, w, n_out)[0, 0] - 7.5) < 1e-15

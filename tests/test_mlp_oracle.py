"""Pins for oracle/mlp_oracle.py: central finite differences (float64) of the
loss reproduce the analytic gradient; a plausible mistake (missing ReLU mask,
transposed weight, dropped bias) fails them.  Plus host-side cost-model
features."""
import numpy as np

from oracle import mlp_oracle as M
from picoquant_hip import costmodel as CM


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    d, h, n = 5, 7, 13
    p = rng.standard_normal(M.n_params(d, h)) * 0.5
    X = rng.standard_normal((n, d))
    y = rng.standard_normal(n)
    loss, g = M.loss_grad(p, X, y, d, h)
    eps = 1e-6
    num = np.zeros_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = eps
        num[i] = (M.loss_grad(p + e, X, y, d, h)[0] - M.loss_grad(p - e, X, y, d, h)[0]) / (2 * eps)
    np.testing.assert_allclose(g, num, rtol=1e-5, atol=1e-8)


def test_forward_closed_form():
    """With identity-like weights the network computes a known value."""
    d, h = 2, 2
    W1 = np.eye(2)
    b1 = np.array([0.0, -10.0])          # second unit always off
    W2 = np.eye(2)
    b2 = np.zeros(2)
    w3 = np.array([3.0, 100.0])
    b3 = 0.5
    p = np.concatenate([W1.ravel(), b1, W2.ravel(), b2, w3, [b3]])
    y, _ = M.forward(p, np.array([[2.0, 1.0], [-1.0, 4.0]]), d, h)
    np.testing.assert_allclose(y, [3 * 2 + 0.5, 0.5])


def test_features_and_merge():
    assert CM.merge([(2, 1, 4), (2, 2, 8), (1, 9, 9), (4, 16, 1)]) == [(4, 1, 4), (4, 16, 1)]
    f = CM.step_features([(4, 1, 4)], [(8, 1, 1)], [(16, 4, 8)], batch=2)
    assert len(f) == CM.N_FEATURES
    assert f[:4] == [2.0, 3.0, 4.0, 1.0]
    assert f[4] == 10.0

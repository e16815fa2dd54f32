"""CPU oracle for the learned cost model's perceptron (TEST INFRASTRUCTURE ONLY:
only tests/ may import it).

The reference has no learned model; this restates the textbook 3-layer
perceptron (ReLU hidden layers, linear output, mean squared error) in numpy,
float64, with the same flat parameter layout as pq_mlp_* in include/pq_hip.h:
    W1[in][H] | b1[H] | W2[H][H] | b2[H] | w3[H] | b3
Pinned by central finite differences (tests/test_mlp_oracle.py).
"""
import numpy as np


def unpack(p, d_in, h):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + d_in * h].reshape(d_in, h); o += d_in * h
    b1 = p[o:o + h]; o += h
    W2 = p[o:o + h * h].reshape(h, h); o += h * h
    b2 = p[o:o + h]; o += h
    w3 = p[o:o + h]; o += h
    b3 = p[o]; o += 1
    assert o == p.size
    return W1, b1, W2, b2, w3, b3


def n_params(d_in, h):
    return d_in * h + h + h * h + h + h + 1


def forward(p, X, d_in, h):
    W1, b1, W2, b2, w3, b3 = unpack(p, d_in, h)
    X = np.asarray(X, dtype=np.float64)
    z1 = X @ W1 + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2 + b2
    h2 = np.maximum(z2, 0)
    y = h2 @ w3 + b3
    return y, (X, z1, h1, z2, h2)


def loss_grad(p, X, y, d_in, h):
    """Mean squared error and its gradient in the flat layout."""
    W1, b1, W2, b2, w3, b3 = unpack(p, d_in, h)
    out, (X, z1, h1, z2, h2) = forward(p, X, d_in, h)
    n = X.shape[0]
    d = out - np.asarray(y, dtype=np.float64)
    loss = float(np.mean(d ** 2))
    dy = 2 * d / n
    g_w3 = h2.T @ dy
    g_b3 = dy.sum()
    dh2 = np.outer(dy, w3) * (z2 > 0)
    g_W2 = h1.T @ dh2
    g_b2 = dh2.sum(0)
    dh1 = (dh2 @ W2.T) * (z1 > 0)
    g_W1 = X.T @ dh1
    g_b1 = dh1.sum(0)
    g = np.concatenate([g_W1.ravel(), g_b1, g_W2.ravel(), g_b2, g_w3, [g_b3]])
    return loss, g

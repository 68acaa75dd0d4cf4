"""CPU oracle of the quotient-range cost model (5 -> 32 -> 32 -> 1, ReLU).

TEST INFRASTRUCTURE ONLY: checks the GPU forward, backward and feature code
of the engine (deleglise-rivat_amd/csrc/mdr_mlp.hip, mdr_mlp.hpp).  The
network has no counterpart in the reference; its parameter layout is
W1[32][5], b1[32], W2[32][32], b2[32], W3[32], b3 and the loss is the mean
squared error over the batch.
"""
import numpy as np

IN, H = 5, 32
W1, B1 = 0, H * IN
W2 = B1 + H
B2 = W2 + H * H
W3 = B2 + H
B3 = W3 + H
NPARAM = B3 + 1


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    return (w[W1:B1].reshape(H, IN), w[B1:W2], w[W2:B2].reshape(H, H), w[B2:W3], w[W3:B3], w[B3])


def forward(w, X):
    w1, b1, w2, b2, w3, b3 = unpack(w)
    X = np.asarray(X, dtype=np.float64)
    z1 = X @ w1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ w2.T + b2
    h2 = np.maximum(z2, 0)
    return h2 @ w3 + b3


def loss_and_grad(w, X, T):
    """Mean squared error and its gradient (backpropagation in float64)."""
    w1, b1, w2, b2, w3, b3 = unpack(w)
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    n = len(X)
    z1 = X @ w1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ w2.T + b2
    h2 = np.maximum(z2, 0)
    out = h2 @ w3 + b3
    diff = out - T
    g = 2 * diff / n
    d2 = np.outer(g, w3) * (z2 > 0)
    d1 = (d2 @ w2) * (z1 > 0)
    grad = np.concatenate([(d1.T @ X).ravel(), d1.sum(0), (d2.T @ h1).ravel(), d2.sum(0), h2.T @ g, [g.sum()]])
    return float(np.mean(diff * diff)), grad


def features(y, start, end, kind):
    """log2(y)/64, log2(start)/64, log2(end)/64, kind/16, log2(y/start)/64, in float32."""
    y = np.asarray(y, dtype=np.float64).astype(np.float32)
    s = np.asarray(start, dtype=np.float64).astype(np.float32)
    e = np.asarray(end, dtype=np.float64).astype(np.float32)
    ly, ls = np.log2(y), np.log2(s)
    return np.stack([ly / 64, ls / 64, np.log2(e) / 64, np.asarray(kind, dtype=np.float32) / 16, (ly - ls) / 64],
                    axis=1).astype(np.float32)

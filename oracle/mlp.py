"""ORACLE — test infrastructure only.

numpy restatement of the rank model that libdmhm_amd.so trains and evaluates
on the GPU (no counterpart in the reference, whose sketch width is the fixed
SampleRank(maxRank) = maxRank + Oversample(), include/dmhm/hmat2d.hpp:32-33):
a perceptron with nf inputs, two hidden layers of 64 ReLU units and one linear
output, trained on the mean squared error. Parameters are one float vector in
the library's layout: W1 [nf x 64], b1 [64], W2 [64 x 64], b2 [64], w3 [64],
b3 [1] (row-major, row = input unit). Computed in float64 here; the tests pin
the gradient by central finite differences and compare the GPU's FP32
forward and backward passes with it.
"""
import numpy as np

H = 64


def num_params(nf):
    return nf * H + H + H * H + H + H + 1


def unpack(p, nf):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + nf * H].reshape(nf, H); o += nf * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    w3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def forward(p, X):
    """Model outputs for the rows of X (n x nf)."""
    X = np.asarray(X, dtype=np.float64)
    W1, b1, W2, b2, w3, b3 = unpack(p, X.shape[1])
    h1 = np.maximum(X @ W1 + b1, 0.0)
    h2 = np.maximum(h1 @ W2 + b2, 0.0)
    return h2 @ w3 + b3


def loss_and_grad(p, X, t):
    """L = mean((y - t)^2) and dL/dp in the parameter layout."""
    X = np.asarray(X, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    n, nf = X.shape
    W1, b1, W2, b2, w3, b3 = unpack(p, nf)
    a1 = X @ W1 + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0.0)
    y = h2 @ w3 + b3
    r = y - t
    L = float(np.mean(r * r))
    dy = 2.0 * r / n
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    g2 = np.outer(dy, w3) * (a2 > 0)
    gW2 = h1.T @ g2
    gb2 = g2.sum(axis=0)
    g1 = (g2 @ W2.T) * (a1 > 0)
    gW1 = X.T @ g1
    gb1 = g1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return L, g


def adam_train(p, X, t, steps, batch, lr, perm):
    """The library's training loop restated: Adam (beta1 0.9, beta2 0.999,
    eps 1e-8) on consecutive mini-batches of the permuted samples."""
    p = np.array(p, dtype=np.float64)
    X = np.asarray(X)[perm]
    t = np.asarray(t)[perm]
    n = len(t)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    b = min(batch, n)
    losses = []
    for st in range(steps):
        off = (st * b) % n
        nb = min(b, n - off)
        L, g = loss_and_grad(p, X[off:off + nb], t[off:off + nb])
        losses.append(L)
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        c1 = 1 - 0.9 ** (st + 1)
        c2 = 1 - 0.999 ** (st + 1)
        p -= lr * (m / c1) / (np.sqrt(v / c2) + 1e-8)
    return p, losses

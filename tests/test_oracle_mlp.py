"""Pins the rank model's CPU oracle (oracle/mlp.py): its gradient against
central finite differences of its loss, and the shape of its parameter
layout. CPU only."""
import numpy as np

from oracle import mlp


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(0)
    nf, n = 12, 40
    p = rng.standard_normal(mlp.num_params(nf)) * 0.3
    X = rng.standard_normal((n, nf))
    t = rng.uniform(0, 16, n)
    L, g = mlp.loss_and_grad(p, X, t)
    assert g.shape == p.shape
    h = 1e-6
    idx = rng.choice(p.size, 200, replace=False)
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mlp.loss_and_grad(p + e, X, t)[0] - mlp.loss_and_grad(p - e, X, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), i
    # the output bias gradient has a closed form: mean of 2 (y - t)
    y = mlp.forward(p, X)
    assert abs(g[-1] - np.mean(2 * (y - t))) < 1e-12


def test_layout_matches_forward():
    rng = np.random.default_rng(1)
    nf = 3
    p = rng.standard_normal(mlp.num_params(nf))
    W1, b1, W2, b2, w3, b3 = mlp.unpack(p, nf)
    x = rng.standard_normal(nf)
    h1 = np.maximum(x @ W1 + b1, 0)
    h2 = np.maximum(h1 @ W2 + b2, 0)
    assert abs(mlp.forward(p, x[None, :])[0] - (h2 @ w3 + b3)) < 1e-12
    assert mlp.num_params(12) == 12 * 64 + 64 + 64 * 64 + 64 + 64 + 1


def test_adam_training_reduces_loss_on_a_learnable_target():
    rng = np.random.default_rng(2)
    nf, n = 12, 512
    X = rng.uniform(-1, 1, (n, nf))
    t = 8 + 4 * X[:, 0] - 3 * X[:, 3] + 2 * np.abs(X[:, 7])
    p0 = rng.standard_normal(mlp.num_params(nf)) * 0.1
    p1, losses = mlp.adam_train(p0, X, t, steps=300, batch=128, lr=3e-3, perm=rng.permutation(n))
    assert losses[-1] < 0.05 * losses[0]

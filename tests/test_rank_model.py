"""The learned rank model on the GPU: its forward pass, gradient and Adam
step against the numpy oracle (oracle/mlp.py, itself pinned by finite
differences); training on ranks revealed by the plain adaptive (tolerance
mode) compression; and width-indirect products driven by it, against the
full-width products and dense arithmetic, including the a-posteriori
fallback. Tolerances: FP32 forward/gradient 1e-4 relative; products 1e-9
relative to the full-width tolerance-mode product (both truncate at
relTol 1e-10)."""
import numpy as np
import pytest

import dmhm_amd
from dmhm_amd import workloads as W
from oracle import mlp

pytestmark = pytest.mark.gpu

NF = 12


@pytest.fixture(scope="module", autouse=True)
def _init():
    dmhm_amd.init(0)
    yield


def _rel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-30))


def test_forward_and_gradient_match_oracle():
    m = dmhm_amd.RankModel(seed=3)
    p = m.params.astype(np.float64)
    rng = np.random.default_rng(0)
    for n in (1, 63, 64, 300):           # partial and several workgroups
        X = rng.standard_normal((n, NF)).astype(np.float32)
        t = rng.uniform(0, 16, n).astype(np.float32)
        yr = mlp.forward(p, X)
        assert _rel(m.predict(X), yr) < 1e-4
        L, g = m.gradient(X, t)
        Lr, gr = mlp.loss_and_grad(p, X, t)
        assert abs(L - Lr) <= 1e-4 * Lr
        assert _rel(g, gr) < 1e-4


def test_one_adam_step_matches_oracle():
    """A full-batch step is independent of the sample order."""
    m = dmhm_amd.RankModel(seed=5)
    p0 = m.params.astype(np.float64)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((200, NF)).astype(np.float32)
    t = rng.uniform(0, 16, 200).astype(np.float32)
    m.train(X, t, steps=1, batch=200, lr=1e-3, seed=0)
    p1r, _ = mlp.adam_train(p0, X, t, steps=1, batch=200, lr=1e-3, perm=np.arange(200))
    # Adam's first step moves each parameter by about lr sign(g).
    assert np.abs(m.params - p1r).max() < 1e-5


def _tol_product(A, B, model=None, margin=1.0, seed=11):
    return A.Multiply(1.0, B, oversample=4, tol_mode=1, rel_tol=1e-10, seed=seed, rank_model=model,
                      rank_margin=margin)


def _operands(n=32, levels=3, rank=16):
    """An operator and the first iterate-like matrix X = A^T A^T A (moderate
    ranks in its admissible blocks), tolerance mode."""
    op = W.small_operator(n, levels, smooth=False)
    args = (op["rowptr"], op["colind"], op["vals"])
    A = dmhm_amd.HMat2d.from_csr(op["sizes"], levels, rank, True, *args)
    At = A.Transpose()
    X = _tol_product(_tol_product(At, At), At)
    return op, A, X


def test_training_on_revealed_ranks_and_learned_width_product():
    op, A, X = _operands()
    # Training pairs: features of products with left operand X and Y, and the
    # ranks the full-width adaptive compression revealed in their results.
    feats, ranks = [], []
    Y = X
    for it in range(4):
        F = Y.rank_features()
        C = _tol_product(Y, A, seed=20 + it)
        feats.append(F)
        ranks.append(C.ranks().astype(np.float32))
        Y = C
    F = np.concatenate(feats)
    t = np.concatenate(ranks)
    m = dmhm_amd.RankModel(seed=7)
    losses = m.train(F, t, steps=400, batch=256, lr=3e-3, seed=1)
    assert losses[-20:].mean() < 0.2 * losses[:20].mean()
    # Learned-width product vs the full-width one: same truncation rule.
    full = _tol_product(X, A)
    learned = _tol_product(X, A, model=m, margin=1.0)
    st = dmhm_amd.last_product_stats()
    assert st["blocks"] == len(X.ranks())
    assert st["width_sum"] < st["full_width_sum"]
    D = W.csr_to_dense(op["rowptr"], op["colind"], op["vals"], 1024)
    exact = X.to_dense() @ D
    err_full = np.linalg.norm(full.to_dense() - exact) / np.linalg.norm(exact)
    err_learned = np.linalg.norm(learned.to_dense() - exact) / np.linalg.norm(exact)
    assert err_learned <= 10 * err_full + 1e-12
    assert _rel(learned.to_dense(), full.to_dense()) < 1e-8


def test_fallback_recomputes_undersized_sketches():
    """An untrained model with a large negative margin predicts the narrowest
    sketch everywhere; blocks whose SVD keeps more than width - oversample
    values are recomputed at full width, so the result matches the full-width
    product."""
    op, A, X = _operands()
    m = dmhm_amd.RankModel(seed=9)
    full = _tol_product(X, A)
    narrow = _tol_product(X, A, model=m, margin=-100.0)
    st = dmhm_amd.last_product_stats()
    assert st["fallbacks"] > 0
    assert st["width_sum"] == 5 * st["blocks"]          # oversample + 1 everywhere
    assert _rel(narrow.to_dense(), full.to_dense()) < 1e-8

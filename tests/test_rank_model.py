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


def _operator(n=32, levels=3, rank=16):
    op = W.small_operator(n, levels, smooth=False)
    args = (op["rowptr"], op["colind"], op["vals"])
    return op, dmhm_amd.HMat2d.from_csr(op["sizes"], levels, rank, True, *args)


def _lowrank(op, r, seed, levels=3, rank=16):
    """X = U V^T imported exactly (ImportLowRank): rank r in every admissible
    block, so X A has admissible blocks of rank at most r."""
    n = int(np.prod(op["sizes"]))
    rng = np.random.default_rng(seed)
    return dmhm_amd.HMat2d.from_lowrank(op["sizes"], levels, rank, True, rng.standard_normal((n, r)),
                                        rng.standard_normal((n, r)))


def test_training_on_revealed_ranks_and_learned_width_product():
    op, A = _operator()
    # Training pairs: features of products X_r A (left operand X_r, whose
    # blocks have rank r) and the ranks the full-width adaptive compression
    # revealed in their results.
    feats, ranks = [], []
    for r, seed in ((2, 1), (5, 2), (9, 3), (3, 4), (7, 5)):
        X = _lowrank(op, r, seed)
        feats.append(X.rank_features())
        ranks.append(_tol_product(X, A, seed=20 + r).ranks().astype(np.float32))
    F = np.concatenate(feats)
    t = np.concatenate(ranks)
    assert 0 < t.max() <= 9
    m = dmhm_amd.RankModel(seed=7)
    losses = m.train(F, t, steps=600, batch=256, lr=3e-3, seed=1)
    assert losses[-20:].mean() < 0.1 * losses[:20].mean()
    # Learned-width product vs the full-width one on a fresh operand.
    X = _lowrank(op, 5, 11)
    full = _tol_product(X, A)
    learned = _tol_product(X, A, model=m, margin=1.0)
    st = dmhm_amd.last_product_stats()
    assert st["blocks"] == len(X.ranks())
    assert st["width_sum"] < 0.7 * st["full_width_sum"]
    D = W.csr_to_dense(op["rowptr"], op["colind"], op["vals"], 1024)
    exact = X.to_dense() @ D
    assert _rel(learned.to_dense(), exact) < 1e-9
    assert _rel(learned.to_dense(), full.to_dense()) < 1e-9
    np.testing.assert_array_equal(learned.ranks(), full.ranks())


def test_fallback_recomputes_undersized_sketches():
    """An untrained model with a large negative margin predicts the narrowest
    sketch (oversample + 1 columns) everywhere; blocks whose SVD keeps more
    than width - oversample values are recomputed at full width, so the result
    matches the full-width product."""
    op, A = _operator()
    X = _lowrank(op, 9, 13)
    m = dmhm_amd.RankModel(seed=9)
    full = _tol_product(X, A)
    assert (full.ranks() >= 2).sum() > 0
    narrow = _tol_product(X, A, model=m, margin=-100.0)
    st = dmhm_amd.last_product_stats()
    assert st["fallbacks"] == int((full.ranks() >= 2).sum())
    assert st["width_sum"] == 5 * st["blocks"]
    assert _rel(narrow.to_dense(), full.to_dense()) < 1e-9

"""Batches of independent H-matrices (one launch per phase for all members):
each member's batched result against the same operation on that member
alone (same seeds; tile engine forced so both sum in the same order), and
every batched inverse against its dense residual ||I - X A||_2. The
reference inverts such operators one by one (HMat3d::SchulzInvert,
src/HMat3d/Invert-incl.hpp:127-188). Tolerances: products and estimates 1e-12
relative to the member's largest entry; inverses 1e-9; dense residuals at
most 1.5x the lone inversion's and below 1e-2 (rank-limited)."""
import numpy as np
import pytest

import dmhm_amd
from dmhm_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    dmhm_amd.init(0)
    prev = dmhm_amd.set_tile_engine("lds")
    yield
    dmhm_amd.set_tile_engine(prev)


# (n per axis, levels, dim, maxRank, coefficient seed)
MEMBERS = [(16, 3, 2, 8, 0), (12, 3, 2, 8, 1), (8, 3, 3, 8, 2), (16, 3, 2, 8, 3)]
INVERT = [(8, 3, 3, 16, 0), (12, 3, 3, 16, 1), (16, 4, 3, 32, 2), (16, 3, 2, 8, 3)]


def _member(n, L, dim, rank, seed):
    op = W.small_operator(n, L, dim=dim, smooth=False, seed=seed)
    cls = dmhm_amd.HMat2d if dim == 2 else dmhm_amd.HMat3d
    H = cls.from_csr(op["sizes"], L, rank, True, op["rowptr"], op["colind"], op["vals"])
    D = W.csr_to_dense(op["rowptr"], op["colind"], op["vals"], int(np.prod(op["sizes"])))
    return H, D


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_batched_product_matches_members():
    hs = [_member(*m)[0] for m in MEMBERS]
    batch = dmhm_amd.HBatch.create(hs)
    C = batch.Multiply(1.0, batch, seed=5)
    assert len(C) == len(hs)
    for i, H in enumerate(hs):
        print("member", i, MEMBERS[i], flush=True)
        ref = H.Multiply(1.0, H, seed=5)
        got = C.member(i)
        assert _rel(got.to_dense(), ref.to_dense()) < 1e-12
        np.testing.assert_array_equal(got.ranks(), ref.ranks())


def test_batched_two_norm_matches_members():
    hs = [_member(*m)[0] for m in MEMBERS]
    est = dmhm_amd.HBatch.create(hs).EstimateTwoNorm(seed=9)
    for i, H in enumerate(hs):
        ref = dmhm_amd.EstimateTwoNorm(H, seed=9)
        assert abs(est[i] - ref) <= 1e-12 * ref


@pytest.mark.parametrize("graph", [False, True])
def test_batched_schulz_matches_members(graph):
    data = [_member(*m) for m in INVERT]
    batch = dmhm_amd.HBatch.create([h for h, _ in data])
    opts = dict(stopTol=1e-9, maxIterations=60, stall_stop=True, graph=graph)
    reps = batch.SchulzInvert(**opts)
    for i, (H, D) in enumerate(data):
        print("member", i, INVERT[i], "iterations", reps[i].iterations, reps[i].residualEstimate, flush=True)
        lone = H.Copy()
        rep = lone.SchulzInvert(**opts)
        assert reps[i].iterations == rep.iterations
        X = batch.member(i).to_dense()
        Xl = lone.to_dense()
        assert _rel(X, Xl) < 1e-9
        n = D.shape[0]
        r = np.linalg.norm(np.eye(n) - X @ D, 2)
        rl = np.linalg.norm(np.eye(n) - Xl @ D, 2)
        print("   dense residual", r, "lone", rl, flush=True)
        assert r <= 1.5 * rl + 1e-12
        assert r < 1e-2


def test_batched_schulz_fixed_iterations():
    data = [_member(*m) for m in MEMBERS[:3]]
    batch = dmhm_amd.HBatch.create([h for h, _ in data])
    reps = batch.SchulzInvert(numIterations=4)
    for i, (H, D) in enumerate(data):
        lone = H.Copy()
        lone.SchulzInvert(numIterations=4)
        assert reps[i].iterations == 4
        assert _rel(batch.member(i).to_dense(), lone.to_dense()) < 1e-9

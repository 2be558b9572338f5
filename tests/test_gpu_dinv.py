"""HMat2d/HMat3d::DirectInvert on the GPU (src/HMat2d/Invert-incl.hpp:17-125):
against the dense inverse where nothing is truncated (maxRank at least every
admissible block's dimension), against the oracle's restatement (itself
pinned to numpy.linalg.inv in tests/test_oracle_hmat.py), and, truncated,
through the residual ||I - X A||_2. Also the Schulz iteration started from
the direct inverse. Tolerances: FP64 1e-9 relative to max|A^-1| without
truncation; FP32 2e-4; truncated residuals within 3x the oracle's (the GPU
compresses each block's sum of terms once by a sketch, the oracle after every
term by RoundedUpdate)."""
import numpy as np
import pytest

import dmhm_amd
from dmhm_amd import workloads as W
import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init():
    dmhm_amd.init(0)
    yield


def _op(n, L, dim, smooth=False, seed=3):
    op = W.small_operator(n, L, dim=dim, smooth=smooth, seed=seed)
    N = int(np.prod(op["sizes"]))
    return op, N, (op["rowptr"], op["colind"], op["vals"])


def _cls(dim):
    return dmhm_amd.HMat2d if dim == 2 else dmhm_amd.HMat3d


@pytest.mark.parametrize("n,L,dim,strong,rank", [(16, 3, 2, True, 16), (16, 3, 2, False, 60), (8, 3, 3, True, 8),
                                                 (12, 3, 2, True, 36)])
def test_direct_invert_exact_without_truncation(n, L, dim, strong, rank):
    print("case", n, L, dim, strong, rank, flush=True)
    op, N, args = _op(n, L, dim)
    A = _cls(dim).from_csr(op["sizes"], L, rank, strong, *args)
    D = W.csr_to_dense(*args, N)
    Xr = np.linalg.inv(D)
    X = A.Copy().DirectInvert().to_dense()
    assert np.abs(X - Xr).max() < 1e-9 * np.abs(Xr).max()
    H = orc.HMat.from_csr(dim, op["sizes"], L, rank, strong, *args)
    Xo = H.direct_invert().to_dense()
    assert np.abs(X - Xo).max() < 1e-9 * np.abs(Xr).max()
    ranks = A.Copy().DirectInvert().ranks()
    assert ranks.max() <= rank


def _dense_csr(D):
    n = D.shape[0]
    rowptr = np.arange(0, n * n + 1, n, dtype=np.int32)
    colind = np.tile(np.arange(n, dtype=np.int32), n)
    return rowptr, colind, np.ascontiguousarray(D, dtype=np.float64).ravel()


@pytest.mark.parametrize("sizes", [(1, 1), (5, 1), (17, 1), (33, 1), (8, 8), (10, 8)])
def test_dense_leaf_inverse_with_pivoting(sizes):
    """One dense leaf (numLevels 1): a non-symmetric Gaussian matrix with a
    weak diagonal, so partial pivoting exchanges rows at most steps (m <= 64:
    the register kernel; 80: the LDS kernel). Against numpy.linalg.inv,
    FP64 within 1e-9 of max|A^-1| scaled by the condition number's size
    (cond <= 1e4 here), FP32 within 1e-3."""
    print("case", sizes, flush=True)
    n = sizes[0] * sizes[1]
    rng = np.random.default_rng(n)
    D = rng.standard_normal((n, n)) + np.diag(1e-3 * np.ones(n))
    if n > 1:
        assert np.linalg.cond(D) < 1e4
    Xr = np.linalg.inv(D)
    args = _dense_csr(D)
    for prec, tol in ((dmhm_amd.DMHM_F64, 1e-9), (dmhm_amd.DMHM_F32, 1e-3)):
        A = dmhm_amd.HMat2d.from_csr([sizes[0], sizes[1]], 1, 4, True, *args, precision=prec)
        X = A.DirectInvert().to_dense()
        assert np.abs(X - Xr).max() < tol * np.abs(Xr).max(), (prec, np.abs(X - Xr).max() / np.abs(Xr).max())


def test_direct_invert_repeated_calls_are_identical():
    """The direct-inversion program and its arena stay with the geometry
    between calls: with one tile engine forced, every later call's result
    equals the first one bit for bit (same kernels and seeds), also when calls
    on another geometry interleave."""
    op, N, args = _op(16, 3, 2, smooth=False, seed=5)
    op2, N2, args2 = _op(12, 3, 2, smooth=False, seed=6)
    A = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 8, True, *args)
    B = dmhm_amd.HMat2d.from_csr(op2["sizes"], 3, 8, True, *args2)
    prev = dmhm_amd.set_tile_engine("lds")     # one engine: tuning would pick per run
    try:
        ref = A.Copy().DirectInvert().to_dense()
        refB = B.Copy().DirectInvert().to_dense()
        for i in range(6):
            print("call", i, flush=True)
            X = A.Copy().DirectInvert().to_dense()
            assert np.array_equal(X, ref), i
            Y = B.Copy().DirectInvert().to_dense()
            assert np.array_equal(Y, refB), i
    finally:
        dmhm_amd.set_tile_engine(prev)
    D = W.csr_to_dense(*args, N)
    assert np.linalg.norm(np.eye(N) - ref @ D, 2) < 1e-2


def test_direct_invert_fp32():
    op, N, args = _op(16, 3, 2)
    A = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 16, True, *args, precision=dmhm_amd.DMHM_F32)
    Xr = np.linalg.inv(W.csr_to_dense(*args, N))
    X = A.DirectInvert().to_dense()
    assert np.abs(X - Xr).max() < 2e-4 * np.abs(Xr).max()


def test_direct_invert_truncated_residual_tracks_oracle():
    """Smooth 32^2 operator, ranks 2, 4, 8: the residual falls with the rank
    and stays within 3x of the oracle's at the same rank."""
    op, N, args = _op(32, 4, 2, smooth=True)
    D = W.csr_to_dense(*args, N)
    res = []
    for rank in (2, 4, 8):
        A = dmhm_amd.HMat2d.from_csr(op["sizes"], 4, rank, True, *args)
        X = A.DirectInvert().to_dense()
        r = np.linalg.norm(np.eye(N) - X @ D, 2)
        H = orc.HMat.from_csr(2, op["sizes"], 4, rank, True, *args)
        ro = np.linalg.norm(np.eye(N) - H.direct_invert().to_dense() @ D, 2)
        print("rank", rank, "residual", r, "oracle", ro, flush=True)
        assert r <= 3.0 * ro + 1e-12
        res.append(r)
    assert res[0] > res[1] > res[2]


def test_program_stats_match_host_lowering():
    op, N, args = _op(16, 3, 2)
    A = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 8, True, *args)
    products, plans, seconds = A.direct_invert_program()
    st = dmhm_amd.check_direct_invert_program(2, op["sizes"], 3, 8, True)
    assert (products, plans) == (st["products"], st["plans"])
    assert seconds >= 0


def test_singular_pivot_is_reported():
    n = 8 * 8
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    vals = np.ones(n)
    vals[5] = 0.0                       # a zero row and column: singular
    A = dmhm_amd.HMat2d.from_csr([8, 8], 2, 4, True, rp, ci, vals)
    with pytest.raises(dmhm_amd.DmhmError) as e:
        A.DirectInvert()
    assert e.value.code == 3


def test_schulz_from_direct_start():
    """The iteration started from the direct inverse needs fewer steps than
    from alpha A^T and ends at least as accurate (64^2 lognormal, rank 8)."""
    op, N, args = _op(64, 4, 2, smooth=False, seed=1)
    D = W.csr_to_dense(*args, N)
    out = {}
    for start in ("transpose", "direct"):
        A = dmhm_amd.HMat2d.from_csr(op["sizes"], 4, 8, True, *args)
        rep = A.SchulzInvert(stopTol=1e-10, maxIterations=60, stall_stop=True, start=start)
        r = np.linalg.norm(np.eye(N) - A.to_dense() @ D, 2)
        out[start] = (rep.iterations, r)
        print(start, rep.iterations, rep.residualEstimate, r, flush=True)
    assert out["direct"][0] < out["transpose"][0]
    assert out["direct"][1] <= 1.5 * out["transpose"][1] + 1e-12

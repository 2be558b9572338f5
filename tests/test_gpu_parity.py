"""GPU parity tests: the gfx950 path (through the C-ABI) against the FP64 CPU
oracle on the same inputs, against dense brute force, and against
size-independent properties. Tolerances are written in each test:
  * FP64 H*X: 1e-12 relative (max-norm) vs the oracle;
  * FP32 H*X: 1e-5 relative;
  * exact-structure H*H products: 1e-12 relative vs the oracle / dense;
  * truncated H*H: within 3x the best blockwise rank-r truncation of the
    exact product (Eckart-Young floor) plus 1e-12.
"""
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


def _cls(dim):
    return dmhm_amd.HMat2d if dim == 2 else dmhm_amd.HMat3d


def _args(op):
    return op["rowptr"], op["colind"], op["vals"]


def _relmax(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


MV_CASES = [
    ("2d-16-strong", lambda: W.small_operator(16, 3), 4, True),
    ("2d-16-weak", lambda: W.small_operator(16, 3), 4, False),
    ("2d-ragged-20x12", lambda: W.rect_operator([20, 12], 3), 4, True),
    ("3d-6-strong", lambda: W.small_operator(6, 2, dim=3), 4, True),
    ("3d-ragged-10x6x5", lambda: W.rect_operator([10, 6, 5], 3), 8, False),
    ("2d-lognormal-32", lambda: W.small_operator(32, 3, smooth=False), 8, True),
]


@pytest.mark.parametrize("name,make,rank,strong", MV_CASES, ids=[c[0] for c in MV_CASES])
@pytest.mark.parametrize("trans", [False, True])
def test_matvec_fp64_matches_oracle(name, make, rank, strong, trans):
    op = make()
    dim = op["dim"]
    L = op["num_levels"]
    H = _cls(dim).from_csr(op["sizes"], L, rank, strong, *_args(op))
    R = orc.HMat.from_csr(dim, op["sizes"], L, rank, strong, *_args(op))
    n = H.Height()
    X = np.random.default_rng(11).standard_normal((n, 5))
    Y = H.TransposeMultiply(1.0, X) if trans else H.Multiply(1.0, X)
    Yr = R.multiply(X, trans=trans)
    assert _relmax(Y, Yr) < 1e-12


def test_import_is_exact():
    op = W.small_operator(16, 3)
    H = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 4, False, *_args(op))
    D = W.csr_to_dense(*_args(op), 256)
    np.testing.assert_array_equal(H.to_dense(), D)


@pytest.mark.parametrize("k", [1, 7, 64, 100])
def test_matvec_many_rhs_and_beta(k):
    op = W.small_operator(32, 3)
    H = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 8, True, *_args(op))
    H.SetToRandom(seed=5)
    D = H.to_dense()
    rng = np.random.default_rng(k)
    X = rng.standard_normal((1024, k))
    Y0 = rng.standard_normal((1024, k))
    Y = H.Multiply(0.5, X, -2.0, Y0.copy(order="F"))
    assert _relmax(Y, 0.5 * D @ X - 2.0 * Y0) < 1e-12


def test_matvec_fp32():
    op = W.small_operator(32, 3, smooth=False)
    H = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 8, True, *_args(op), precision=dmhm_amd.DMHM_F32)
    R = orc.HMat.from_csr(2, op["sizes"], 3, 8, True, *_args(op))
    X = np.random.default_rng(3).standard_normal((1024, 3))
    assert _relmax(H.Multiply(1.0, X), R.multiply(X)) < 1e-5


def test_multiply_writes_numpy_y_in_place():
    """Multiply(alpha, X, beta, Y) updates Y in place also when Y is C-ordered
    or of another dtype (the wrapper copies and writes back)."""
    op = W.small_operator(16, 3)
    H = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 4, True, *_args(op))
    rng = np.random.default_rng(11)
    X = rng.standard_normal((256, 3))
    Y0 = rng.standard_normal((256, 3))
    want = 2.0 * H.Multiply(1.0, X) + 0.5 * Y0
    for Y in (np.array(Y0, order="C"), np.array(Y0, order="F"), np.array(Y0, dtype=np.float32)):
        out = H.Multiply(2.0, X, 0.5, Y)
        assert out is Y
        tol = 1e-6 if Y.dtype == np.float32 else 1e-13
        assert _relmax(Y.astype(np.float64), want) < tol
    y = rng.standard_normal(256)
    want1 = H.Multiply(1.0, X[:, 0]) - y
    H.Multiply(1.0, X[:, 0], -1.0, y)
    assert _relmax(y, want1) < 1e-13


def test_set_to_random_ranks_and_consistency():
    op = W.small_operator(32, 3)
    H = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 6, True, *_args(op))
    H.SetToRandom(seed=9)
    assert np.all(H.ranks() == 6)
    x = np.random.default_rng(1).standard_normal(1024)
    assert _relmax(H.Multiply(1.0, x), H.to_dense() @ x) < 1e-12


def test_whole_matrix_routines():
    op = W.small_operator(16, 3)
    H = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 4, True, *_args(op))
    H.SetToRandom(seed=3)
    D = H.to_dense()
    T = H.Transpose()
    np.testing.assert_allclose(T.to_dense(), D.T, atol=1e-13)
    C = H.Copy()
    C.Scale(-2.5)
    np.testing.assert_allclose(C.to_dense(), -2.5 * D, atol=1e-12)
    C.AddConstantToDiagonal(3.0)
    np.testing.assert_allclose(C.to_dense(), -2.5 * D + 3.0 * np.eye(256), atol=1e-12)
    C.SetToIdentity()
    np.testing.assert_array_equal(C.to_dense(), np.eye(256))
    C.Scale(0.0)
    assert np.all(C.ranks() == 0)
    F = H.Convert(dmhm_amd.DMHM_F32)
    np.testing.assert_allclose(F.to_dense(), D, rtol=1e-5, atol=1e-5 * np.abs(D).max())


@pytest.mark.parametrize("name,n,L,rank,strong", [("c1-64", 64, 4, 8, True), ("16-weak-r16", 16, 3, 16, False)])
def test_hh_sparse_operands_match_oracle(name, n, L, rank, strong):
    """A*A of an imported stencil: an admissible block of A^2 couples at most
    two grid lines on each side of a box interface (rank <= 2 x 8 = 16 here;
    strong admissibility makes them zero), so both implementations are exact
    and must agree."""
    op = W.small_operator(n, L)
    H = dmhm_amd.HMat2d.from_csr(op["sizes"], L, rank, strong, *_args(op))
    R = orc.HMat.from_csr(2, op["sizes"], L, rank, strong, *_args(op))
    C = H.Multiply(-1.0, H)
    orc.rng_initialize()
    Cr = R.matmul(R, alpha=-1.0)
    assert _relmax(C.to_dense(), Cr.to_dense()) < 1e-12


def test_hh_truncating_product_flat_spectrum():
    """Weak admissibility at rank 4: A^2's admissible blocks have flat spectra
    (interface couplings), the hard case for a randomized range finder with
    oversampling 4. The expected Frobenius error of the sketch is at most
    sqrt(1 + k/(p-1)) times the optimum; the test allows 3x the blockwise
    Eckart-Young floor. The reference algorithm's sequential rounding is
    reported for comparison, not asserted."""
    op = W.small_operator(16, 3)
    H = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 4, False, *_args(op))
    D = W.csr_to_dense(*_args(op), 256)
    P = D @ D
    C = H.Multiply(1.0, H)
    leaves = dmhm_amd.geometry_leaves(op["sizes"], 3, False)
    floor = _blockwise_floor(P, leaves, 4)
    assert np.linalg.norm(C.to_dense() - P) <= 3.0 * floor


@pytest.mark.parametrize("dim,n,L,rank,strong", [(2, 16, 3, 16, True), (2, 8, 3, 16, False), (3, 8, 3, 8, True)])
def test_hh_exact_when_rank_covers_blocks(dim, n, L, rank, strong):
    op = W.small_operator(n, L, dim=dim)
    cls = _cls(dim)
    A = cls.from_csr(op["sizes"], L, rank, strong, *_args(op))
    B = cls.from_csr(op["sizes"], L, rank, strong, *_args(op))
    A.SetToRandom(seed=1)
    B.SetToRandom(seed=2)
    C = A.Multiply(0.75, B)
    ref = 0.75 * A.to_dense() @ B.to_dense()
    assert np.linalg.norm(C.to_dense() - ref) <= 1e-12 * np.linalg.norm(ref)


def _blockwise_floor(P, leaves, rank):
    """Frobenius error of the best rank-`rank` truncation of every admissible
    block of P (Eckart-Young), dense blocks exact."""
    err2 = 0.0
    for typ, lvl, r0, m, c0, n_, idx in leaves:
        if typ == 1:
            continue
        s = np.linalg.svd(P[r0:r0 + m, c0:c0 + n_], compute_uv=False)
        err2 += float(np.sum(s[rank:] ** 2))
    return np.sqrt(err2)


@pytest.mark.parametrize("rank", [4, 8])
def test_hh_truncation_near_eckart_young_floor(rank):
    """X = H-approximate inverse (Schulz), C = X * A. Each admissible block of
    C is compressed from a Gaussian sketch of the exact restricted product; the
    error must stay within 3x the blockwise optimum."""
    op = W.small_operator(32, 3)
    A = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, rank, True, *_args(op))
    X = A.Copy()
    X.SchulzInvert(stopTol=1e-30, maxIterations=12)
    Xd, Ad = X.to_dense(), A.to_dense()
    P = Xd @ Ad
    C = X.Multiply(1.0, A, seed=7)
    leaves = dmhm_amd.geometry_leaves(op["sizes"], 3, True)
    floor = _blockwise_floor(P, leaves, rank)
    err = np.linalg.norm(C.to_dense() - P)
    assert err <= 3.0 * floor + 1e-12 * np.linalg.norm(P)


def test_hh_tolerance_mode_adapts_ranks():
    op = W.small_operator(32, 3)
    A = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 16, True, *_args(op))
    X = A.Copy()
    X.SchulzInvert(stopTol=1e-30, maxIterations=14)
    P = X.to_dense() @ A.to_dense()
    C = X.Multiply(1.0, A, tol_mode=1, rel_tol=1e-6)
    r = C.ranks()
    assert r.max() <= 16 and r.min() < 16
    leaves = dmhm_amd.geometry_leaves(op["sizes"], 3, True)
    Cd = C.to_dense()
    for typ, lvl, r0, m, c0, n_, idx in leaves:
        if typ == 0 and r[idx] < 16:          # blocks not held at the rank cap
            blk = P[r0:r0 + m, c0:c0 + n_]
            nb = np.linalg.norm(blk, 2)
            if nb > 0:
                assert np.linalg.norm(Cd[r0:r0 + m, c0:c0 + n_] - blk, 2) <= 3e-6 * nb + 1e-14


def test_estimate_two_norm_brackets_true_norm():
    op = W.small_operator(16, 3)
    H = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 4, True, *_args(op))
    s = np.linalg.norm(W.csr_to_dense(*_args(op), 256), 2)
    est = H.EstimateTwoNorm(1.5, 6.0, seed=3)
    assert s * (1 - 1e-9) <= est <= 1.5 * s * (1 + 1e-9)


def test_schulz_exact_rank_reaches_roundoff():
    op = W.small_operator(16, 3)
    H = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 16, True, *_args(op))
    rep = H.SchulzInvert(stopTol=1e-9, maxIterations=60)
    D = W.csr_to_dense(*_args(op), 256)
    res = np.linalg.norm(np.eye(256) - D @ H.to_dense(), 2)
    assert rep.residualEstimate < 1e-9
    assert res < 1e-9


def test_schulz_bf16_on_well_conditioned_operator():
    """Mixed Schulz with BF16 tile operands in the early FP32 steps (the
    iteration's own quadratic convergence absorbs their rounding) on the
    smooth-coefficient 32^2 operator: it ends where the FP32 mixed form ends,
    within 10x on the dense residual ||I - A X||_2."""
    op = W.small_operator(32, 3)
    D = W.csr_to_dense(*_args(op), 1024)
    out = {}
    for bf16 in (False, True):
        H = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 16, True, *_args(op))
        rep = H.SchulzInvert(stopTol=1e-8, maxIterations=60, mixed=True, bf16=bf16, stall_stop=True)
        r = np.linalg.norm(np.eye(1024) - D @ H.to_dense(), 2)
        print("bf16", bf16, "iterations", rep.iterations, "bf16 steps", rep.iterationsBf16,
              "estimate", rep.residualEstimate, "dense residual", r, flush=True)
        assert np.isfinite(rep.residualEstimate)
        out[bf16] = (rep, r)
    assert out[True][0].iterationsBf16 > 0
    assert out[True][1] <= 10 * out[False][1] + 1e-12


@pytest.mark.parametrize("mixed", [False, True])
def test_schulz_graph_replay_matches_eager(mixed):
    """Steps replayed from captured HIP graphs launch what the eager steps
    launch (same plans, same counter-offset seeds): every step's residual
    estimate and the final inverse agree. Each replay is followed by the
    step's readback and host decision, so replays interleave with eager work."""
    op = W.small_operator(32, 3, smooth=False)
    kw = dict(numIterations=-1, stopTol=1e-13, maxIterations=14, mixed=mixed, switchTol=1e-2, seed=7)
    # One fixed tile engine for both runs (autotuning would time a different
    # candidate engine in each run's first steps).
    prev = dmhm_amd.set_tile_engine("mfma")
    try:
        He = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 8, True, *_args(op))
        Hg = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 8, True, *_args(op))
        re = He.SchulzInvert(graph=False, **kw)
        rg = Hg.SchulzInvert(graph=True, **kw)
    finally:
        dmhm_amd.set_tile_engine(prev)
    assert re.graphIterations == 0
    assert rg.graphIterations >= 8
    assert re.iterations == rg.iterations == 14
    assert re.iterationsLowPrecision == rg.iterationsLowPrecision
    he = np.array(re.history[:re.iterations])
    hg = np.array(rg.history[:rg.iterations])
    assert np.all(np.abs(he - hg) <= 1e-9 * np.abs(he))
    De, Dg = He.to_dense(), Hg.to_dense()
    assert _relmax(Dg, De) < 1e-12
    # 14 steps leave this operator in the linear phase: I - A X contracts
    # (spectral norm below 1) and is the same for both runs.
    D = W.csr_to_dense(*_args(op), 1024)
    rg_norm = np.linalg.norm(np.eye(1024) - D @ Dg, 2)
    assert rg_norm < 1.0
    assert abs(rg_norm - np.linalg.norm(np.eye(1024) - D @ De, 2)) < 1e-10


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_every_tile_engine_gives_the_same_product(prec):
    """The LDS kernel, the MFMA kernel and its prefetching form (and the BF16
    forms, which fall back to FP32/FP64 MFMA outside a mixed Schulz loop)
    compute the same exact product: rank-16 operands on a 20 x 12 grid, whose
    level-2 boxes hold 15 points, so no block needs more than rank 16; against
    the dense product (FP64 1e-12, FP32 1e-5)."""
    op = W.rect_operator([20, 12], 3)      # level-2 boxes of 15 points: rank-16 products are exact
    sizes = op["sizes"]
    n = int(np.prod(sizes))
    p = dmhm_amd.DMHM_F64 if prec == "f64" else dmhm_amd.DMHM_F32
    A = dmhm_amd.HMat2d.from_csr(sizes, 3, 16, True, *_args(op))
    A.SetToRandom(seed=3)
    B = dmhm_amd.HMat2d.from_csr(sizes, 3, 16, True, *_args(op))
    B.SetToRandom(seed=4)
    ref = A.to_dense() @ B.to_dense()
    if p == dmhm_amd.DMHM_F32:
        A, B = A.Convert(p), B.Convert(p)
    tol = 1e-12 if p == dmhm_amd.DMHM_F64 else 1e-5
    prev = dmhm_amd.set_tile_engine("lds")
    try:
        for eng in ("lds", "mfma", "mfma_pf", "bf16", "bf16_pf", "auto", "auto", "auto", "auto"):
            dmhm_amd.set_tile_engine(eng)
            C = A.Multiply(0.5, B, oversample=4)
            assert _relmax(C.to_dense(), 0.5 * ref) < tol, eng
    finally:
        dmhm_amd.set_tile_engine(prev)


def _inverse_floor(D, leaves, rank):
    """||I - A X_opt||_2 for X_opt = blockwise rank-`rank` truncation of inv(A)."""
    Ainv = np.linalg.inv(D)
    X = Ainv.copy()
    for typ, lvl, r0, m, c0, n_, idx in leaves:
        if typ == 1:
            continue
        U, s, VT = np.linalg.svd(Ainv[r0:r0 + m, c0:c0 + n_])
        X[r0:r0 + m, c0:c0 + n_] = (U[:, :rank] * s[:rank]) @ VT[:rank]
    return np.linalg.norm(np.eye(D.shape[0]) - D @ X, 2)


@pytest.mark.parametrize("mixed", [False, True])
def test_schulz_reaches_truncation_floor(mixed):
    """32^2 grid, maxRank 8: the reachable ||I - AX||_2 is set by the best
    blockwise rank-8 approximation of inv(A); the GPU iteration must land
    within 2x of it (the reference algorithm's sequential rounding lands an
    order of magnitude above it on this case)."""
    op = W.small_operator(32, 3)
    H = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 8, True, *_args(op))
    rep = H.SchulzInvert(stopTol=1e-30, maxIterations=40, mixed=mixed, switchTol=1e-3)
    D = W.csr_to_dense(*_args(op), 1024)
    leaves = dmhm_amd.geometry_leaves(op["sizes"], 3, True)
    floor = _inverse_floor(D, leaves, 8)
    res = np.linalg.norm(np.eye(1024) - D @ H.to_dense(), 2)
    assert res <= 2.0 * floor
    if mixed:
        assert rep.iterationsLowPrecision > 0


def test_c1_full_size_parity():
    cfg = W.config_operator("c1")
    args = _args(cfg)
    H = dmhm_amd.HMat2d.from_csr(cfg["sizes"], 4, 8, True, *args)
    R = orc.HMat.from_csr(2, cfg["sizes"], 4, 8, True, *args)
    X = np.random.default_rng(2).standard_normal((4096, 16))
    assert _relmax(H.Multiply(1.0, X), R.multiply(X)) < 1e-12
    C = H.Multiply(1.0, H)
    Cr = R.matmul(R)
    assert _relmax(C.to_dense(), Cr.to_dense()) < 1e-12


def test_c2_full_size_matvec_parity():
    cfg = W.config_operator("c2")
    args = _args(cfg)
    H = dmhm_amd.HMat2d.from_csr(cfg["sizes"], 6, 16, True, *args)
    R = orc.HMat.from_csr(2, cfg["sizes"], 6, 16, True, *args)
    X = np.random.default_rng(4).standard_normal((65536, 2))
    assert _relmax(H.Multiply(1.0, X), R.multiply(X)) < 1e-12
    assert _relmax(H.TransposeMultiply(1.0, X), R.multiply(X, trans=True)) < 1e-12


def test_nonconformal_product_raises():
    a = W.small_operator(16, 3)
    b = W.small_operator(32, 3)
    A = dmhm_amd.HMat2d.from_csr(a["sizes"], 3, 4, True, *_args(a))
    B = dmhm_amd.HMat2d.from_csr(b["sizes"], 3, 4, True, *_args(b))
    with pytest.raises(dmhm_amd.NonConformalError):
        A.Multiply(1.0, B)


def test_bad_arguments_raise():
    op = W.small_operator(16, 3)
    with pytest.raises(dmhm_amd.DmhmError):
        dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 99, True, *_args(op))
    H = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 4, True, *_args(op))
    with pytest.raises(dmhm_amd.DmhmError):
        H.EstimateTwoNorm(theta=1.0)


# --- update forms, low-rank import, RoundedUpdate / Compress -----------------

@pytest.mark.parametrize("name,n,L,rank,strong", [("c1-64", 64, 4, 8, True), ("16-weak-r16", 16, 3, 16, False)])
def test_hh_update_form_matches_oracle(name, n, L, rank, strong):
    """C := -A A + 0.5 C with C = A (MultiplyHMat-incl.hpp:189-380); the
    operands are exact at these ranks, so both implementations agree."""
    op = W.small_operator(n, L)
    H = dmhm_amd.HMat2d.from_csr(op["sizes"], L, rank, strong, *_args(op))
    C = H.Copy()
    H.Multiply(-1.0, H, 0.5, C)
    R = orc.HMat.from_csr(2, op["sizes"], L, rank, strong, *_args(op))
    Cr = R.copy()
    orc.rng_initialize()
    Cr.multiply_update(-1.0, R, R, 0.5)
    assert _relmax(C.to_dense(), Cr.to_dense()) < 1e-12


@pytest.mark.parametrize("dim,n,L,rank,strong", [(2, 16, 3, 16, True), (3, 8, 3, 8, True)])
def test_hh_update_form_exact_on_random_operands(dim, n, L, rank, strong):
    op = W.small_operator(n, L, dim=dim)
    cls = _cls(dim)
    A, B, C = (cls.from_csr(op["sizes"], L, rank, strong, *_args(op)) for _ in range(3))
    A.SetToRandom(seed=1)
    B.SetToRandom(seed=2)
    C.SetToRandom(seed=3)
    ref = 0.75 * A.to_dense() @ B.to_dense() - 0.25 * C.to_dense()
    A.Multiply(0.75, B, -0.25, C)
    assert np.linalg.norm(C.to_dense() - ref) <= 1e-12 * np.linalg.norm(ref)


def test_update_with_exact_sum():
    """A := A + alpha B at full block rank: exact."""
    op = W.small_operator(16, 3)
    A = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 16, True, *_args(op))
    B = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 16, True, *_args(op))
    A.SetToRandom(seed=5)
    B.SetToRandom(seed=6)
    ref = A.to_dense() + 0.3 * B.to_dense()
    A.UpdateWith(0.3, B)
    assert np.linalg.norm(A.to_dense() - ref) <= 1e-12 * np.linalg.norm(ref)


def test_update_with_truncation_is_optimal():
    """Two rank-4 operands summed and rounded to rank 4: the sketch (4 + 4
    oversampling columns) spans the whole rank-8 range of every block, so the
    result is the best rank-4 truncation (RoundedUpdate's QR + SVD result):
    the error equals the blockwise Eckart-Young floor."""
    op = W.small_operator(16, 3)
    A = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 4, False, *_args(op))
    B = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 4, False, *_args(op))
    A.SetToRandom(seed=7)
    B.SetToRandom(seed=8)
    P = A.to_dense() - 0.6 * B.to_dense()
    A.UpdateWith(-0.6, B)
    leaves = dmhm_amd.geometry_leaves(op["sizes"], 3, False)
    floor = _blockwise_floor(P, leaves, 4)
    err = np.linalg.norm(A.to_dense() - P)
    assert floor > 1e-3 * np.linalg.norm(P)
    assert abs(err - floor) <= 1e-9 * np.linalg.norm(P)


@pytest.mark.parametrize("r,max_rank", [(3, 4), (6, 4)])
def test_import_lowrank_is_exact(r, max_rank):
    """HMat2d(LowRank F, ...) (HMat2d.cpp:124-142, ImportLowRank): admissible
    blocks keep the exact factors (rank r even above maxRank), dense leaves
    hold U_t V_s^T, so the H-matrix equals U V^T."""
    rng = np.random.default_rng(11)
    n = 256
    U = rng.standard_normal((n, r))
    V = rng.standard_normal((n, r))
    H = dmhm_amd.HMat2d.from_lowrank([16, 16], 3, max_rank, True, U, V)
    assert _relmax(H.to_dense(), U @ V.T) < 1e-13
    assert set(np.unique(H.ranks())) == {r}
    X = rng.standard_normal((n, 3))
    assert _relmax(H.Multiply(1.0, X), U @ (V.T @ X)) < 1e-12


@pytest.mark.parametrize("m,n,rA,rB,max_rank", [(40, 30, 5, 7, 6), (5, 40, 4, 8, 3), (64, 64, 30, 34, 16)])
def test_rounded_update_matches_oracle(m, n, rA, rB, max_rank):
    rng = np.random.default_rng(m + n + rA)
    AU, AV = rng.standard_normal((m, rA)), rng.standard_normal((n, rA))
    BU, BV = rng.standard_normal((m, rB)), rng.standard_normal((n, rB))
    U, V = dmhm_amd.RoundedUpdate(max_rank, 0.7, (AU, AV), -1.3, (BU, BV))
    Ur, Vr = orc.rounded_update(max_rank, 0.7, AU, AV, -1.3, BU, BV)
    assert U.shape[1] == Ur.shape[1] == min(max_rank, m, n)
    assert _relmax(U @ V.T, Ur @ Vr.T) < 1e-11
    assert np.abs(U.T @ U - np.eye(U.shape[1])).max() < 1e-12


def test_rounded_update_early_exit_concatenates():
    rng = np.random.default_rng(3)
    AU, AV = rng.standard_normal((20, 3)), rng.standard_normal((25, 3))
    BU, BV = rng.standard_normal((20, 2)), rng.standard_normal((25, 2))
    U, V = dmhm_amd.RoundedUpdate(8, 2.0, (AU, AV), 0.5, (BU, BV))
    Ur, Vr = orc.rounded_update(8, 2.0, AU, AV, 0.5, BU, BV)
    assert np.array_equal(U, Ur) and np.array_equal(V, Vr)


# Tall panels reach every register class of the QR (m <= 64, 256, 512 and
# 1024 rows; 24- and 40-column variants) and the wide fallback.
@pytest.mark.parametrize("m,n,max_rank", [(50, 20, 8), (10, 30, 6), (20, 90, 5), (64, 64, 64), (63, 17, 17),
                                          (200, 20, 12), (200, 36, 30), (400, 36, 20), (500, 23, 23),
                                          (900, 20, 16), (1000, 24, 10)])
def test_compress_matches_oracle(m, n, max_rank):
    rng = np.random.default_rng(m * n)
    D = rng.standard_normal((m, n)) * np.logspace(0, -6, n)
    U, V = dmhm_amd.Compress(max_rank, D)
    Ur, Vr = orc.compress_dense(max_rank, D)
    assert U.shape[1] == Ur.shape[1] == min(max_rank, m, n)
    assert _relmax(U @ V.T, Ur @ Vr.T) < 1e-11


# --- Pack / Unpack ----------------------------------------------------------

@pytest.mark.parametrize("kind", ["import", "random64", "random32", "product"])
def test_pack_unpack_round_trip(kind):
    """Unpack(Pack(H)) reproduces H exactly and packs to the same bytes (the
    reference's own hard check, tests/HMat2d/Pack.cpp:217)."""
    op = W.small_operator(16, 3)
    H = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 8, False, *_args(op))
    if kind.startswith("random"):
        H.SetToRandom(seed=3)
        if kind == "random32":
            H = H.Convert(dmhm_amd.DMHM_F32)
    elif kind == "product":
        H.SetToRandom(seed=4)
        H = H.Multiply(1.0, H)
    img = H.Pack()
    assert len(img) == H.PackedSize()
    G = dmhm_amd.HMat2d.Unpack(img)
    assert np.array_equal(G.to_dense(), H.to_dense())
    assert np.array_equal(G.ranks(), H.ranks())
    assert G.Pack() == img
    x = np.random.default_rng(0).standard_normal(256)
    assert np.array_equal(G.Multiply(1.0, x), H.Multiply(1.0, x))


def test_unpack_rejects_bad_images():
    op = W.small_operator(16, 3)
    H = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, 8, True, *_args(op))
    img = H.Pack()
    with pytest.raises(dmhm_amd.DmhmError):
        dmhm_amd.HMat2d.Unpack(img[:-8])
    with pytest.raises(dmhm_amd.DmhmError):
        dmhm_amd.HMat2d.Unpack(b"XXXXXXXX" + img[8:])


@pytest.mark.parametrize("sizes,L,strong", [([16, 16], 3, True), ([16, 16], 3, False), ([20, 12], 3, False),
                                            ([10, 6, 5], 3, False), ([6, 6, 6], 2, True), ([33, 17], 4, False)])
def test_device_import_matches_oracle_import(sizes, L, strong):
    """The CSR import on the GPU equals the oracle's ImportSparse restatement
    exactly (dense leaves and the unit/value low-rank terms expand to the same
    matrix, bitwise), and every admissible block's rank is its nonzero count
    (the oracle's import rank)."""
    op = W.rect_operator(sizes, L)
    dim = len(sizes)
    n = int(np.prod(sizes))
    H = _cls(dim).from_csr(sizes, L, 4, strong, *_args(op))
    R = orc.HMat.from_csr(dim, sizes, L, 4, strong, *_args(op))
    np.testing.assert_array_equal(H.to_dense(), R.to_dense())
    np.testing.assert_array_equal(H.to_dense(), W.csr_to_dense(*_args(op), n))
    leaves = dmhm_amd.geometry_leaves(sizes, L, strong)
    rp, ci = op["rowptr"], op["colind"]
    ranks = H.ranks()
    for typ, lvl, r0, m, c0, nn, idx in leaves:
        if typ != 0:
            continue
        cnt = 0
        for i in range(r0, r0 + m):
            cols = ci[rp[i]:rp[i + 1]]
            cnt += int(np.count_nonzero((cols >= c0) & (cols < c0 + nn)))
        assert ranks[idx] == cnt


@pytest.mark.parametrize("sizes,L", [([16, 16], 3), ([20, 12], 3), ([7, 9], 2), ([10, 6, 5], 3), ([64, 64], 4),
                                     ([33, 33, 17], 4)])
def test_device_natural_map_matches_host(sizes, L):
    d = dmhm_amd.build_natural_to_hierarchical_map_device(sizes, L)
    h = dmhm_amd.build_natural_to_hierarchical_map(sizes, L)
    np.testing.assert_array_equal(d, h)
    assert sorted(d.tolist()) == list(range(int(np.prod(sizes))))

"""Pins the oracle's LAPACK substitutes (oracle/la.cpp) against numpy.linalg
(LAPACK) on the same inputs, including rank-deficient and wide cases."""
import numpy as np
import pytest

import oracle as orc


def _cases():
    rng = np.random.default_rng(7)
    out = [rng.standard_normal((40, 12)), rng.standard_normal((12, 12)), rng.standard_normal((9, 20))]
    low = rng.standard_normal((50, 3)) @ rng.standard_normal((3, 16))     # rank 3
    out.append(low)
    out.append(np.zeros((6, 4)))
    return out


@pytest.mark.parametrize("A", _cases())
def test_qr_matches_lapack(A):
    QR, tau = orc.qr(A)
    m, n = A.shape
    k = min(m, n)
    R = np.triu(QR)[:k, :]
    Q = orc.form_q(QR[:, :k], tau, k)
    np.testing.assert_allclose(Q @ R, A, atol=1e-12 * max(1, np.abs(A).max()))
    np.testing.assert_allclose(Q.T @ Q, np.eye(k), atol=1e-13)
    Rl = np.linalg.qr(A, mode="r")[:k, :]
    np.testing.assert_allclose(np.abs(R), np.abs(Rl), atol=1e-11 * max(1, np.abs(A).max()))


@pytest.mark.parametrize("A", _cases())
def test_pivoted_qr_reveals_rank(A):
    QR, jpvt, tau = orc.pivoted_qr(A)
    m, n = A.shape
    k = min(m, n)
    R = np.triu(QR)[:k, :]
    Q = orc.form_q(QR[:, :k], tau, k)
    np.testing.assert_allclose(Q @ R, A[:, jpvt], atol=1e-12 * max(1, np.abs(A).max()))
    d = np.abs(np.diag(R))
    assert np.all(np.diff(d) <= 1e-12 * max(1, d.max()))       # non-increasing diagonal
    rank = np.linalg.matrix_rank(A)
    if rank < k:
        assert d[rank] < 1e-10 * max(1, d[0])


@pytest.mark.parametrize("A", _cases())
def test_svd_matches_lapack(A):
    U, s, VT = orc.svd(A)
    sl = np.linalg.svd(A, compute_uv=False)
    np.testing.assert_allclose(s, sl, atol=1e-12 * max(1, sl.max() if sl.size else 1))
    np.testing.assert_allclose((U * s) @ VT, A, atol=1e-12 * max(1, np.abs(A).max()))
    r = int(np.sum(s > 1e-12 * max(s.max(), 1e-300))) if s.size else 0
    np.testing.assert_allclose(U[:, :r].T @ U[:, :r], np.eye(r), atol=1e-12)
    np.testing.assert_allclose(VT[:r] @ VT[:r].T, np.eye(r), atol=1e-12)


def test_gemm_transposes():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((5, 7))
    B = rng.standard_normal((5, 4))
    C = np.asfortranarray(rng.standard_normal((7, 4)))
    import ctypes
    out = C.copy(order="F")
    orc.lib().orc_gemm(b"T", b"N", 7, 4, 5, 2.0, orc._d(np.asfortranarray(A)), 5, orc._d(np.asfortranarray(B)),
                       5, 0.5, orc._d(out), 7)
    np.testing.assert_allclose(out, 2.0 * A.T @ B + 0.5 * C, atol=1e-13)

"""The allocation-free H*X entry point (dmhm_hmat_multiply_dense_noalloc)
against the allocating one (dmhm_hmat_multiply_dense) through the C-ABI, on
device operands: two calls in a row, the second on a smaller operator with
fewer right-hand sides, and its error paths (no prepared plan, workspace above
the scratch maximum)."""
import ctypes

import numpy as np
import pytest

import dmhm_amd
from dmhm_amd import _lib
from dmhm_amd import workloads as W

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc / hipMemcpy / hipFree through ctypes (the test owns these buffers)."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []

    def to_device(self, a):
        a = np.asfortranarray(a)
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0    # host -> device
        return p

    def to_host(self, p, like):
        out = np.zeros(like.shape, dtype=like.dtype, order="F")
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, 2) == 0  # device -> host
        return out

    def free(self, p):
        assert self.L.hipFree(p) == 0


def _operator(side, max_rank, prec):
    op = W.small_operator(side, 3)
    H = dmhm_amd.HMat2d.from_csr(op["sizes"], 3, max_rank, True, op["rowptr"], op["colind"], op["vals"],
                                 precision=prec)
    H.SetToRandom(seed=side)         # low-rank blocks at rank max_rank: the projections are used
    return H, side * side


def _both(hip, H, n, k, prec, seed):
    dt = np.float64 if prec == dmhm_amd.DMHM_F64 else np.float32
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, k)).astype(dt)
    Y0 = rng.standard_normal((n, k)).astype(dt)
    dX, dY1, dY2 = hip.to_device(X), hip.to_device(Y0), hip.to_device(Y0)
    L = _lib.load()
    try:
        _lib.check(L.dmhm_hmat_multiply_dense(H._h, b"N", 1.5, dX, n, k, -0.5, dY1, n, _lib.DMHM_DEVICE))
        _lib.check(L.dmhm_hmat_multiply_dense_noalloc(H._h, 1.5, dX, n, k, -0.5, dY2, n))
        return hip.to_host(dY1, Y0), hip.to_host(dY2, Y0)
    finally:
        for p in (dX, dY1, dY2):
            hip.free(p)


def _rel(a, b):
    return float(np.abs(a.astype(np.float64) - b.astype(np.float64)).max() / np.abs(b).max())


def test_noalloc_matches_allocating_entry_point_twice_shrinking():
    hip = _Hip()
    for prec, tol in ((dmhm_amd.DMHM_F64, 1e-13), (dmhm_amd.DMHM_F32, 1e-5)):
        H1, n1 = _operator(32, 8, prec)            # first call: 1024 unknowns, 5 columns
        ref, got = _both(hip, H1, n1, 5, prec, 1)
        assert _rel(got, ref) < tol
        H2, n2 = _operator(16, 4, prec)            # second call: smaller operator, 2 columns
        ref, got = _both(hip, H2, n2, 2, prec, 2)
        assert _rel(got, ref) < tol


def test_noalloc_without_prepared_plan_is_rejected():
    hip = _Hip()
    H, n = _operator(16, 4, dmhm_amd.DMHM_F64)
    X = np.ones((n, 7))
    dX, dY = hip.to_device(X), hip.to_device(X)
    try:
        code = _lib.load().dmhm_hmat_multiply_dense_noalloc(H._h, 1.0, dX, n, 7, 0.0, dY, n)
        assert code == 1 and b"prepared plan" in _lib.load().dmhm_last_error()
        code = _lib.load().dmhm_hmat_multiply_dense_noalloc(H._h, 1.0, dX, n, 9, 0.0, dY, n)
        assert code == 1
    finally:
        hip.free(dX)
        hip.free(dY)


def test_noalloc_rejects_workspace_above_scratch_maximum():
    # 3D 32^3, maxRank 32, FP64, 8 columns: the projections need
    # 56,448 blocks x 32 x 8 x 8 B = 115.6 MB > 64 MiB.
    op = W.config_operator("c3", seed=0)
    H = dmhm_amd.HMat3d.from_csr(op["sizes"], op["num_levels"], op["max_rank"], True, op["rowptr"], op["colind"],
                                 op["vals"])
    n, k = int(np.prod(op["sizes"])), 8
    hip = _Hip()
    X = np.zeros((n, k))
    dX, dY = hip.to_device(X), hip.to_device(X)
    try:
        _lib.check(_lib.load().dmhm_hmat_multiply_dense(H._h, b"N", 1.0, dX, n, k, 0.0, dY, n, _lib.DMHM_DEVICE))
        code = _lib.load().dmhm_hmat_multiply_dense_noalloc(H._h, 1.0, dX, n, k, 0.0, dY, n)
        assert code == 1 and b"scratch maximum" in _lib.load().dmhm_last_error()
    finally:
        hip.free(dX)
        hip.free(dY)

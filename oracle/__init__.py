"""ORACLE — test infrastructure only.

ctypes wrapper over ``oracle/liboracle.so``, the FP64 CPU restatement of the
reference's sequential HMat2d/HMat3d + hmat_tools arithmetic (see
``oracle/hmat.hpp`` for the reference files each routine follows and for how
each one is pinned). Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg import this module, and only as the
checker; the product (``dmhm_amd``) never imports, links or calls it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_LAPACK_PATH = os.path.join(_HERE, "liboracle_lapack.so")
_lib = None
_lapack = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    """The self-contained oracle (textbook dense routines): the checker."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _bind(ctypes.CDLL(_LIB_PATH))
    return _lib


def lapack_lib():
    """The same oracle with its dense routines bound to MKL's sequential
    BLAS/LAPACK (liboracle_lapack.so), or None when it was not built (no MKL)."""
    global _lapack
    if _lapack is None and os.path.exists(_LAPACK_PATH):
        _lapack = _bind(ctypes.CDLL(_LAPACK_PATH, mode=ctypes.RTLD_GLOBAL))  # MKL loads its CPU kernels by name
    return _lapack


def _bind(L):
    if L is not None:
        sig = {
            "orc_set_options": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int]),
            "orc_rng_initialize": (None, []),
            "orc_rng_set_state": (None, [ctypes.c_uint64, ctypes.c_uint64]),
            "orc_rng_serial_raw": (None, [ctypes.c_int, _u64p]),
            "orc_rng_serial_uniform": (None, [ctypes.c_int, _dp]),
            "orc_rng_serial_gaussian_vector": (None, [ctypes.c_int, _dp]),
            "orc_rng_serial_gaussian_vectors": (None, [ctypes.c_int, ctypes.c_int, _dp]),
            "orc_gemm": (None, [ctypes.c_char, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_double, _dp, ctypes.c_int, _dp, ctypes.c_int,
                                ctypes.c_double, _dp, ctypes.c_int]),
            "orc_qr": (None, [ctypes.c_int, ctypes.c_int, _dp, _dp]),
            "orc_pivoted_qr": (None, [ctypes.c_int, ctypes.c_int, _dp, _ip, _dp]),
            "orc_form_q": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]),
            "orc_apply_q": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]),
            "orc_svd": (None, [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]),
            "orc_build_map": (None, [ctypes.c_int, _ip, ctypes.c_int, _ip]),
            "orc_hmat_create": (_vp, [ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      _ip, _ip, _dp]),
            "orc_hmat_free": (None, [_vp]),
            "orc_hmat_height": (ctypes.c_int, [_vp]),
            "orc_hmat_set_to_random": (None, [_vp]),
            "orc_hmat_multiply_dense": (None, [_vp, ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_double, _dp, ctypes.c_int]),
            "orc_hmat_to_dense": (None, [_vp, _dp]),
            "orc_hmat_count_blocks": (None, [_vp, _lp, _lp, ctypes.c_int]),
            "orc_hmat_rank_stats": (None, [_vp, _lp]),
            "orc_hmat_multiply": (_vp, [ctypes.c_double, _vp, _vp]),
            "orc_hmat_multiply_update": (None, [ctypes.c_double, _vp, _vp, ctypes.c_double, _vp]),
            "orc_hmat_copy": (_vp, [_vp]),
            "orc_hmat_direct_invert": (None, [_vp]),
            "orc_invert": (ctypes.c_int, [ctypes.c_int, _dp]),
            "orc_hmat_transpose": (_vp, [_vp]),
            "orc_hmat_scale": (None, [_vp, ctypes.c_double]),
            "orc_hmat_add_diag": (None, [_vp, ctypes.c_double]),
            "orc_hmat_set_identity": (None, [_vp]),
            "orc_estimate_two_norm": (ctypes.c_double, [_vp, ctypes.c_double, ctypes.c_double]),
            "orc_hmat_schulz": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_int, _dp, ctypes.c_int]),
            "orc_rounded_update": (None, [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_int, _dp,
                                          _dp, _dp, _dp, _ip]),
            "orc_compress_dense": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _ip]),
            "orc_randomized_multiply": (None, [ctypes.c_int, ctypes.c_double, _vp, _vp, ctypes.c_int,
                                               _dp, _dp, _ip]),
            "orc_hmat_multiply_child_row": (_vp, [ctypes.c_double, _vp, _vp, ctypes.c_int]),
            "orc_hmat_multiply_child": (_vp, [ctypes.c_double, _vp, _vp, ctypes.c_int, ctypes.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
    return L


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def set_options(oversample=4, tol_mode=0, rel_tol=1e-10, ref_quirk=True):
    lib().orc_set_options(oversample, tol_mode, rel_tol, 1 if ref_quirk else 0)


def rng_initialize():
    lib().orc_rng_initialize()


def rng_serial_raw(n):
    out = np.zeros(n, dtype=np.uint64)
    lib().orc_rng_serial_raw(n, out.ctypes.data_as(_u64p))
    return out


def rng_serial_uniform(n):
    out = np.zeros(n)
    lib().orc_rng_serial_uniform(n, _d(out))
    return out


def gaussian_vector(n):
    out = np.zeros(n)
    lib().orc_rng_serial_gaussian_vector(n, _d(out))
    return out


def gaussian_vectors(m, n):
    out = np.zeros((m, n), order="F")
    lib().orc_rng_serial_gaussian_vectors(m, n, _d(out))
    return out


def qr(A):
    A = np.array(A, dtype=np.float64, order="F")
    m, n = A.shape
    tau = np.zeros(min(m, n))
    lib().orc_qr(m, n, _d(A), _d(tau))
    return A, tau


def form_q(QR, tau, ncols):
    m = QR.shape[0]
    A = np.zeros((m, max(ncols, QR.shape[1])), order="F")
    A[:, :QR.shape[1]] = QR
    lib().orc_form_q(m, ncols, len(tau), _d(A), _d(np.ascontiguousarray(tau)))
    return A[:, :ncols].copy()


def pivoted_qr(A):
    A = np.array(A, dtype=np.float64, order="F")
    m, n = A.shape
    tau = np.zeros(min(m, n))
    jpvt = np.zeros(n, dtype=np.int32)
    lib().orc_pivoted_qr(m, n, _d(A), _i(jpvt), _d(tau))
    return A, jpvt, tau


def svd(A):
    A = np.array(A, dtype=np.float64, order="F")
    m, n = A.shape
    k = min(m, n)
    U = np.zeros((m, k), order="F")
    s = np.zeros(k)
    VT = np.zeros((k, n), order="F")
    lib().orc_svd(m, n, _d(A), _d(U), _d(s), _d(VT))
    return U, s, VT


def build_map(dim, sizes, num_levels):
    sizes = np.array(list(sizes) + [1] * (3 - len(sizes)), dtype=np.int32)
    n = int(np.prod(sizes[:dim]))
    out = np.zeros(n, dtype=np.int32)
    lib().orc_build_map(dim, _i(sizes), num_levels, _i(out))
    return out


def rounded_update(max_rank, alpha, AU, AV, beta, BU, BV):
    AU = np.asfortranarray(AU, dtype=np.float64); AV = np.asfortranarray(AV, dtype=np.float64)
    BU = np.asfortranarray(BU, dtype=np.float64); BV = np.asfortranarray(BV, dtype=np.float64)
    m, rA = AU.shape
    n = AV.shape[0]
    rB = BU.shape[1]
    cap = rA + rB
    U = np.zeros(m * cap)
    V = np.zeros(n * cap)
    r = ctypes.c_int(0)
    lib().orc_rounded_update(max_rank, alpha, m, n, rA, _d(AU), _d(AV), beta, rB, _d(BU), _d(BV),
                             _d(U), _d(V), ctypes.byref(r))
    r = r.value
    return U[:m * r].reshape((m, r), order="F"), V[:n * r].reshape((n, r), order="F")


def invert(A):
    """Inverse of a square matrix by the oracle's Gauss-Jordan (partial pivoting)."""
    A = np.array(A, dtype=np.float64, order="F")
    if lib().orc_invert(A.shape[0], _d(A)) != 0:
        raise np.linalg.LinAlgError("singular matrix")
    return A


def compress_dense(max_rank, D):
    D = np.asfortranarray(D, dtype=np.float64)
    m, n = D.shape
    k = min(m, n)
    U = np.zeros(m * k)
    V = np.zeros(n * k)
    r = ctypes.c_int(0)
    lib().orc_compress_dense(max_rank, m, n, _d(D), _d(U), _d(V), ctypes.byref(r))
    r = r.value
    return U[:m * r].reshape((m, r), order="F"), V[:n * r].reshape((n, r), order="F")


class HMat:
    """Owning handle on an oracle H-matrix (dim 2 or 3). ``L`` is the oracle
    library that owns it (``lib()`` by default, ``lapack_lib()`` for timing)."""

    def __init__(self, ptr, L=None):
        self.ptr = ptr
        self.L = L if L is not None else lib()

    @classmethod
    def from_csr(cls, dim, sizes, num_levels, max_rank, strong, rowptr, colind, vals, L=None):
        L = L if L is not None else lib()
        sizes = np.array(list(sizes) + [1] * (3 - len(sizes)), dtype=np.int32)
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        colind = np.ascontiguousarray(colind, dtype=np.int32)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        p = L.orc_hmat_create(dim, _i(sizes), num_levels, max_rank, 1 if strong else 0,
                              _i(rowptr), _i(colind), _d(vals))
        return cls(p, L)

    def __del__(self):
        if getattr(self, "ptr", None) and getattr(self, "L", None) is not None:
            self.L.orc_hmat_free(self.ptr)
            self.ptr = None

    @property
    def n(self):
        return self.L.orc_hmat_height(self.ptr)

    def set_to_random(self):
        self.L.orc_hmat_set_to_random(self.ptr)

    def multiply(self, X, alpha=1.0, trans=False):
        X = np.asfortranarray(X, dtype=np.float64)
        if X.ndim == 1:
            X = X.reshape(-1, 1, order="F")
        k = X.shape[1]
        Y = np.zeros((self.n, k), order="F")
        self.L.orc_hmat_multiply_dense(self.ptr, 1 if trans else 0, alpha, _d(X), X.shape[0], k,
                                      0.0, _d(Y), self.n)
        return Y

    def to_dense(self):
        n = self.n
        out = np.zeros((n, n), order="F")
        self.L.orc_hmat_to_dense(self.ptr, _d(out))
        return out

    def count_blocks(self, max_levels=12):
        d = np.zeros(max_levels, dtype=np.int64)
        f = np.zeros(max_levels, dtype=np.int64)
        self.L.orc_hmat_count_blocks(self.ptr, d.ctypes.data_as(_lp), f.ctypes.data_as(_lp), max_levels)
        return d, f

    def rank_stats(self):
        out = np.zeros(3, dtype=np.int64)
        self.L.orc_hmat_rank_stats(self.ptr, out.ctypes.data_as(_lp))
        return {"sum_rank": int(out[0]), "max_rank": int(out[1]), "storage": int(out[2])}

    def matmul(self, other, alpha=1.0):
        return HMat(self.L.orc_hmat_multiply(alpha, self.ptr, other.ptr), self.L)

    def matmul_child(self, other, t, s, alpha=1.0):
        """Child (t, s) of alpha self other only (a timing sample)."""
        return HMat(self.L.orc_hmat_multiply_child(alpha, self.ptr, other.ptr, t, s), self.L)

    def matmul_child_row(self, other, t, alpha=1.0):
        """Children (t, *) of alpha self other only (a timing sample)."""
        return HMat(self.L.orc_hmat_multiply_child_row(alpha, self.ptr, other.ptr, t), self.L)

    def multiply_update(self, alpha, A, B, beta):
        """self := alpha A B + beta self (in place)."""
        self.L.orc_hmat_multiply_update(alpha, A.ptr, B.ptr, beta, self.ptr)
        return self

    def copy(self):
        return HMat(self.L.orc_hmat_copy(self.ptr), self.L)

    def direct_invert(self):
        """self := inv(self) by recursive block Gauss-Jordan (DirectInvert)."""
        self.L.orc_hmat_direct_invert(self.ptr)
        return self

    def transpose(self):
        return HMat(self.L.orc_hmat_transpose(self.ptr), self.L)

    def scale(self, a):
        self.L.orc_hmat_scale(self.ptr, a)

    def add_diag(self, a):
        self.L.orc_hmat_add_diag(self.ptr, a)

    def estimate_two_norm(self, theta=1.5, confidence=6.0):
        return self.L.orc_estimate_two_norm(self.ptr, theta, confidence)

    def schulz_invert(self, num_iterations=-1, theta=1.5, confidence=6.0, stop_tol=1e-4,
                      max_iter=30):
        hist = np.zeros(max(max_iter, 64))
        it = self.L.orc_hmat_schulz(self.ptr, num_iterations, theta, confidence, stop_tol,
                                   max_iter, _d(hist), len(hist))
        return it, [h for h in hist if h >= 0]

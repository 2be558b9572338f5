"""Python mirror of the reference's sequential H-matrix interface.

``HMat2d`` / ``HMat3d`` keep the method names, argument order and meaning of
``dmhm::HMat2d<Scalar>`` / ``HMat3d<Scalar>`` (include/dmhm/hmat2d.hpp:22-275,
include/dmhm/hmat3d.hpp) and ``hmat_tools::EstimateTwoNorm``
(include/dmhm/hmat_tools.hpp:1504-1544). Every call goes through the C-ABI of
libdmhm_amd.so to the gfx950 kernels; nothing is computed in Python.

Dense operands may be numpy arrays (host memory; copied through the library's
staging buffers) or torch tensors on the GPU (passed by device pointer). torch
bundles its own HIP runtime and the library binds whichever libamdhip64.so.7
the process loaded first: torch tensors work when torch is imported before the
library is loaded (both then share torch's runtime); imported afterwards, torch
does not see the GPU.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, DMHM_F64, DMHM_F32, DMHM_HOST, DMHM_DEVICE


def build_natural_to_hierarchical_map(sizes, num_levels):
    """HMat2d/HMat3d::BuildNaturalToHierarchicalMap (HMat2d.cpp:69-87)."""
    sizes = np.ascontiguousarray(sizes, dtype=np.int32)
    out = np.zeros(int(np.prod(sizes)), dtype=np.int32)
    check(_lib.load().dmhm_build_natural_to_hierarchical_map(
        len(sizes), sizes.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), num_levels,
        out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
    return out


def build_natural_to_hierarchical_map_device(sizes, num_levels):
    """The natural -> hierarchical map computed on the GPU."""
    _lib.init()
    sizes = np.ascontiguousarray(sizes, dtype=np.int32)
    out = np.zeros(int(np.prod(sizes)), dtype=np.int32)
    check(_lib.load().dmhm_build_natural_to_hierarchical_map_device(
        len(sizes), sizes.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), num_levels,
        out.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
    return out


def geometry_block_counts(sizes, num_levels, strongly_admissible, max_levels=12):
    """Dense / low-rank block counts per level of the block tree (host only)."""
    sizes = np.ascontiguousarray(sizes, dtype=np.int32)
    d = np.zeros(max_levels, dtype=np.int64)
    f = np.zeros(max_levels, dtype=np.int64)
    lp = ctypes.POINTER(ctypes.c_longlong)
    check(_lib.load().dmhm_geometry_block_counts(len(sizes), sizes.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                 num_levels, 1 if strongly_admissible else 0,
                                                 d.ctypes.data_as(lp), f.ctypes.data_as(lp), max_levels))
    return d, f


def geometry_leaves(sizes, num_levels, strongly_admissible):
    """Leaf blocks (DFS order) as an int array of rows
    (type: 0 low rank / 1 dense, level, row_off, rows, col_off, cols, index)."""
    sizes = np.ascontiguousarray(sizes, dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    cnt = ctypes.c_int()
    check(_lib.load().dmhm_geometry_leaves(len(sizes), sizes.ctypes.data_as(ip), num_levels,
                                           1 if strongly_admissible else 0, None, 0, ctypes.byref(cnt)))
    out = np.zeros((cnt.value, 7), dtype=np.int32)
    check(_lib.load().dmhm_geometry_leaves(len(sizes), sizes.ctypes.data_as(ip), num_levels,
                                           1 if strongly_admissible else 0, out.ctypes.data_as(ip), cnt.value,
                                           ctypes.byref(cnt)))
    return out


def geometry_product_terms(sizes, num_levels, strongly_admissible):
    """Term counts of one H*H product on this geometry (host only): dict with
    sketch (NODE*NODE into admissible), dd_to_lowrank, dd_to_dense, lowrank,
    pushdown."""
    sizes = np.ascontiguousarray(sizes, dtype=np.int32)
    out = np.zeros(5, dtype=np.int64)
    check(_lib.load().dmhm_geometry_product_terms(len(sizes), sizes.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                  num_levels, 1 if strongly_admissible else 0,
                                                  out.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))))
    return dict(zip(["sketch", "dd_to_lowrank", "dd_to_dense", "lowrank", "pushdown"], out.tolist()))


def _np_dtype(prec):
    return np.float64 if prec == DMHM_F64 else np.float32


class _HMatBase:
    dim = None

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle) if not isinstance(handle, ctypes.c_void_p) else handle

    # -- construction ----------------------------------------------------
    @classmethod
    def from_csr(cls, sizes, num_levels, max_rank, strongly_admissible, row_offsets, column_indices,
                 nonzeros, precision=DMHM_F64):
        """HMat2d(const Sparse<Scalar>&, numLevels, maxRank, strong, xSize, ySize)."""
        _lib.init()
        sizes = np.ascontiguousarray(sizes, dtype=np.int32)
        if len(sizes) != cls.dim:
            raise ValueError("expected %d grid sizes" % cls.dim)
        rp = np.ascontiguousarray(row_offsets, dtype=np.int32)
        ci = np.ascontiguousarray(column_indices, dtype=np.int32)
        nz = np.ascontiguousarray(nonzeros, dtype=np.float64)
        ip = ctypes.POINTER(ctypes.c_int)
        out = ctypes.c_void_p()
        check(_lib.load().dmhm_hmat_create_from_csr(
            cls.dim, sizes.ctypes.data_as(ip), num_levels, max_rank, 1 if strongly_admissible else 0,
            precision, rp.ctypes.data_as(ip), ci.ctypes.data_as(ip),
            nz.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(out)))
        return cls(out)

    @classmethod
    def zeros(cls, sizes, num_levels, max_rank, strongly_admissible, precision=DMHM_F64):
        _lib.init()
        sizes = np.ascontiguousarray(sizes, dtype=np.int32)
        out = ctypes.c_void_p()
        check(_lib.load().dmhm_hmat_create_zero(cls.dim, sizes.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                num_levels, max_rank, 1 if strongly_admissible else 0,
                                                precision, ctypes.byref(out)))
        return cls(out)

    @classmethod
    def from_lowrank(cls, sizes, num_levels, max_rank, strongly_admissible, U, V, precision=DMHM_F64):
        """HMat2d(const LowRank<Scalar>&, numLevels, maxRank, strong, xSize, ySize)
        (src/HMat2d/HMat2d.cpp:124-142, ImportLowRank :811-904): F = U V^T with
        rows in the matrix's hierarchical ordering; each admissible block takes
        the matching row slices of U and V exactly (rank r, not truncated),
        each dense leaf their product."""
        _lib.init()
        sizes = np.ascontiguousarray(sizes, dtype=np.int32)
        if len(sizes) != cls.dim:
            raise ValueError("expected %d grid sizes" % cls.dim)
        n = int(np.prod(sizes))
        U = np.asfortranarray(U, dtype=np.float64)
        V = np.asfortranarray(V, dtype=np.float64)
        if U.ndim != 2 or V.ndim != 2 or U.shape[0] != n or V.shape[0] != n or U.shape[1] != V.shape[1]:
            raise ValueError("U and V must be n x r with n = %d" % n)
        dp = ctypes.POINTER(ctypes.c_double)
        out = ctypes.c_void_p()
        check(_lib.load().dmhm_hmat_create_from_lowrank(
            cls.dim, sizes.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), num_levels, max_rank,
            1 if strongly_admissible else 0, precision, U.shape[1], U.ctypes.data_as(dp), V.ctypes.data_as(dp),
            ctypes.byref(out)))
        return cls(out)

    @classmethod
    def Unpack(cls, packed):
        """Rebuild an H-matrix from Pack()'s bytes (the HMat2d(const Vector<byte>&)
        constructor, src/HMat2d/Pack-incl.hpp)."""
        _lib.init()
        buf = bytes(packed)
        out = ctypes.c_void_p()
        check(_lib.load().dmhm_hmat_unpack(ctypes.c_char_p(buf), len(buf), ctypes.byref(out)))
        return cls(out)

    def PackedSize(self):
        n = ctypes.c_uint64()
        check(_lib.load().dmhm_hmat_packed_size(self._h, ctypes.byref(n)))
        return n.value

    def Pack(self):
        """Packed byte image (HMat2d::Pack, src/HMat2d/Pack-incl.hpp:17-364)."""
        n = self.PackedSize()
        buf = ctypes.create_string_buffer(max(n, 1))
        w = ctypes.c_uint64()
        check(_lib.load().dmhm_hmat_pack(self._h, buf, n, ctypes.byref(w)))
        return buf.raw[:w.value]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.dmhm_hmat_destroy(h)
            self._h = None

    # -- AbstractHMat fulfilments (abstract_hmat.hpp:20-93) -----------------
    def info(self):
        inf = _lib.Info()
        check(_lib.load().dmhm_hmat_info(self._h, ctypes.byref(inf)))
        return inf

    def Height(self):
        return self.info().n

    Width = Height

    def NumLevels(self):
        return self.info().numLevels

    def MaxRank(self):
        return self.info().maxRank

    def StronglyAdmissible(self):
        return bool(self.info().strong)

    @property
    def precision(self):
        return self.info().precision

    def block_counts(self, max_levels=12):
        d = np.zeros(max_levels, dtype=np.int64)
        f = np.zeros(max_levels, dtype=np.int64)
        lp = ctypes.POINTER(ctypes.c_longlong)
        check(_lib.load().dmhm_hmat_block_counts(self._h, d.ctypes.data_as(lp), f.ctypes.data_as(lp),
                                                 max_levels))
        return d, f

    def ranks(self):
        n = self.info().numLowRank
        out = np.zeros(n, dtype=np.int32)
        check(_lib.load().dmhm_hmat_ranks(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), int(n)))
        return out

    def matvec_bytes(self, k):
        b = ctypes.c_double()
        f = ctypes.c_double()
        check(_lib.load().dmhm_hmat_matvec_bytes(self._h, k, ctypes.byref(b), ctypes.byref(f)))
        return b.value, f.value

    # -- products ------------------------------------------------------------
    def _mult(self, op, alpha, X, beta, Y):
        prec = self.precision
        n = self.Height()
        if hasattr(X, "data_ptr"):          # torch tensors on the GPU
            import torch
            want = torch.float64 if prec == DMHM_F64 else torch.float32
            if X.dtype != want or (Y is not None and Y.dtype != want):
                raise TypeError("operand dtype must match the H-matrix precision")
            Xc = X if X.dim() == 2 else X.reshape(-1, 1)
            k = Xc.shape[1]
            if Y is None:
                Y = torch.zeros((k, n), dtype=want, device=X.device).t()
                beta = 0.0
            Yc = Y if Y.dim() == 2 else Y.reshape(-1, 1)
            # column-major views: stride(0) == 1
            if Xc.stride(0) != 1 or Yc.stride(0) != 1:
                raise ValueError("torch operands must be column-major (use .t().contiguous().t())")
            ldx = Xc.stride(1) if k > 1 else n
            ldy = Yc.stride(1) if k > 1 else n
            # The library works on its own stream: finish torch's pending work on
            # X and Y (e.g. the fill of a fresh Y) before the library reads or
            # writes them. The call returns after the library's work is done.
            torch.cuda.current_stream(X.device).synchronize()
            check(_lib.load().dmhm_hmat_multiply_dense(self._h, op.encode(), alpha, ctypes.c_void_p(Xc.data_ptr()),
                                                       ldx, k, beta, ctypes.c_void_p(Yc.data_ptr()), ldy,
                                                       DMHM_DEVICE))
            return Y
        dt = _np_dtype(prec)
        vec = np.ndim(X) == 1
        Xc = np.asfortranarray(np.reshape(X, (n, -1)), dtype=dt)
        k = Xc.shape[1]
        if Y is None:
            Yc = np.zeros((n, k), dtype=dt, order="F")
            beta = 0.0
        else:
            if not isinstance(Y, np.ndarray) or Y.size != n * k:
                raise ValueError("Y must be a numpy array with n x k entries")
            Yc = np.asfortranarray(np.reshape(Y, (n, -1)), dtype=dt)
        check(_lib.load().dmhm_hmat_multiply_dense(self._h, op.encode(), alpha,
                                                   Xc.ctypes.data_as(ctypes.c_void_p), n, k, beta,
                                                   Yc.ctypes.data_as(ctypes.c_void_p), n, DMHM_HOST))
        if Y is not None:
            # The reference updates y in place (MultiplyVector-incl.hpp:12-64):
            # write the result back when Y was not already a Fortran-ordered
            # array of the matrix's precision.
            if not np.shares_memory(Yc, Y):
                Y[...] = Yc.reshape(Y.shape)
            return Y
        return Yc.reshape(-1) if vec else Yc

    def Multiply(self, alpha, B, beta=None, C=None, **hh_options):
        """Multiply(alpha, x|X, [beta, y|Y]) -> alpha H X (+ beta Y), or
        Multiply(alpha, HMat B, [beta, HMat C], **hh_options) -> C := alpha H B
        (+ beta C) (MultiplyHMat-incl.hpp:12-380)."""
        if isinstance(B, _HMatBase):
            return self.MultiplyHMat(alpha, B, beta, C, **hh_options)
        return self._mult("N", alpha, B, 0.0 if beta is None else beta, C)

    def TransposeMultiply(self, alpha, B, beta=None, C=None):
        return self._mult("T", alpha, B, 0.0 if beta is None else beta, C)

    AdjointMultiply = TransposeMultiply

    def MultiplyHMat(self, alpha, B, beta=None, C=None, oversample=4, tol_mode=0, rel_tol=1e-10, seed=1,
                     rank_model=None, rank_margin=1.0):
        """C := alpha A B (new matrix), or C := alpha A B + beta C in place when
        ``C`` is given (MultiplyHMat-incl.hpp:12-380). With ``rank_model`` (a
        RankModel) and tol_mode=1 the sketch of every admissible block is
        sized by the model's predicted rank + rank_margin + oversample."""
        o = _hh_opts(oversample, tol_mode, rel_tol, seed, rank_model, rank_margin)
        if C is not None:
            check(_lib.load().dmhm_hmat_multiply_hmat_update(alpha, self._h, B._h, 0.0 if beta is None else beta,
                                                             C._h, ctypes.byref(o)))
            return C
        out = ctypes.c_void_p()
        check(_lib.load().dmhm_hmat_multiply_hmat(alpha, self._h, B._h, ctypes.byref(out), ctypes.byref(o)))
        return type(self)(out)

    def UpdateWith(self, alpha, B, oversample=4, tol_mode=0, rel_tol=1e-10, seed=1):
        """A := A + alpha B with recompression of admissible blocks
        (src/HMat2d/HMat2d.cpp:746-784)."""
        o = _hh_opts(oversample, tol_mode, rel_tol, seed)
        check(_lib.load().dmhm_hmat_update_with(self._h, alpha, B._h, ctypes.byref(o)))
        return self

    # -- whole-matrix routines (HMat2d.cpp:331-744) --------------------------
    def Copy(self):
        out = ctypes.c_void_p()
        check(_lib.load().dmhm_hmat_copy(self._h, ctypes.byref(out)))
        return type(self)(out)

    def Transpose(self):
        out = ctypes.c_void_p()
        check(_lib.load().dmhm_hmat_transpose(self._h, ctypes.byref(out)))
        return type(self)(out)

    Adjoint = Transpose

    def Convert(self, precision):
        out = ctypes.c_void_p()
        check(_lib.load().dmhm_hmat_convert(self._h, precision, ctypes.byref(out)))
        return type(self)(out)

    def Scale(self, alpha):
        check(_lib.load().dmhm_hmat_scale(self._h, alpha))

    def SetToIdentity(self):
        check(_lib.load().dmhm_hmat_set_to_identity(self._h))

    def AddConstantToDiagonal(self, alpha):
        check(_lib.load().dmhm_hmat_add_constant_to_diagonal(self._h, alpha))

    def SetToRandom(self, seed=1):
        check(_lib.load().dmhm_hmat_set_to_random(self._h, seed))

    def EstimateTwoNorm(self, theta=1.5, confidence=6.0, seed=1):
        est = ctypes.c_double()
        check(_lib.load().dmhm_hmat_estimate_two_norm(self._h, theta, confidence, seed, ctypes.byref(est)))
        return est.value

    def SchulzInvert(self, numIterations=-1, theta=1.5, confidence=6.0, stopTol=1e-4, maxIterations=30,
                     mixed=False, switchTol=1e-3, oversample=4, tol_mode=0, rel_tol=1e-10, seed=1,
                     raise_on_noconv=False, stall_stop=False, graph=True, bf16=False, bf16Tol=0.5,
                     rank_model=None, rank_margin=1.0, start="transpose"):
        """A := inv(A) by Schulz iteration (Invert-incl.hpp:128-196). Returns the
        report. graph=True replays captured HIP graphs for the steady steps;
        mixed=True runs FP32 steps until the estimate is below switchTol, with
        BF16 tile operands while it is at least bf16Tol when bf16=True;
        rank_model (with tol_mode=1) sizes every product's sketches.
        start="transpose" is the reference's X_0 = alpha A^T; start="direct"
        takes X_0 from DirectInvert, which the iteration then refines."""
        if start not in ("transpose", "direct"):
            raise ValueError("start must be 'transpose' or 'direct'")
        o = _lib.SchulzOpts(numIterations, maxIterations, theta, confidence, stopTol, 1 if mixed else 0,
                            switchTol, _hh_opts(oversample, tol_mode, rel_tol, seed, rank_model, rank_margin),
                            1 if stall_stop else 0, 1 if graph else 0, 1 if bf16 else 0, bf16Tol,
                            1 if start == "direct" else 0)
        rep = _lib.SchulzReport()
        code = _lib.load().dmhm_hmat_schulz_invert(self._h, ctypes.byref(o), ctypes.byref(rep))
        if code == 4 and not raise_on_noconv:
            return rep
        check(code)
        return rep

    def DirectInvert(self, oversample=4, tol_mode=0, rel_tol=1e-10, seed=1):
        """A := inv(A) by recursive block Gauss-Jordan elimination
        (HMat2d::DirectInvert, src/HMat2d/Invert-incl.hpp:17-125); every
        sub-matrix product is truncated to maxRank (or by rel_tol). Raises
        DmhmError (DMHM_ESINGULAR) on a zero pivot."""
        o = _hh_opts(oversample, tol_mode, rel_tol, seed)
        check(_lib.load().dmhm_hmat_direct_invert(self._h, ctypes.byref(o)))
        return self

    def direct_invert_program(self, oversample=4):
        """(sub-matrix products, merged plans, host build seconds) of the
        cached direct-inversion program of this geometry and rank."""
        p, q = ctypes.c_longlong(), ctypes.c_longlong()
        s = ctypes.c_double()
        check(_lib.load().dmhm_direct_invert_program_stats(self._h, oversample, ctypes.byref(p), ctypes.byref(q),
                                                            ctypes.byref(s)))
        return p.value, q.value, s.value

    def profile_matvec(self, k=1, reps=20):
        """HIP-event timing of H*X on the library stream: (ms, bytes) arrays."""
        ms = np.zeros(3)
        by = np.zeros(2)
        dp = ctypes.POINTER(ctypes.c_double)
        check(_lib.load().dmhm_profile_matvec(self._h, k, reps, ms.ctypes.data_as(dp), by.ctypes.data_as(dp)))
        return ms, by

    def profile_hh(self, B, alpha=1.0, oversample=4, tol_mode=0, rel_tol=1e-10, seed=1):
        ms = np.zeros(6)
        fl = np.zeros(2)
        dp = ctypes.POINTER(ctypes.c_double)
        o = _hh_opts(oversample, tol_mode, rel_tol, seed)
        check(_lib.load().dmhm_profile_hh(alpha, self._h, B._h, ctypes.byref(o), ms.ctypes.data_as(dp),
                                          fl.ctypes.data_as(dp)))
        return ms, fl

    @staticmethod
    def last_hh_phases(cap=64):
        """Per-phase (kind, ms, flops) of the last profile_hh call."""
        ms = np.zeros(cap)
        fl = np.zeros(cap)
        kinds = np.zeros(cap, dtype=np.int32)
        cnt = ctypes.c_int()
        dp = ctypes.POINTER(ctypes.c_double)
        check(_lib.load().dmhm_profile_hh_phases(cap, ms.ctypes.data_as(dp), fl.ctypes.data_as(dp),
                                                 kinds.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                                 ctypes.byref(cnt)))
        n = min(cnt.value, cap)
        return [(int(kinds[i]), float(ms[i]), float(fl[i])) for i in range(n)]

    def rank_features(self):
        """Features (numLowRank x DMHM_RANK_FEATURES, float32) of the admissible
        blocks of a product whose left operand is this matrix: the input of
        the rank model in that product."""
        n = int(self.info().numLowRank)
        out = np.zeros((n, _lib.DMHM_RANK_FEATURES), dtype=np.float32)
        check(_lib.load().dmhm_hmat_rank_features(self._h, out.ctypes.data_as(_lib._fp), out.size))
        return out

    def to_dense(self):
        n = self.Height()
        out = np.zeros((n, n), order="F")
        check(_lib.load().dmhm_hmat_to_dense(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n))
        return out


def _hh_opts(oversample, tol_mode, rel_tol, seed, rank_model=None, rank_margin=1.0):
    return _lib.HHOpts(oversample, tol_mode, rel_tol, seed, rank_model._h if rank_model is not None else None,
                       rank_margin)


def last_product_stats():
    """The last product that used a rank model: dict of admissible blocks, sum
    of the predicted sketch widths, blocks recomputed at full width, and
    blocks x full width."""
    out = np.zeros(4, dtype=np.int64)
    check(_lib.load().dmhm_last_product_stats(out.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))))
    return dict(zip(["blocks", "width_sum", "fallbacks", "full_width_sum"], out.tolist()))


class RankModel:
    """The learned rank model (3-layer perceptron over DMHM_RANK_FEATURES block
    features, two hidden layers of 64 ReLU units), trained and evaluated by the
    library's GPU kernels."""

    def __init__(self, seed=0):
        _lib.init()
        h = ctypes.c_void_p()
        check(_lib.load().dmhm_rank_model_create(seed, ctypes.byref(h)))
        self._h = h
        self.num_params = _lib.load().dmhm_rank_model_num_params()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.dmhm_rank_model_destroy(h)
            self._h = None

    @property
    def params(self):
        p = np.zeros(self.num_params, dtype=np.float32)
        check(_lib.load().dmhm_rank_model_get_params(self._h, p.ctypes.data_as(_lib._fp), p.size))
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(_lib.load().dmhm_rank_model_set_params(self._h, p.ctypes.data_as(_lib._fp), p.size))

    @staticmethod
    def _xt(X, t=None):
        X = np.ascontiguousarray(X, dtype=np.float32)
        if X.ndim != 2 or X.shape[1] != _lib.DMHM_RANK_FEATURES:
            raise ValueError("X must be n x %d" % _lib.DMHM_RANK_FEATURES)
        if t is None:
            return X, None
        t = np.ascontiguousarray(t, dtype=np.float32)
        if t.shape != (X.shape[0],):
            raise ValueError("t must have one target per row of X")
        return X, t

    def gradient(self, X, t):
        """(loss, dloss/dparams) of mean (y - t)^2 on the GPU."""
        X, t = self._xt(X, t)
        g = np.zeros(self.num_params, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        check(_lib.load().dmhm_rank_model_gradient(self._h, X.ctypes.data_as(_lib._fp), t.ctypes.data_as(_lib._fp),
                                                   X.shape[0], g.ctypes.data_as(_lib._fp),
                                                   loss.ctypes.data_as(_lib._fp)))
        return float(loss[0]), g

    def predict(self, X):
        X, _ = self._xt(X)
        y = np.zeros(X.shape[0], dtype=np.float32)
        check(_lib.load().dmhm_rank_model_predict(self._h, X.ctypes.data_as(_lib._fp), X.shape[0],
                                                  y.ctypes.data_as(_lib._fp)))
        return y

    def train(self, X, t, steps, batch=1024, lr=1e-3, seed=0):
        """Adam on the GPU; returns the per-step mini-batch losses."""
        X, t = self._xt(X, t)
        losses = np.zeros(max(steps, 1), dtype=np.float32)
        check(_lib.load().dmhm_rank_model_train(self._h, X.ctypes.data_as(_lib._fp), t.ctypes.data_as(_lib._fp),
                                                X.shape[0], steps, batch, lr, seed,
                                                losses.ctypes.data_as(_lib._fp)))
        return losses[:steps]


class HMat2d(_HMatBase):
    dim = 2


class HMat3d(_HMatBase):
    dim = 3


def _fortran(a):
    return np.asfortranarray(a, dtype=np.float64)


def RoundedUpdate(max_rank, alpha, A, beta, B):
    """hmat_tools::RoundedUpdate (src/hmat_tools/RoundedUpdate.cpp:14-162):
    returns the low-rank factors (U, V) of alpha A + beta B truncated to
    max_rank, where A = (AU, AV) and B = (BU, BV) are low-rank pairs.
    Exact concatenation [beta B, alpha A] when the summed rank fits."""
    _lib.init()
    AU, AV = map(_fortran, A)
    BU, BV = map(_fortran, B)
    m, n = AU.shape[0], AV.shape[0]
    if BU.shape[0] != m or BV.shape[0] != n or AU.shape[1] != AV.shape[1] or BU.shape[1] != BV.shape[1]:
        raise ValueError("nonconformal low-rank operands")
    rA, rB = AU.shape[1], BU.shape[1]
    r = rA + rB
    outU = np.zeros((m, max(r, 1)), order="F")
    outV = np.zeros((n, max(r, 1)), order="F")
    keep = ctypes.c_int()
    dp = ctypes.POINTER(ctypes.c_double)
    check(_lib.load().dmhm_rounded_update(max_rank, alpha, m, n, rA, AU.ctypes.data_as(dp), AV.ctypes.data_as(dp),
                                          beta, rB, BU.ctypes.data_as(dp), BV.ctypes.data_as(dp),
                                          outU.ctypes.data_as(dp), outV.ctypes.data_as(dp), ctypes.byref(keep)))
    return outU[:, :keep.value].copy(order="F"), outV[:, :keep.value].copy(order="F")


def Compress(max_rank, D):
    """hmat_tools::Compress of a dense block (src/hmat_tools/Compress.cpp:14-53):
    returns (U, V) with D ~= U V^T truncated to max_rank by an SVD
    (V carries the singular values)."""
    _lib.init()
    D = _fortran(D)
    m, n = D.shape
    k = min(m, n)
    outU = np.zeros((m, max(k, 1)), order="F")
    outV = np.zeros((n, max(k, 1)), order="F")
    keep = ctypes.c_int()
    dp = ctypes.POINTER(ctypes.c_double)
    check(_lib.load().dmhm_compress_dense(max_rank, m, n, D.ctypes.data_as(dp), outU.ctypes.data_as(dp),
                                          outV.ctypes.data_as(dp), ctypes.byref(keep)))
    return outU[:, :keep.value].copy(order="F"), outV[:, :keep.value].copy(order="F")


def EstimateTwoNorm(A, theta=1.5, confidence=6.0, seed=1):
    """hmat_tools::EstimateTwoNorm (include/dmhm/hmat_tools.hpp:1504-1544)."""
    return A.EstimateTwoNorm(theta, confidence, seed)


class HBatch:
    """Independent H-matrices worked on together (dmhm_hbatch_*): members of
    one class share one launch per phase, members whose products fill the
    chip alone run one after another. No reference counterpart: the
    reference inverts such operators one by one
    (src/HMat3d/Invert-incl.hpp:127-188)."""

    def __init__(self, handle, cls_of):
        self._h = ctypes.c_void_p(handle) if not isinstance(handle, ctypes.c_void_p) else handle
        self._cls = list(cls_of)

    @classmethod
    def create(cls, mats):
        _lib.init()
        arr = (ctypes.c_void_p * len(mats))(*[m._h.value for m in mats])
        out = ctypes.c_void_p()
        check(_lib.load().dmhm_hbatch_create(len(mats), arr, ctypes.byref(out)))
        return cls(out, [type(m) for m in mats])

    def __del__(self):
        try:
            if self._h and self._h.value:
                _lib.load().dmhm_hbatch_destroy(self._h)
                self._h = ctypes.c_void_p()
        except Exception:
            pass

    def __len__(self):
        return len(self._cls)

    def member(self, i):
        """A new handle holding a copy of member i."""
        out = ctypes.c_void_p()
        check(_lib.load().dmhm_hbatch_member(self._h, i, ctypes.byref(out)))
        return self._cls[i](out)

    def Multiply(self, alpha, B, oversample=4, tol_mode=0, rel_tol=1e-10, seed=1):
        """C_i = alpha A_i B_i for every member (HMat2d::Multiply per member)."""
        o = _hh_opts(oversample, tol_mode, rel_tol, seed)
        out = ctypes.c_void_p()
        check(_lib.load().dmhm_hbatch_multiply_hmat(alpha, self._h, B._h, ctypes.byref(o), ctypes.byref(out)))
        return HBatch(out, self._cls)

    def EstimateTwoNorm(self, theta=1.5, confidence=6.0, seed=1):
        est = np.zeros(len(self))
        check(_lib.load().dmhm_hbatch_estimate_two_norm(self._h, theta, confidence, seed,
                                                         est.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return est

    def SchulzInvert(self, numIterations=-1, theta=1.5, confidence=6.0, stopTol=1e-4, maxIterations=30,
                     oversample=4, tol_mode=0, rel_tol=1e-10, seed=1, stall_stop=False, graph=True,
                     raise_on_noconv=False):
        """Every member := its inverse by Schulz iteration, each stopping on
        its own residual estimate. Returns the per-member reports."""
        o = _lib.SchulzOpts(numIterations, maxIterations, theta, confidence, stopTol, 0, 1e-3,
                            _hh_opts(oversample, tol_mode, rel_tol, seed), 1 if stall_stop else 0,
                            1 if graph else 0, 0, 0.5, 0)
        reps = (_lib.SchulzReport * len(self))()
        code = _lib.load().dmhm_hbatch_schulz_invert(self._h, ctypes.byref(o), reps)
        if code == 4 and not raise_on_noconv:
            return list(reps)
        check(code)
        return list(reps)


def check_direct_invert_program(dim, sizes, num_levels, max_rank, strong, oversample=4):
    """Host only: lowers the direct inversion of a geometry at max_rank and
    checks every task's bounds. Returns the program's counts."""
    sz = np.zeros(3, dtype=np.int32)
    sz[:len(sizes)] = sizes
    st = np.zeros(7, dtype=np.int64)
    check(_lib.load().dmhm_check_direct_invert_program(dim, sz.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), num_levels,
                                                        max_rank, 1 if strong else 0, oversample,
                                                        st.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))))
    return dict(products=int(st[0]), plans=int(st[1]), steps=int(st[2]), leaf_inversions=int(st[3]),
                workspace=int(st[4]), phases=int(st[5]), identity_phases=int(st[6]))

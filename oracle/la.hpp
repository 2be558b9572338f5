// ORACLE — test infrastructure only. Nothing in the product links, loads or
// calls this code; tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg use it as the checker.
//
// Dense column-major containers and the LAPACK/BLAS substitutes the reference
// calls on small blocks (include/dmhm/core/blas.hpp:596-820 Gemm,
// include/dmhm/core/lapack.hpp:424-530 QR, :534-640 PivotedQR, :647-780 ApplyQ,
// :784-880 FormQ, :886-1080 SVD). The reference binds Fortran LAPACK (no pinned
// version, not vendored); these are textbook restatements (Householder QR,
// column-pivoted Householder QR, one-sided Jacobi SVD) pinned in tests against
// numpy.linalg (LAPACK) on the same inputs.
#pragma once
#include <cmath>
#include <cstring>
#include <vector>
#include <algorithm>
#include <stdexcept>

namespace orc {

// Owning column-major matrix with ldim == height (reference Dense<Scalar>,
// include/dmhm/core/dense.hpp:24-88; views are replaced by explicit offsets).
struct Dense {
    int m = 0, n = 0;
    std::vector<double> a;
    Dense() {}
    Dense(int m_, int n_) : m(m_), n(n_), a((size_t)m_ * n_, 0.0) {}
    void resize(int m_, int n_) { m = m_; n = n_; a.assign((size_t)m_ * n_, 0.0); }
    double& operator()(int i, int j) { return a[i + (size_t)j * m]; }
    double operator()(int i, int j) const { return a[i + (size_t)j * m]; }
    double* col(int j) { return a.data() + (size_t)j * m; }
    const double* col(int j) const { return a.data() + (size_t)j * m; }
    // Keep the leading columns (reference Dense::Resize shrink semantics,
    // include/dmhm/core/dense.hpp:256-282, used for in-place truncation).
    void keep_cols(int r) { a.resize((size_t)m * r); n = r; }
};

// Low-rank A = U V^T, transpose (not adjoint) convention
// (include/dmhm/core/low_rank.hpp:23-35).
struct LowRank {
    Dense U, V;
    int rank() const { return U.n; }
    int height() const { return U.m; }
    int width() const { return V.m; }
};

// C(m x n) = alpha op(A) op(B) + beta C, plain loops (BLAS Gemm semantics).
void gemm(char ta, char tb, int m, int n, int k, double alpha,
          const double* A, int lda, const double* B, int ldb,
          double beta, double* C, int ldc);

// In-place inverse of an n x n matrix (hmat_tools::Invert on a Dense,
// src/hmat_tools/Invert.cpp:14-42, there LAPACK getrf + getri): Gauss-Jordan
// elimination with partial pivoting. Throws on a zero pivot.
void invert(int n, double* A, int lda);

// Householder QR in place (geqrf semantics: R in the upper triangle,
// reflectors below with v(0)=1 implicit, tau per reflector).
void qr(int m, int n, double* A, int lda, double* tau);
// Column-pivoted Householder QR (geqp3 semantics; jpvt 0-based output).
void pivoted_qr(int m, int n, double* A, int lda, int* jpvt, double* tau);
// Form the first n columns of Q from k reflectors (orgqr semantics).
void form_q(int m, int n, int k, double* A, int lda, const double* tau);
// C(m x n) := Q C with Q from k reflectors stored in A (ormqr 'L','N').
void apply_q(int m, int n, int k, const double* A, int lda, const double* tau,
             double* C, int ldc);
// Economic SVD A = U diag(s) VT (gesvd 'S','S' semantics, s descending),
// via one-sided Jacobi on the columns (of A or A^T).
// U: m x min(m,n), VT: min(m,n) x n.
void svd(int m, int n, const double* A, int lda, double* U, double* s, double* VT);

} // namespace orc

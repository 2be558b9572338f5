// ORACLE — test infrastructure only (see la.hpp header).
//
// The dense routines of la.hpp bound to a Fortran BLAS/LAPACK (LP64), as the
// reference binds them (include/dmhm/core/blas.hpp:25-200,
// include/dmhm/core/lapack.hpp:21-360): dgemm, dgeqrf, dgeqp3, dorgqr,
// dormqr, dgesvd, dgetrf + dgetri (hmat_tools::Invert). Built into liboracle_lapack.so (with MKL's sequential
// layer) for the timed CPU baseline; tests check it against the
// self-contained liboracle.so on the same inputs.
#include "la.hpp"

extern "C" {
void dgemm_(const char* ta, const char* tb, const int* m, const int* n, const int* k, const double* alpha,
            const double* A, const int* lda, const double* B, const int* ldb, const double* beta, double* C,
            const int* ldc);
void dgeqrf_(const int* m, const int* n, double* A, const int* lda, double* tau, double* work,
             const int* lwork, int* info);
void dgeqp3_(const int* m, const int* n, double* A, const int* lda, int* jpvt, double* tau, double* work,
             const int* lwork, int* info);
void dorgqr_(const int* m, const int* n, const int* k, double* A, const int* lda, const double* tau,
             double* work, const int* lwork, int* info);
void dormqr_(const char* side, const char* trans, const int* m, const int* n, const int* k, const double* A,
             const int* lda, const double* tau, double* C, const int* ldc, double* work, const int* lwork,
             int* info);
void dgetrf_(const int* m, const int* n, double* A, const int* lda, int* ipiv, int* info);
void dgetri_(const int* n, double* A, const int* lda, const int* ipiv, double* work, const int* lwork, int* info);
void dgesvd_(const char* jobu, const char* jobvt, const int* m, const int* n, double* A, const int* lda,
             double* s, double* U, const int* ldu, double* VT, const int* ldvt, double* work, const int* lwork,
             int* info);
}

namespace orc {

namespace {
void check(int info, const char* what)
{
    if (info != 0) throw std::runtime_error(std::string(what) + " failed, info " + std::to_string(info));
}
// Workspace query then call (LAPACK's lwork = -1 protocol).
template <typename F>
void with_work(F f, const char* what)
{
    double q = 0;
    int info = 0, lw = -1;
    f(&q, &lw, &info);
    check(info, what);
    lw = std::max(1, (int)q);
    std::vector<double> w(lw);
    f(w.data(), &lw, &info);
    check(info, what);
}
} // namespace

void gemm(char ta, char tb, int m, int n, int k, double alpha, const double* A, int lda, const double* B,
          int ldb, double beta, double* C, int ldc)
{
    if (m == 0 || n == 0) return;
    lda = std::max(lda, 1); ldb = std::max(ldb, 1); ldc = std::max(ldc, 1);
    dgemm_(&ta, &tb, &m, &n, &k, &alpha, A, &lda, B, &ldb, &beta, C, &ldc);
}

void qr(int m, int n, double* A, int lda, double* tau)
{
    if (m == 0 || n == 0) return;
    with_work([&](double* w, int* lw, int* info) { dgeqrf_(&m, &n, A, &lda, tau, w, lw, info); }, "dgeqrf");
}

void pivoted_qr(int m, int n, double* A, int lda, int* jpvt, double* tau)
{
    for (int j = 0; j < n; ++j) jpvt[j] = 0;           // all columns free
    if (m > 0 && n > 0)
        with_work([&](double* w, int* lw, int* info) { dgeqp3_(&m, &n, A, &lda, jpvt, tau, w, lw, info); },
                  "dgeqp3");
    else
        for (int j = 0; j < n; ++j) jpvt[j] = j + 1;
    for (int j = 0; j < n; ++j) jpvt[j] -= 1;           // 0-based like la.cpp
}

void form_q(int m, int n, int k, double* A, int lda, const double* tau)
{
    if (m == 0 || n == 0) return;
    with_work([&](double* w, int* lw, int* info) { dorgqr_(&m, &n, &k, A, &lda, tau, w, lw, info); }, "dorgqr");
}

void apply_q(int m, int n, int k, const double* A, int lda, const double* tau, double* C, int ldc)
{
    if (m == 0 || n == 0 || k == 0) return;
    const char side = 'L', trans = 'N';
    with_work([&](double* w, int* lw, int* info) {
        dormqr_(&side, &trans, &m, &n, &k, A, &lda, tau, C, &ldc, w, lw, info);
    }, "dormqr");
}

void svd(int m, int n, const double* A, int lda, double* U, double* s, double* VT)
{
    const int k = std::min(m, n);
    if (k == 0) return;
    std::vector<double> a((size_t)m * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) a[i + (size_t)j * m] = A[i + (size_t)j * lda];
    const char job = 'S';
    int ldu = m, ldvt = k, lda2 = m;
    with_work([&](double* w, int* lw, int* info) {
        dgesvd_(&job, &job, &m, &n, a.data(), &lda2, s, U, &ldu, VT, &ldvt, w, lw, info);
    }, "dgesvd");
}

void invert(int n, double* A, int lda)
{
    if (n == 0) return;
    std::vector<int> ipiv(n);
    int info = 0;
    dgetrf_(&n, &n, A, &lda, ipiv.data(), &info);
    if (info > 0) throw std::runtime_error("singular matrix");
    check(info, "dgetrf");
    with_work([&](double* w, int* lw, int* inf) { dgetri_(&n, A, &lda, ipiv.data(), w, lw, inf); }, "dgetri");
}

} // namespace orc

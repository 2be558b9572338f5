// ORACLE — test infrastructure only (see la.hpp header).
#include "la.hpp"

namespace orc {

void gemm(char ta, char tb, int m, int n, int k, double alpha,
          const double* A, int lda, const double* B, int ldb,
          double beta, double* C, int ldc)
{
    for (int j = 0; j < n; ++j) {
        double* c = C + (size_t)j * ldc;
        if (beta == 0.0) for (int i = 0; i < m; ++i) c[i] = 0.0;
        else if (beta != 1.0) for (int i = 0; i < m; ++i) c[i] *= beta;
    }
    if (alpha == 0.0 || k == 0) return;
    const bool tA = (ta == 'T' || ta == 't' || ta == 'C' || ta == 'c');
    const bool tB = (tb == 'T' || tb == 't' || tb == 'C' || tb == 'c');
    for (int j = 0; j < n; ++j) {
        double* c = C + (size_t)j * ldc;
        for (int p = 0; p < k; ++p) {
            const double b = tB ? B[j + (size_t)p * ldb] : B[p + (size_t)j * ldb];
            if (b == 0.0) continue;
            const double ab = alpha * b;
            if (!tA) {
                const double* a = A + (size_t)p * lda;
                for (int i = 0; i < m; ++i) c[i] += ab * a[i];
            } else {
                for (int i = 0; i < m; ++i) c[i] += ab * A[p + (size_t)i * lda];
            }
        }
    }
}

// Generate an elementary reflector for x = [alpha; x(1:)] (dlarfg convention).
// The column is scaled by its largest magnitude first (as dlarfg rescales
// tiny columns): squares of entries below ~1e-154 would otherwise fall into
// the subnormal range, and the reflector would lose its orthogonality.
static void make_reflector(int len, double* x, double& tau)
{
    if (len <= 1) { tau = 0.0; return; }
    double amax = 0.0;
    for (int i = 0; i < len; ++i) amax = std::max(amax, std::fabs(x[i]));
    if (amax == 0.0) { tau = 0.0; return; }
    const double inv = 1.0 / amax;
    double xnorm2 = 0.0;
    for (int i = 1; i < len; ++i) xnorm2 += (x[i] * inv) * (x[i] * inv);
    if (xnorm2 == 0.0) { tau = 0.0; return; }
    const double alpha = x[0] * inv;
    double beta = std::sqrt(alpha * alpha + xnorm2);
    if (alpha > 0) beta = -beta;
    tau = (beta - alpha) / beta;
    const double scal = inv / (alpha - beta);
    for (int i = 1; i < len; ++i) x[i] *= scal;
    x[0] = beta * amax;
}

// A(j:m, c0:c1) := (I - tau v v^T) A(j:m, c0:c1), v = [1; A(j+1:m, j)].
static void apply_reflector(int m, int j, const double* v, double tau,
                            double* A, int lda, int c0, int c1)
{
    if (tau == 0.0) return;
    for (int c = c0; c < c1; ++c) {
        double* a = A + (size_t)c * lda;
        double w = a[j];
        for (int i = j + 1; i < m; ++i) w += v[i] * a[i];
        w *= tau;
        a[j] -= w;
        for (int i = j + 1; i < m; ++i) a[i] -= w * v[i];
    }
}

void qr(int m, int n, double* A, int lda, double* tau)
{
    const int k = std::min(m, n);
    for (int j = 0; j < k; ++j) {
        double* col = A + (size_t)j * lda;
        make_reflector(m - j, col + j, tau[j]);
        apply_reflector(m, j, col, tau[j], A, lda, j + 1, n);
    }
}

void pivoted_qr(int m, int n, double* A, int lda, int* jpvt, double* tau)
{
    const int k = std::min(m, n);
    for (int j = 0; j < n; ++j) jpvt[j] = j;
    for (int j = 0; j < k; ++j) {
        // Pick the remaining column with the largest trailing norm.
        int best = j; double bestNorm = -1.0;
        for (int c = j; c < n; ++c) {
            const double* a = A + (size_t)c * lda;
            double s = 0.0;
            for (int i = j; i < m; ++i) s += a[i] * a[i];
            if (s > bestNorm) { bestNorm = s; best = c; }
        }
        if (best != j) {
            double* a = A + (size_t)j * lda;
            double* b = A + (size_t)best * lda;
            for (int i = 0; i < m; ++i) std::swap(a[i], b[i]);
            std::swap(jpvt[j], jpvt[best]);
        }
        double* col = A + (size_t)j * lda;
        make_reflector(m - j, col + j, tau[j]);
        apply_reflector(m, j, col, tau[j], A, lda, j + 1, n);
    }
}

void form_q(int m, int n, int k, double* A, int lda, const double* tau)
{
    // Reflector vectors are read from a copy because Q overwrites A.
    std::vector<double> refl((size_t)m * k);
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < m; ++i) refl[i + (size_t)j * m] = A[i + (size_t)j * lda];
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) A[i + (size_t)j * lda] = (i == j) ? 1.0 : 0.0;
    for (int j = k - 1; j >= 0; --j)
        apply_reflector(m, j, refl.data() + (size_t)j * m, tau[j], A, lda, j, n);
}

void apply_q(int m, int n, int k, const double* A, int lda, const double* tau,
             double* C, int ldc)
{
    for (int j = k - 1; j >= 0; --j)
        apply_reflector(m, j, A + (size_t)j * lda, tau[j], C, ldc, 0, n);
}

// One-sided Jacobi SVD of a tall matrix W (m >= n), in place: on exit the
// columns of W are U*diag(s) (unsorted) and V holds the accumulated rotations.
static void jacobi_tall(int m, int n, double* W, double* V)
{
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) V[i + (size_t)j * n] = (i == j) ? 1.0 : 0.0;
    const double eps = 2.2e-16;
    for (int sweep = 0; sweep < 80; ++sweep) {
        bool rotated = false;
        for (int p = 0; p < n - 1; ++p) {
            for (int q = p + 1; q < n; ++q) {
                double* wp = W + (size_t)p * m;
                double* wq = W + (size_t)q * m;
                double alpha = 0, beta = 0, gamma = 0;
                for (int i = 0; i < m; ++i) {
                    alpha += wp[i] * wp[i]; beta += wq[i] * wq[i]; gamma += wp[i] * wq[i];
                }
                if (gamma == 0.0 || std::fabs(gamma) <= eps * std::sqrt(alpha * beta))
                    continue;
                rotated = true;
                const double zeta = (beta - alpha) / (2.0 * gamma);
                const double t = (zeta >= 0 ? 1.0 : -1.0) /
                                 (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t);
                const double s = c * t;
                for (int i = 0; i < m; ++i) {
                    const double x = wp[i], y = wq[i];
                    wp[i] = c * x - s * y; wq[i] = s * x + c * y;
                }
                double* vp = V + (size_t)p * n;
                double* vq = V + (size_t)q * n;
                for (int i = 0; i < n; ++i) {
                    const double x = vp[i], y = vq[i];
                    vp[i] = c * x - s * y; vq[i] = s * x + c * y;
                }
            }
        }
        if (!rotated) break;
    }
}

void svd(int m, int n, const double* A, int lda, double* U, double* s, double* VT)
{
    const int k = std::min(m, n);
    if (k == 0) return;
    const bool tall = (m >= n);
    const int mm = tall ? m : n, nn = tall ? n : m;
    std::vector<double> W((size_t)mm * nn), V((size_t)nn * nn);
    for (int j = 0; j < nn; ++j)
        for (int i = 0; i < mm; ++i)
            W[i + (size_t)j * mm] = tall ? A[i + (size_t)j * lda] : A[j + (size_t)i * lda];
    jacobi_tall(mm, nn, W.data(), V.data());
    std::vector<double> sig(nn);
    std::vector<int> order(nn);
    for (int j = 0; j < nn; ++j) {
        double t = 0;
        for (int i = 0; i < mm; ++i) t += W[i + (size_t)j * mm] * W[i + (size_t)j * mm];
        sig[j] = std::sqrt(t);
        order[j] = j;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sig[a] > sig[b]; });
    // Left vectors of the tall problem: W_j / s_j; right vectors: V_j.
    for (int r = 0; r < k; ++r) {
        const int j = order[r];
        s[r] = sig[j];
        const double inv = sig[j] > 0 ? 1.0 / sig[j] : 0.0;
        for (int i = 0; i < mm; ++i) {
            const double u = W[i + (size_t)j * mm] * inv;
            const double v = (i < nn) ? V[i + (size_t)j * nn] : 0.0;
            if (tall) {
                U[i + (size_t)r * m] = u;
                if (i < nn) VT[r + (size_t)i * k] = v;
            } else {
                // A^T = W V^T  =>  A = V W^T: U_A = V, VT_A = (W/s)^T
                VT[r + (size_t)i * k] = u;
                if (i < nn) U[i + (size_t)r * m] = v;
            }
        }
    }
}

void invert(int n, double* A, int lda)
{
    // Step k: bring the largest |a(i,k)|, i >= k, into row k; then
    // a(k,:) /= a(k,k) with a(k,k) := 1/a(k,k), and for i != k
    // a(i,:) -= a(i,k) a(k,:) with a(i,k) := -a(i,k)/a(k,k). The row swaps
    // are undone as column swaps in reverse order.
    std::vector<int> piv(n);
    auto a = [&](int i, int j) -> double& { return A[i + (size_t)j * lda]; };
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(a(i, k)) > std::fabs(a(p, k))) p = i;
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(a(k, j), a(p, j));
        const double d = a(k, k);
        if (d == 0.0) throw std::runtime_error("singular matrix");
        const double inv = 1.0 / d;
        for (int j = 0; j < n; ++j) a(k, j) *= inv;
        a(k, k) = inv;
        for (int i = 0; i < n; ++i) {
            if (i == k) continue;
            const double f = a(i, k);
            if (f == 0.0) continue;
            for (int j = 0; j < n; ++j)
                if (j != k) a(i, j) -= f * a(k, j);
            a(i, k) = -f * inv;
        }
    }
    for (int k = n - 1; k >= 0; --k)
        if (piv[k] != k)
            for (int i = 0; i < n; ++i) std::swap(a(i, k), a(i, piv[k]));
}

} // namespace orc

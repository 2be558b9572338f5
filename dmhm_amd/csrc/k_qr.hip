// gfx950 batched Householder QR of tall-skinny panels, Q formed in place
// (geqrf + orgqr semantics, one workgroup per panel, FP64 arithmetic).
#include "kcommon.hpp"

namespace dmhm {

// A column whose squared norm (diagonal entry included) is below this is left
// unreflected (tau = 0): its entries are below ~1e-140, and squares of such
// entries reach the subnormal range, where a reflector computed from them is
// no longer orthogonal and would corrupt the later columns. The entries it
// leaves below the diagonal are that small in absolute terms.
constexpr double kTinyColumn = 1e-280;

// ---------------------------------------------------------------------------
// Batched Householder QR (one workgroup per panel, FP64 arithmetic).
// ---------------------------------------------------------------------------

// dots[c] = sum_{i > j} A(i, j) * A(i, c) for c in [c0, k)
template <typename T>
__device__ void column_dots(const T* A, int m, int j, int c0, int k, double* dots,
                            double (*red)[16])
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int cb = c0; cb < k; cb += 16) {
        double p[16];
#pragma unroll
        for (int cc = 0; cc < 16; ++cc) p[cc] = 0.0;
        for (int i = j + 1 + tid; i < m; i += 256) {
            const double x = (double)A[i + (int64_t)j * m];
#pragma unroll
            for (int cc = 0; cc < 16; ++cc)
                if (cb + cc < k) p[cc] += x * (double)A[i + (int64_t)(cb + cc) * m];
        }
#pragma unroll
        for (int cc = 0; cc < 16; ++cc) {
            const double v = wave_sum(p[cc]);
            if (lane == 0) red[wave][cc] = v;
        }
        __syncthreads();
        if (tid < 16 && cb + tid < k) dots[cb + tid] = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(256) qr_kernel(Bases b, const QrTask* __restrict__ tasks)
{
    __shared__ double red[4][16];
    __shared__ double dots[64];
    __shared__ double wv[64];
    __shared__ double tau_s[64];
    __shared__ double scale_s;
    const QrTask tk = tasks[blockIdx.x];
    T* A = reinterpret_cast<T*>(b.val[tk.slot]) + tk.off;
    const int m = tk.m, kstat = tk.k;
    const int k = tk.krank >= 0 ? min(kstat, b.rank[tk.kslot][tk.krank]) : kstat, kk = min(m, k);
    const int tid = threadIdx.x;
    if (m <= 0 || k <= 0) return;
    // Factor.
    for (int j = 0; j < kk; ++j) {
        column_dots(A, m, j, j, k, dots, red);
        if (tid == 0) {
            const double alpha = (double)A[j + (int64_t)j * m];
            const double xn2 = dots[j];
            double tau = 0.0, scale = 0.0, beta = alpha;
            if (xn2 > 0.0 && alpha * alpha + xn2 > kTinyColumn) {
                beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
                tau = (beta - alpha) / beta;
                scale = 1.0 / (alpha - beta);
            }
            tau_s[j] = tau; scale_s = scale;
            A[j + (int64_t)j * m] = (T)beta;
        }
        __syncthreads();
        const double scale = scale_s, tau = tau_s[j];
        if (tid > j && tid < k) {
            const int c = tid;
            const double w = tau * ((double)A[j + (int64_t)c * m] + scale * dots[c]);
            wv[c] = w;
            A[j + (int64_t)c * m] = (T)((double)A[j + (int64_t)c * m] - w);
        }
        __syncthreads();
        for (int i = j + 1 + tid; i < m; i += 256) {
            const double v = (double)A[i + (int64_t)j * m] * scale;
            for (int c = j + 1; c < k; ++c)
                A[i + (int64_t)c * m] = (T)((double)A[i + (int64_t)c * m] - wv[c] * v);
            A[i + (int64_t)j * m] = (T)v;
        }
        __syncthreads();
    }
    // R (kstat x k, ld rld: upper triangle, rows >= kk zero).
    if (tk.roff >= 0) {
        T* R = reinterpret_cast<T*>(b.val[SLOT_WS]) + tk.roff;
        for (int e = tid; e < kstat * k; e += 256) {
            const int r = e % kstat, c = e / kstat;
            R[r + (int64_t)c * tk.rld] = (r <= c && r < kk) ? A[r + (int64_t)c * m] : T(0);
        }
    }
    __syncthreads();
    // Form Q in place (orgqr); columns >= kk are zero.
    for (int c = kk + (tid >> 6); c < k; c += 4)
        for (int i = tid & 63; i < m; i += 64) A[i + (int64_t)c * m] = T(0);
    __syncthreads();
    for (int j = kk - 1; j >= 0; --j) {
        const double tau = tau_s[j];
        if (j + 1 < k) column_dots(A, m, j, j + 1, k, dots, red);
        if (tid > j && tid < k) {
            const int c = tid;
            const double w = tau * ((double)A[j + (int64_t)c * m] + dots[c]);
            wv[c] = w;
            A[j + (int64_t)c * m] = (T)((double)A[j + (int64_t)c * m] - w);
        }
        __syncthreads();
        for (int i = j + 1 + tid; i < m; i += 256) {
            const double v = (double)A[i + (int64_t)j * m];
            for (int c = j + 1; c < k; ++c)
                A[i + (int64_t)c * m] = (T)((double)A[i + (int64_t)c * m] - wv[c] * v);
            A[i + (int64_t)j * m] = (T)(-tau * v);
        }
        for (int i = tid; i < j; i += 256) A[i + (int64_t)j * m] = T(0);
        if (tid == 0) A[j + (int64_t)j * m] = (T)(1.0 - tau);
        __syncthreads();
    }
}

// Wave sum with a uniform result: DPP sums inside each row of 16 lanes, then
// the four row sums through scalar registers (no LDS round trip).
__device__ inline double wave_sum_uniform(double v)
{
    v = group_sum<16>(v);
    return (bcast(v, 0) + bcast(v, 16)) + (bcast(v, 32) + bcast(v, 48));
}

// Householder QR of one m x k panel (m <= 64 RR, k <= KMAX), lane = row (rows
// lane + 64 r), the KMAX columns split over NWV waves (wave w holds columns
// [w CW, (w + 1) CW) of every row in registers). Step j: the wave that owns
// column j forms the reflector (its column's sum of squares below the
// diagonal is one wave reduction), publishes v and tau through LDS, and
// every wave then applies it to its own columns right of j (one wave
// reduction v^T a_c per column). One barrier per step; buffers alternate by
// step parity, so a step's writes never meet the reads of the step before.
// The explicit Q is accumulated backwards the same way (orgqr order). Same
// reflector convention and outputs as qr_kernel (Q over the panel, R to the
// workspace); the panel and R may be strided (leading dimensions ld, rld), as
// the row blocks of a TSQR are. FP64 arithmetic.
template <typename T, int KMAX, int NWV, int RR>
__global__ void __launch_bounds__(64 * NWV) qr_split_kernel(Bases b, const QrTask* __restrict__ tasks)
{
    constexpr int CW = KMAX / NWV;
    static_assert(KMAX % NWV == 0, "whole columns per wave");
    __shared__ double vbuf[2][64 * RR];
    __shared__ double tbuf[2];
    __shared__ double taus[KMAX];
    const QrTask tk = tasks[blockIdx.x];
    T* A = reinterpret_cast<T*>(b.val[tk.slot]) + tk.off;
    const int m = tk.m, kstat = tk.k, ld = tk.ld;
    const int k = tk.krank >= 0 ? min(kstat, b.rank[tk.kslot][tk.krank]) : kstat, kk = min(m, k);
    if (k <= 0) return;
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double a[CW][RR];
#pragma unroll
    for (int jj = 0; jj < CW; ++jj)
#pragma unroll
        for (int r = 0; r < RR; ++r) {
            const int c = w * CW + jj, i = lane + 64 * r;
            a[jj][r] = (i < m && c < k) ? (double)A[i + (int64_t)c * ld] : 0.0;
        }
    // Applies the published reflector (v, tau) of step j to this wave's columns c, j < c < k.
    auto apply = [&](int j, int par) {
        if (w * CW + CW - 1 <= j || w * CW >= k) return;       // no column of this wave is right of j
        const double t = tbuf[par];
        double v[RR];
#pragma unroll
        for (int r = 0; r < RR; ++r) v[r] = vbuf[par][lane + 64 * r];
        double dot[CW];
#pragma unroll
        for (int jj = 0; jj < CW; ++jj) {
            double s = 0.0;
#pragma unroll
            for (int r = 0; r < RR; ++r) s += v[r] * a[jj][r];
            dot[jj] = wave_sum_uniform(s);
        }
#pragma unroll
        for (int jj = 0; jj < CW; ++jj) {
            const int c = w * CW + jj;
            if (c > j && c < k) {
                const double f = t * dot[jj];
#pragma unroll
                for (int r = 0; r < RR; ++r) a[jj][r] -= f * v[r];
            }
        }
    };
    int step = 0;
#pragma unroll 1
    for (int j = 0; j < kk; ++j, ++step) {
        const int par = step & 1;
        if (w == j / CW) {
            const int js = j - w * CW;
            double x[RR];
#pragma unroll
            for (int r = 0; r < RR; ++r) {
                x[r] = 0.0;
#pragma unroll
                for (int jj = 0; jj < CW; ++jj) x[r] = jj == js ? a[jj][r] : x[r];
            }
            double s = 0.0, xj = 0.0;
#pragma unroll
            for (int r = 0; r < RR; ++r) {
                const int i = lane + 64 * r;
                const double xb = (i > j && i < m) ? x[r] : 0.0;
                s += xb * xb;
                xj = r == (j >> 6) ? x[r] : xj;
            }
            const double xn2 = wave_sum_uniform(s);
            const double alpha = bcast(xj, j & 63);
            double t = 0.0, scale = 0.0, beta = alpha;
            if (xn2 > 0.0 && alpha * alpha + xn2 > kTinyColumn) {
                beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
                t = (beta - alpha) / beta;
                scale = 1.0 / (alpha - beta);
            }
#pragma unroll
            for (int r = 0; r < RR; ++r) {
                const int i = lane + 64 * r;
                const bool below = i > j && i < m;
                vbuf[par][i] = below ? x[r] * scale : (i == j ? 1.0 : 0.0);
                const double nx = i == j ? beta : (i > j ? x[r] * scale : x[r]);
#pragma unroll
                for (int jj = 0; jj < CW; ++jj) a[jj][r] = jj == js ? nx : a[jj][r];
            }
            if (lane == 0) { tbuf[par] = t; taus[j] = t; }
        }
        __syncthreads();
        apply(j, par);
    }
    if (tk.roff >= 0) {
        T* R = reinterpret_cast<T*>(b.val[SLOT_WS]) + tk.roff;
#pragma unroll
        for (int jj = 0; jj < CW; ++jj)
#pragma unroll
            for (int r = 0; r < RR; ++r) {
                const int c = w * CW + jj, i = lane + 64 * r;
                if (c < k && i < kstat) R[i + (int64_t)c * tk.rld] = (i <= c && i < kk) ? (T)a[jj][r] : T(0);
            }
    }
#pragma unroll
    for (int jj = 0; jj < CW; ++jj)
        if (w * CW + jj >= kk) {
#pragma unroll
            for (int r = 0; r < RR; ++r) a[jj][r] = 0.0;
        }
#pragma unroll 1
    for (int j = kk - 1; j >= 0; --j, ++step) {
        const int par = step & 1;
        if (w == j / CW) {
            const int js = j - w * CW;
            const double t = taus[j];
#pragma unroll
            for (int r = 0; r < RR; ++r) {
                const int i = lane + 64 * r;
                double x = 0.0;
#pragma unroll
                for (int jj = 0; jj < CW; ++jj) x = jj == js ? a[jj][r] : x;
                const double v = i == j ? 1.0 : ((i > j && i < m) ? x : 0.0);
                vbuf[par][i] = v;
                const double nx = i == j ? 1.0 - t : (i > j ? -t * v : 0.0);
#pragma unroll
                for (int jj = 0; jj < CW; ++jj) a[jj][r] = jj == js ? nx : a[jj][r];
            }
            if (lane == 0) tbuf[par] = t;
        }
        __syncthreads();
        apply(j, par);
    }
#pragma unroll
    for (int jj = 0; jj < CW; ++jj)
#pragma unroll
        for (int r = 0; r < RR; ++r) {
            const int c = w * CW + jj, i = lane + 64 * r;
            if (c < k && i < m) A[i + (int64_t)c * ld] = (T)a[jj][r];
        }
}

// Panel classes of a QR phase (tasks are sorted by class at upload):
// 0: m <= 64, 1: m <= 256, 2: m <= 512 (k <= 40), 3: m <= 1024 with k <= 24,
// 4: the rest (qr_kernel). Classes 0-3 take the column-split kernel with
// 64, 256, 512 and 1024 rows in registers. Measured (FP64, k = 20, one launch):
// 80 panels of 64 rows 21 us (64 us with lane = column and rows over 8 waves,
// 162 us with one wave per panel), 2,000 of them 93 us with 8 waves (103 us
// with 12, 193 us with one wave), 80 panels of 256 rows 28 us (166 us with
// 4 waves of lane = row and the column sums through LDS), 100 of 1024 rows
// 72 us (374 us).
constexpr int kQrManyPanels = 512;

template <typename T, int KMAX>
static void launch_qr_classes(const QrTask* tasks, const int* classEnd, const Bases& b, hipStream_t s)
{
    constexpr int NWV = KMAX == 24 ? 12 : 10;      // 2 or 4 columns per wave
    int begin = 0;
    for (int c = 0; c < 4; ++c) {
        const int n = classEnd[c] - begin;
        if (n > 0) {
            const QrTask* t = tasks + begin;
            if (c == 0) {
                if (KMAX == 24 && n >= kQrManyPanels) qr_split_kernel<T, 24, 8, 1><<<n, 512, 0, s>>>(b, t);
                else qr_split_kernel<T, KMAX, NWV, 1><<<n, 64 * NWV, 0, s>>>(b, t);
            } else if (c == 1) {
                qr_split_kernel<T, KMAX, NWV, 4><<<n, 64 * NWV, 0, s>>>(b, t);
            } else if (c == 2) {
                qr_split_kernel<T, KMAX, NWV, 8><<<n, 64 * NWV, 0, s>>>(b, t);
            } else {
                qr_split_kernel<T, 24, 12, 16><<<n, 768, 0, s>>>(b, t);    // class 3 has k <= 24
            }
        }
        begin = classEnd[c];
    }
}

template <typename T>
void launch_qr_phase(const Phase& ph, const Bases& b, hipStream_t s)
{
    const int n = (int)ph.qr.size();
    if (n == 0) return;
    const QrTask* tasks = static_cast<const QrTask*>(ph.dTasks);
    if (ph.qrClassEnd[3] > 0) {
        if (ph.qrRegK <= 24) launch_qr_classes<T, 24>(tasks, ph.qrClassEnd, b, s);
        else launch_qr_classes<T, 40>(tasks, ph.qrClassEnd, b, s);
    }
    const int big = ph.qrClassEnd[3];
    if (n > big) qr_kernel<T><<<n - big, 256, 0, s>>>(b, tasks + big);
    HIP_LAUNCH_CHECK();
}
template void launch_qr_phase<double>(const Phase&, const Bases&, hipStream_t);
template void launch_qr_phase<float>(const Phase&, const Bases&, hipStream_t);

} // namespace dmhm

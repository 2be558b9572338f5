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

// Register-resident Householder QR for panels of m <= 64*NW rows and
// k <= KMAX columns: one workgroup of NW waves per panel, thread t holding
// row t of the panel in registers. The reflector loop runs at run time; the
// row's registers stay statically indexed because column j is read and
// written through unrolled compare-selects over the KMAX columns. Each
// reflector step needs one reduction round: the squared norm of the column
// below the diagonal and its products with the later columns are reduced
// together (wave sums, then the NW wave partials through LDS), and the pivot
// row is published through LDS in the same round, so a step costs one
// barrier (the LDS buffers alternate between steps). Q is formed in place
// (orgqr order). FP64 arithmetic; same reflector convention and outputs as
// qr_kernel; the panel and R may be strided (leading dimensions ld, rld), as
// the row blocks of a TSQR are.
template <int KMAX>
__device__ __forceinline__ double pick_col(const double (&a)[KMAX], int j)
{
    double v = 0.0;
#pragma unroll
    for (int c = 0; c < KMAX; ++c) v = (c == j) ? a[c] : v;
    return v;
}

template <typename T, int KMAX, int NW, bool BATCHED>
__global__ void __launch_bounds__(64 * NW) qr_reg_kernel(Bases b, const QrTask* __restrict__ tasks)
{
    __shared__ double part[2][NW][KMAX];
    __shared__ double rowj[2][KMAX];
    __shared__ double taus[KMAX];
    const QrTask tk = tasks[blockIdx.x];
    T* A = reinterpret_cast<T*>(b.val[tk.slot]) + tk.off;
    const int m = tk.m, kstat = tk.k, ld = tk.ld;
    // Effective width (width-indirect sketches): columns >= k are left alone.
    const int k = tk.krank >= 0 ? min(kstat, b.rank[tk.kslot][tk.krank]) : kstat, kk = min(m, k);
    if (k <= 0) return;
    const int i = threadIdx.x, w = i >> 6, lane = i & 63;
    const bool inPanel = i < m;
    // BATCHED: all column sums of a step reduced together (lowest latency,
    // more registers: occupancy 2), for phases of few panels; else one
    // column at a time (more panels in flight), for large phases.
    constexpr bool kBatchedSums = BATCHED;
    double a[KMAX];
#pragma unroll
    for (int c = 0; c < KMAX; ++c) a[c] = (inPanel && c < k) ? (double)A[i + (int64_t)c * ld] : 0.0;

#pragma unroll 1
    for (int j = 0; j < kk; ++j) {
        const int buf = j & 1;
        const bool below = i > j && inPanel;
        const double aj = pick_col(a, j);
        if (i == j) {
#pragma unroll
            for (int c = 0; c < KMAX; ++c)
                if (c >= j) rowj[buf][c] = a[c];
        }
        // Column products (zero outside columns j..k-1): reduced together
        // where the registers allow it (latency paid once per butterfly
        // step), else one column at a time.
        if constexpr (kBatchedSums) {
            double sv[KMAX];
#pragma unroll
            for (int c = 0; c < KMAX; ++c) sv[c] = (below && c >= j && c < k) ? aj * a[c] : 0.0;
            wave_sum_all<KMAX>(sv);
            if (lane == 0) {
#pragma unroll
                for (int c = 0; c < KMAX; ++c) part[buf][w][c] = sv[c];
            }
        } else {
#pragma unroll
            for (int c = 0; c < KMAX; ++c) {
                if (c >= j && c < k) {
                    const double sum = wave_sum(below ? aj * a[c] : 0.0);
                    if (lane == 0) part[buf][w][c] = sum;
                }
            }
        }
        __syncthreads();
        double xn2 = 0.0;
#pragma unroll
        for (int q = 0; q < NW; ++q) xn2 += part[buf][q][j];
        const double alpha = rowj[buf][j];
        double t = 0.0, scale = 0.0, beta = alpha;
        if (xn2 > 0.0 && alpha * alpha + xn2 > kTinyColumn) {
            beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
            t = (beta - alpha) / beta;
            scale = 1.0 / (alpha - beta);
        }
        const double vi = aj * scale;          // reflector entry of this row (rows below j)
#pragma unroll
        for (int c = 0; c < KMAX; ++c) {
            if (c > j && c < k) {
                double d = 0.0;
#pragma unroll
                for (int q = 0; q < NW; ++q) d += part[buf][q][c];
                const double wc = t * (rowj[buf][c] + scale * d);
                if (i == j) a[c] -= wc;
                else if (below) a[c] -= wc * vi;
            }
        }
        const double newj = (i == j) ? beta : (below ? vi : aj);
#pragma unroll
        for (int c = 0; c < KMAX; ++c)
            if (c == j) a[c] = newj;
        if (i == 0) taus[j] = t;
    }
    __syncthreads();
    // R: rows 0..kstat-1 of the first k columns (rows >= kk zero, so stacked R
    // factors of a TSQR carry zero rows below a narrower effective width).
    if (tk.roff >= 0 && i < kstat) {
        T* R = reinterpret_cast<T*>(b.val[SLOT_WS]) + tk.roff;
#pragma unroll
        for (int c = 0; c < KMAX; ++c)
            if (c < k) R[i + (int64_t)c * tk.rld] = (i <= c && i < kk) ? (T)a[c] : T(0);
    }
    // Q in place: columns >= kk start at zero, then H_j for j = kk-1 .. 0.
#pragma unroll
    for (int c = 0; c < KMAX; ++c)
        if (c >= kk) a[c] = 0.0;
#pragma unroll 1
    for (int j = kk - 1; j >= 0; --j) {
        const int buf = j & 1;
        const double t = taus[j];
        const double aj = pick_col(a, j);
        const double v = (i == j) ? 1.0 : ((i > j && inPanel) ? aj : 0.0);
        if constexpr (kBatchedSums) {
            double sv[KMAX];
#pragma unroll
            for (int c = 0; c < KMAX; ++c) sv[c] = (c > j && c < k) ? v * a[c] : 0.0;
            wave_sum_all<KMAX>(sv);
            if (lane == 0) {
#pragma unroll
                for (int c = 0; c < KMAX; ++c) part[buf][w][c] = sv[c];
            }
        } else {
#pragma unroll
            for (int c = 0; c < KMAX; ++c) {
                if (c > j && c < k) {
                    const double sum = wave_sum(v * a[c]);
                    if (lane == 0) part[buf][w][c] = sum;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < KMAX; ++c) {
            if (c > j && c < k) {
                double d = 0.0;
#pragma unroll
                for (int q = 0; q < NW; ++q) d += part[buf][q][c];
                a[c] -= t * d * v;
            }
        }
        const double newj = (i == j) ? 1.0 - t : ((i > j && inPanel) ? -t * aj : 0.0);
#pragma unroll
        for (int c = 0; c < KMAX; ++c)
            if (c == j) a[c] = newj;
    }
    if (inPanel) {
#pragma unroll
        for (int c = 0; c < KMAX; ++c)
            if (c < k) A[i + (int64_t)c * ld] = (T)a[c];
    }
}

// Householder QR of a panel of m <= NW*RW rows and k <= 64 columns by NW
// waves with the panel in registers, transposed to the other kernels: lane c
// holds column c, wave w rows RW*w .. RW*w+RW-1. Step j: each lane reads the
// reflector column j from lane j of its wave (readlane, uniform) and forms
// its column's product with it over the wave's rows; the NW wave partials
// (column j's own product is the squared norm below the diagonal) are summed
// through LDS, where the wave holding row j also publishes that row. One
// barrier per step (the LDS buffers alternate) and no cross-lane reduction.
// Q is formed in place in orgqr order the same way. Same reflector
// convention, R and Q as qr_kernel.
template <typename T, int RW, int NW>
__global__ void __launch_bounds__(64 * NW) qr_col_kernel(Bases b, const QrTask* __restrict__ tasks)
{
    __shared__ double part[2][NW][64];
    __shared__ double rowj[2][64];
    __shared__ double taus[64];
    const QrTask tk = tasks[blockIdx.x];
    T* A = reinterpret_cast<T*>(b.val[tk.slot]) + tk.off;
    const int m = tk.m, kstat = tk.k, ld = tk.ld;
    const int k = tk.krank >= 0 ? min(kstat, b.rank[tk.kslot][tk.krank]) : kstat, kk = min(m, k);
    if (k <= 0) return;
    const int c = threadIdx.x & 63, w = threadIdx.x >> 6, r0 = w * RW;
    double a[RW];
#pragma unroll
    for (int r = 0; r < RW; ++r) a[r] = (c < k && r0 + r < m) ? (double)A[r0 + r + (int64_t)c * ld] : 0.0;
#pragma unroll 1
    for (int j = 0; j < kk; ++j) {
        const int buf = j & 1;
        double dot = 0.0;
#pragma unroll
        for (int r = 0; r < RW; ++r)
            if (r0 + r > j) dot += bcast(a[r], j) * a[r];
        part[buf][w][c] = dot;
        if (w == j / RW) {
            double v = 0.0;
#pragma unroll
            for (int r = 0; r < RW; ++r) v = (r0 + r == j) ? a[r] : v;
            rowj[buf][c] = v;
        }
        __syncthreads();
        double xn2 = 0.0, dc = 0.0;
#pragma unroll
        for (int q = 0; q < NW; ++q) { xn2 += part[buf][q][j]; dc += part[buf][q][c]; }
        const double alpha = rowj[buf][j];
        double t = 0.0, scale = 0.0, beta = alpha;
        if (xn2 > 0.0 && alpha * alpha + xn2 > kTinyColumn) {
            beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
            t = (beta - alpha) / beta;
            scale = 1.0 / (alpha - beta);
        }
        const double wc = t * (rowj[buf][c] + scale * dc);
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            const int i = r0 + r;
            const double vi = bcast(a[r], j) * scale;     // reflector entry of row i (rows below j)
            if (c == j) a[r] = (i == j) ? beta : ((i > j) ? vi : a[r]);
            else if (c > j && i == j) a[r] -= wc;
            else if (c > j && i > j) a[r] -= wc * vi;
        }
        if (threadIdx.x == 0) taus[j] = t;
    }
    __syncthreads();
    if (tk.roff >= 0 && c < k) {
        T* R = reinterpret_cast<T*>(b.val[SLOT_WS]) + tk.roff;
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            const int i = r0 + r;
            if (i < kstat) R[i + (int64_t)c * tk.rld] = (i <= c && i < kk) ? (T)a[r] : T(0);
        }
    }
    if (c >= kk) {
#pragma unroll
        for (int r = 0; r < RW; ++r) a[r] = 0.0;
    }
#pragma unroll 1
    for (int j = kk - 1; j >= 0; --j) {
        const int buf = j & 1;
        const double t = taus[j];
        double dot = 0.0;
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            const int i = r0 + r;
            const double v = (i == j) ? 1.0 : ((i > j) ? bcast(a[r], j) : 0.0);
            dot += v * a[r];
        }
        part[buf][w][c] = dot;
        __syncthreads();
        double dc = 0.0;
#pragma unroll
        for (int q = 0; q < NW; ++q) dc += part[buf][q][c];
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            const int i = r0 + r;
            const double v = (i == j) ? 1.0 : ((i > j) ? bcast(a[r], j) : 0.0);
            if (c == j) a[r] = (i == j) ? 1.0 - t : ((i > j) ? -t * v : 0.0);
            else if (c > j && c < k) a[r] -= t * dc * v;
        }
    }
    if (c < k) {
#pragma unroll
        for (int r = 0; r < RW; ++r)
            if (r0 + r < m) A[r0 + r + (int64_t)c * ld] = (T)a[r];
    }
}

// Panel classes of a QR phase (tasks are sorted by class at upload):
// 0: m <= 64, 1: m <= 256, 2: m <= 512 (k <= 40), 3: m <= 1024 with k <= 24
// (16 waves of at most 128 registers fill a CU), 4: the rest (qr_kernel).
// Latency form (batched column sums) for classes of fewer panels than this:
// they fit the chip with room to spare, so the time is one panel's latency.
constexpr int kQrLatencyPanels = 512;

template <typename T, int KMAX, int NW>
static void launch_qr_reg(int n, const QrTask* t, const Bases& b, hipStream_t s)
{
    if constexpr (KMAX == 24 && NW <= 8) {
        if (n < kQrLatencyPanels) {
            qr_reg_kernel<T, KMAX, NW, true><<<n, 64 * NW, 0, s>>>(b, t);
            return;
        }
    }
    qr_reg_kernel<T, KMAX, NW, false><<<n, 64 * NW, 0, s>>>(b, t);
}

template <typename T, int KMAX>
static void launch_qr_classes(const QrTask* tasks, const int* classEnd, const Bases& b, hipStream_t s)
{
    int begin = 0;
    for (int c = 0; c < 4; ++c) {
        const int n = classEnd[c] - begin;
        if (n > 0) {
            const QrTask* t = tasks + begin;
            // Measured per class (FP64, k = 20 and 36, 80 and 2,000 panels):
            // few panels (latency) of <= 64 rows and few wide panels of <= 256
            // rows take the column-per-lane kernel; otherwise the row-per-lane
            // kernel (with batched column sums for few 24-column panels),
            // except for many wide panels of <= 64 rows.
            const bool few = n < kQrLatencyPanels;
            if (c == 0) {
                if (few || KMAX > 24) qr_col_kernel<T, 8, 8><<<n, 512, 0, s>>>(b, t);
                else qr_reg_kernel<T, KMAX, 1, false><<<n, 64, 0, s>>>(b, t);
            } else if (c == 1) {
                if (few && KMAX > 24) qr_col_kernel<T, 16, 16><<<n, 1024, 0, s>>>(b, t);
                else launch_qr_reg<T, KMAX, 4>(n, t, b, s);
            }
            else if (c == 2) launch_qr_reg<T, KMAX, 8>(n, t, b, s);
            else launch_qr_reg<T, 24, 16>(n, t, b, s);    // class 3 has k <= 24
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

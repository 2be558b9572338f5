// gfx950 batched one-sided Jacobi SVD of small square factors.
#include "kcommon.hpp"

namespace dmhm {

// ---------------------------------------------------------------------------
// Batched one-sided Jacobi SVD of k x k (k <= 64), FP64 in LDS.
// R = U S V^T; outputs G1 = V S (first `keep` columns, rest zero) and
// G2 = U (same), both k x cap, and the kept rank. The rotations act on the
// columns of R^T (the rows of the triangular R): Jacobi on the transposed
// triangular factor of a QR converges in fewer sweeps (7-8 against 10-15 on
// sketch factors of decaying spectra, 20 columns). With R^T W = A (columns
// of A orthogonal, norms s), R = W S (A/s)^T: G1 = A and G2 = W, and no
// column is divided by its norm.
// ---------------------------------------------------------------------------
// KMAX bounds k (LDS sized for it); NT threads, TPP (a power of two <= 8)
// per column pair.
// Rotation (c, s) of a column pair with squared norms al, be and inner
// product ga that makes the pair orthogonal: t = tan(theta) is the root of
// t^2 + 2 zeta t - 1 = 0 of smaller magnitude, zeta = (be - al) / (2 ga),
// written as t = sgn(d) g / (|d| + sqrt(d^2 + g^2)) with d = be - al and
// g = 2 ga. t does not change when d and g are scaled together, so both are
// scaled by an estimate of 1 / max(|d|, |g|) first (no overflow or underflow
// in the squares). c = 1 / sqrt(1 + t^2), s = c t. The reciprocals and
// reciprocal square roots are the hardware estimates (v_rcp_f64, v_rsq_f64)
// refined by two Newton steps each: a shorter dependent chain than IEEE
// division and square root, and c^2 + s^2 = 1 to rounding either way.
__device__ inline double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}
__device__ inline double rsq_nr(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
    return fma(0.5 * y, fma(-x * y, y, 1.0), y);
}
__device__ inline void jacobi_rotation(double al, double be, double ga, double& cs, double& sn)
{
    const double d0 = be - al, g0 = 2.0 * ga;
    const double rm = __builtin_amdgcn_rcp(fmax(fabs(d0), fabs(g0)));
    const double dd = d0 * rm, gg = g0 * rm;
    const double h = fma(dd, dd, gg * gg);
    const double den = fabs(dd) + h * rsq_nr(h);
    const double t = (dd >= 0.0 ? gg : -gg) * rcp_nr(den);
    cs = rsq_nr(fma(t, t, 1.0));
    sn = cs * t;
}

template <typename T, int KMAX, int NT, int TPP = 8>
__global__ void __launch_bounds__(NT) svd_kernel(Bases b, const SvdTask* __restrict__ tasks,
                                                 int tolMode, double relTol)
{
    static_assert(NT >= TPP * (KMAX / 2), "one TPP-thread group per column pair");
    __shared__ double As[KMAX][KMAX + 1];
    __shared__ double Vs[KMAX][KMAX + 1];
    __shared__ double sig[KMAX];
    __shared__ int order[KMAX];
    __shared__ int rotated;
    __shared__ int keep_s;
    __shared__ double fro2;
    const SvdTask tk = tasks[blockIdx.x];
    const int ld = tk.ld > 0 ? tk.ld : tk.k;
    const int k = tk.krank >= 0 ? min(tk.k, b.rank[SLOT_WIDTH][tk.krank]) : tk.k;
    const int tid = threadIdx.x;
    T* ws = reinterpret_cast<T*>(b.val[SLOT_WS]);
    for (int e = tid; e < k * k; e += NT) {
        const int r = e % k, c = e / k;
        As[c][r] = (double)ws[tk.roff + c + (int64_t)r * ld];    // column c of R^T
        Vs[c][r] = (r == c) ? 1.0 : 0.0;
    }
    // ||R||_F^2: a pair whose inner product is below 1e-30 ||R||_F^2 is taken
    // as orthogonal. Rank-deficient factors (sketches wider than the block's
    // rank) carry columns of rounding noise that the relative test alone
    // would keep rotating until the sweep limit.
    if (tid == 0) fro2 = 0.0;
    __syncthreads();
    {
        double s = 0;
        for (int e = tid; e < k * k; e += NT) {
            const double x = As[e / k][e % k];
            s += x * x;
        }
        s = wave_sum(s);
        if ((tid & 63) == 0) atomicAdd(&fro2, s);
    }
    const int kp = k + (k & 1);
    constexpr int LOG = TPP == 16 ? 4 : (TPP == 8 ? 3 : (TPP == 4 ? 2 : (TPP == 2 ? 1 : 0)));
    static_assert((1 << LOG) == TPP, "TPP must be 1, 2, 4, 8 or 16");
    constexpr int RPT = (KMAX + TPP - 1) / TPP;     // rows per thread
    const int pi = tid >> LOG, sub = tid & (TPP - 1);
    __syncthreads();
    const double froTol = 1e-30 * fro2;
    // Round-robin order: seat 0 keeps its column, the others advance by one
    // seat per round; pair pi takes seats pi and kp - 1 - pi. The column at a
    // seat after `round` rounds is 1 + (seat - 1 + round) mod (kp - 1), kept
    // here incrementally.
    auto next = [&](int c) { return c == 0 ? 0 : (c == kp - 1 ? 1 : c + 1); };
    for (int sweep = 0; sweep < 60; ++sweep) {
        __syncthreads();
        if (tid == 0) rotated = 0;
        __syncthreads();
        int a0 = pi, a1 = kp - 1 - pi;                  // round 0: column = seat
        for (int round = 0; round < kp - 1; ++round) {
            if (pi < kp / 2) {
                const int p = min(a0, a1), q = max(a0, a1);
                if (q < k) {
                    // All of the group's LDS reads first, so their latencies overlap.
                    double x[RPT], y[RPT], vx[RPT], vy[RPT];
#pragma unroll
                    for (int i = 0; i < RPT; ++i) {
                        const int r = sub + i * TPP;
                        const bool in = r < k;
                        x[i] = in ? As[p][r] : 0.0;
                        y[i] = in ? As[q][r] : 0.0;
                        vx[i] = in ? Vs[p][r] : 0.0;
                        vy[i] = in ? Vs[q][r] : 0.0;
                    }
                    double al = 0, be = 0, ga = 0;
#pragma unroll
                    for (int i = 0; i < RPT; ++i) {
                        al += x[i] * x[i]; be += y[i] * y[i]; ga += x[i] * y[i];
                    }
                    al = group_sum<TPP>(al);
                    be = group_sum<TPP>(be);
                    ga = group_sum<TPP>(ga);
                    if (ga * ga > 1e-30 * (al * be) && fabs(ga) > froTol) {
                        double cs, sn;
                        jacobi_rotation(al, be, ga, cs, sn);
#pragma unroll
                        for (int i = 0; i < RPT; ++i) {
                            const int r = sub + i * TPP;
                            if (r < k) {
                                As[p][r] = cs * x[i] - sn * y[i]; As[q][r] = sn * x[i] + cs * y[i];
                                Vs[p][r] = cs * vx[i] - sn * vy[i]; Vs[q][r] = sn * vx[i] + cs * vy[i];
                            }
                        }
                        if (sub == 0) rotated = 1;
                    }
                }
            }
            a0 = next(a0); a1 = next(a1);
            __syncthreads();
        }
        const int rot = rotated;
        if (!rot) break;
    }
    __syncthreads();
    if (tid < k) {
        double s = 0;
        for (int r = 0; r < k; ++r) s += As[tid][r] * As[tid][r];
        sig[tid] = sqrt(s);
    }
    __syncthreads();
    if (tid == 0) {
        for (int c = 0; c < k; ++c) order[c] = c;
        for (int a = 0; a < k; ++a) {          // selection sort, descending
            int best = a;
            for (int c = a + 1; c < k; ++c)
                if (sig[order[c]] > sig[order[best]]) best = c;
            const int tmp = order[a]; order[a] = order[best]; order[best] = tmp;
        }
        int keep = min(tk.maxKeep, k);
        if (tolMode == 1) {
            const double s0 = sig[order[0]];
            if (s0 < relTol * relTol) keep = min(1, keep);
            else {
                int r = 0;
                while (r < keep && sig[order[r]] > relTol * s0) ++r;
                keep = r;
            }
        }
        keep_s = keep;
        b.rank[SLOT_C][tk.rankIndex] = keep;
        // A-posteriori check of a narrowed sketch: fewer than `spare` of its
        // singular values fell below the truncation rule, so the block may
        // hold more numerical rank than the sketch saw.
        if (tk.krank >= 0 && k < tk.k && keep > k - tk.spare && b.rank[SLOT_FLAG])
            b.rank[SLOT_FLAG][tk.rankIndex] = 1;
    }
    __syncthreads();
    const int keep = keep_s;
    const int kst = tk.k;                 // rows k..kst-1 of G1, G2 are zero
    for (int e = tid; e < kst * tk.cap; e += NT) {
        const int r = e % kst, j = e / kst;
        double g1 = 0.0, g2 = 0.0;
        if (j < keep && r < k) {
            const int c = order[j];
            g1 = As[c][r];
            g2 = Vs[c][r];
        }
        ws[tk.g1off + r + (int64_t)j * ld] = (T)g1;
        ws[tk.g2off + r + (int64_t)j * ld] = (T)g2;
    }
}

constexpr int kSvdLatencyTasks = 256;

template <typename T>
void launch_svd_phase(const Phase& ph, const Bases& b, TruncPolicy pol, hipStream_t s)
{
    const int n = (int)ph.svd.size();
    if (n == 0) return;
    const SvdTask* t = static_cast<const SvdTask*>(ph.dTasks);
    if (ph.svdK > 64) throw std::runtime_error("SVD factor wider than 64");
    // 24-column factors: 16 threads per column pair for phases of few factors
    // (one workgroup per CU, the time is one factor's latency: 89 us against
    // 99 us with 8 threads for 80 factors of k = 20), 8 threads beyond
    // (120 us against 136 us for 850 factors).
    if (ph.svdK <= 24 && n < kSvdLatencyTasks) svd_kernel<T, 24, 192, 16><<<n, 192, 0, s>>>(b, t, pol.tolMode, pol.relTol);
    else if (ph.svdK <= 24) svd_kernel<T, 24, 128, 8><<<n, 128, 0, s>>>(b, t, pol.tolMode, pol.relTol);
    else if (ph.svdK <= 40) svd_kernel<T, 40, 192><<<n, 192, 0, s>>>(b, t, pol.tolMode, pol.relTol);
    else svd_kernel<T, 64, 256><<<n, 256, 0, s>>>(b, t, pol.tolMode, pol.relTol);
    HIP_LAUNCH_CHECK();
}

template void launch_svd_phase<double>(const Phase&, const Bases&, TruncPolicy, hipStream_t);
template void launch_svd_phase<float>(const Phase&, const Bases&, TruncPolicy, hipStream_t);

} // namespace dmhm

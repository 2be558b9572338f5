// gfx950 tile GEMM engine: grouped output-owned small GEMMs (H*X, sub-H
// products, projections, dense-leaf accumulation, factor formation).
#include "kcommon.hpp"

namespace dmhm {

// ---------------------------------------------------------------------------
// Tile GEMM engine.
// Workgroup = 256 threads; a tile is at most MR x NC (MR = 64 or 16 rows,
// NC = 16, 32 or 64 columns). Thread (ti, tj) = (tid % RT, tid / RT) owns
// the TM x TN micro-tile at rows TM*ti.., columns TN*tj.., accumulated in
// registers (TM + TN LDS reads per TM*TN FMAs). L and R are staged through
// LDS in chunks of QC along the inner dimension, double-buffered: the global
// loads of the next chunk (possibly of the next contribution) are issued
// before the current chunk is multiplied, so one barrier per chunk suffices
// and load latency overlaps the arithmetic. Out-of-range entries are staged
// as zeros, so every chunk multiplies a full QC slice. R is pre-scaled by
// the contribution's alpha when it is stored to LDS.
// ---------------------------------------------------------------------------
// Inner-dimension chunk. (Chunks twice as deep for 16-row tiles measured the
// same on the 256^2 rank-16 product: 117 ms FP64 either way.)
template <typename T, int MR> struct TileCfg { static constexpr int QC = 32; };
template <int MR> struct TileCfg<double, MR> { static constexpr int QC = 16; };

__device__ __forceinline__ int contrib_q(const Bases& b, const Contrib& C)
{
    int q = C.q;
    if (C.qrank >= 0) q = min(q, b.rank[C.qslot][C.qrank]);
    return q;
}

template <typename T, int MR, int NC, int QC, int LN, int RN>
__device__ __forceinline__ void tile_load(const Bases& b, const Contrib& C, int p0, int q, int m, int k,
                                          T alpha, T beta, T (&lr)[LN], T (&rr)[RN], T& a)
{
    const int tid = threadIdx.x;
    const int qc = min(QC, q - p0);
    const T* __restrict__ L = reinterpret_cast<const T*>(b.val[C.lslot]) + C.loff;
    const T* __restrict__ R = reinterpret_cast<const T*>(b.val[C.rslot]) + C.roff;
    a = (C.flags & CF_RUNTIME_ALPHA) ? alpha * T(C.alpha)
      : (C.flags & CF_RUNTIME_BETA) ? beta * T(C.alpha) : T(C.alpha);
    const bool tL = C.flags & CF_TRANS_L, tR = C.flags & CF_TRANS_R;
#pragma unroll
    for (int u = 0; u < LN; ++u) {
        const int e = tid + 256 * u;
        int ii, pp;
        if (!tL) { ii = e % MR; pp = e / MR; } else { pp = e % QC; ii = e / QC; }
        const bool ok = e < MR * QC && ii < m && pp < qc;
        lr[u] = ok ? (!tL ? L[ii + (int64_t)(p0 + pp) * C.ldl] : L[(p0 + pp) + (int64_t)ii * C.ldl]) : T(0);
    }
#pragma unroll
    for (int u = 0; u < RN; ++u) {
        const int e = tid + 256 * u;
        int jj, pp;
        if (!tR) { pp = e % QC; jj = e / QC; } else { jj = e % NC; pp = e / NC; }
        const bool ok = e < QC * NC && jj < k && pp < qc;
        rr[u] = ok ? (!tR ? R[(p0 + pp) + (int64_t)jj * C.ldr] : R[jj + (int64_t)(p0 + pp) * C.ldr]) : T(0);
    }
}

template <typename T, int MR, int NC, int QC, int LN, int RN, int LP, int RP>
__device__ __forceinline__ void tile_store(T* __restrict__ Lb, T* __restrict__ Rb, const T (&lr)[LN],
                                           const T (&rr)[RN], T a, uint8_t flags)
{
    const int tid = threadIdx.x;
    const bool tL = flags & CF_TRANS_L, tR = flags & CF_TRANS_R;
#pragma unroll
    for (int u = 0; u < LN; ++u) {
        const int e = tid + 256 * u;
        if (e < MR * QC) {
            int ii, pp;
            if (!tL) { ii = e % MR; pp = e / MR; } else { pp = e % QC; ii = e / QC; }
            Lb[pp * LP + ii] = lr[u];
        }
    }
#pragma unroll
    for (int u = 0; u < RN; ++u) {
        const int e = tid + 256 * u;
        if (e < QC * NC) {
            int jj, pp;
            if (!tR) { pp = e % QC; jj = e / QC; } else { jj = e % NC; pp = e / NC; }
            Rb[pp * RP + jj] = a * rr[u];
        }
    }
}

template <typename T, int MR, int NC, int TM, int TN, int ROLE>
__global__ void __launch_bounds__(256) tile_kernel(Bases b, const Tile* __restrict__ tiles,
                                                   const Contrib* __restrict__ cs, T alpha, T beta)
{
    constexpr int QC = TileCfg<T, MR>::QC;
    constexpr int RT = MR / TM;
    static_assert(RT * (NC / TN) == 256, "the thread grid must cover 256 threads");
    constexpr int LN = (MR * QC + 255) / 256;
    constexpr int RN = (QC * NC + 255) / 256;
    constexpr int LP = MR + 4, RP = NC + 4;     // padded LDS rows (16-byte multiples)
    __shared__ T Ls[2][QC * LP];
    __shared__ T Rs[2][QC * RP];
    const Tile tl = tiles[blockIdx.x];
    int m = tl.m, k = tl.k;
    if (tl.mrank >= 0) m = min(m, max(0, b.rank[tl.mslot][tl.mrank] - tl.moff));
    if (tl.krank >= 0) k = min(k, max(0, b.rank[tl.kslot][tl.krank] - tl.koff));
    if (m <= 0 || k <= 0) return;
    const int tid = threadIdx.x, ti = tid % RT, tj = tid / RT;
    T acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[i][j] = T(0);

    // Chunk iterator over (contribution c, offset p0 < q); uniform across the workgroup.
    int c = tl.cb, q = 0, p0 = 0;
    for (; c < tl.ce; ++c) {
        q = contrib_q(b, cs[c]);
        if (q > 0) break;
    }
    T lr[LN], rr[RN], a = T(0);
    uint8_t flags = 0;
    if (c < tl.ce) {
        const Contrib C = cs[c];
        tile_load<T, MR, NC, QC, LN, RN>(b, C, p0, q, m, k, alpha, beta, lr, rr, a);
        flags = C.flags;
    }
    int buf = 0;
    while (c < tl.ce) {
        tile_store<T, MR, NC, QC, LN, RN, LP, RP>(Ls[buf], Rs[buf], lr, rr, a, flags);
        __syncthreads();
        // Advance and prefetch the next chunk while this one is multiplied.
        p0 += QC;
        if (p0 >= q) {
            p0 = 0;
            for (++c; c < tl.ce; ++c) {
                q = contrib_q(b, cs[c]);
                if (q > 0) break;
            }
        }
        if (c < tl.ce) {
            const Contrib C = cs[c];
            tile_load<T, MR, NC, QC, LN, RN>(b, C, p0, q, m, k, alpha, beta, lr, rr, a);
            flags = C.flags;
        }
        const T* __restrict__ Lb = Ls[buf];
        const T* __restrict__ Rb = Rs[buf];
#pragma unroll 4
        for (int p = 0; p < QC; ++p) {
            T lv[TM], rv[TN];
#pragma unroll
            for (int i = 0; i < TM; ++i) lv[i] = Lb[p * LP + ti * TM + i];
#pragma unroll
            for (int j = 0; j < TN; ++j) rv[j] = Rb[p * RP + tj * TN + j];
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j) acc[i][j] += lv[i] * rv[j];
        }
        buf ^= 1;
    }
    T* out = reinterpret_cast<T*>(b.val[tl.slot]) + tl.off;
    const bool useBeta = (tl.flags & TF_RUNTIME_BETA) && beta != T(0);
    const bool accum = (tl.flags & TF_ACCUMULATE);
#pragma unroll
    for (int i = 0; i < TM; ++i) {
        const int r = ti * TM + i;
        if (r >= m) continue;
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int col = tj * TN + j;
            if (col < k) {
                T* o = out + r + (int64_t)col * tl.ld;
                if (useBeta) *o = beta * *o + acc[i][j];
                else if (accum) *o += acc[i][j];
                else *o = acc[i][j];
            }
        }
    }
}

template <typename T, int ROLE, int MR>
static void launch_tile_mr(int n, int kmax, const Bases& b, const Tile* t, const Contrib* c, T alpha, T beta,
                           hipStream_t s)
{
    constexpr int TM = MR == 64 ? 4 : 1;
    if (kmax <= 16) tile_kernel<T, MR, 16, TM, 1, ROLE><<<n, 256, 0, s>>>(b, t, c, alpha, beta);
    else if (kmax <= 32) tile_kernel<T, MR, 32, TM, 2, ROLE><<<n, 256, 0, s>>>(b, t, c, alpha, beta);
    else tile_kernel<T, MR, 64, TM, 4, ROLE><<<n, 256, 0, s>>>(b, t, c, alpha, beta);
}

template <typename T>
void launch_tile_phase(const Phase& ph, const Bases& b, T alpha, T beta, hipStream_t s)
{
    const int n = (int)ph.tiles.size();
    if (n == 0) return;
    const Tile* t = static_cast<const Tile*>(ph.dTiles);
    const Contrib* c = static_cast<const Contrib*>(ph.dContribs);
    const int kmax = ph.maxKTile;
    if (kmax > 64 || ph.maxMTile > 64) throw std::runtime_error("tile larger than 64 x 64");
    // ROLE only distinguishes projection and accumulation launches in profiles;
    // MR = 16 serves phases whose tiles all have at most 16 rows.
    if (ph.maxMTile <= 16) {
        if (ph.role == 1) launch_tile_mr<T, 1, 16>(n, kmax, b, t, c, alpha, beta, s);
        else launch_tile_mr<T, 0, 16>(n, kmax, b, t, c, alpha, beta, s);
    } else {
        if (ph.role == 1) launch_tile_mr<T, 1, 64>(n, kmax, b, t, c, alpha, beta, s);
        else launch_tile_mr<T, 0, 64>(n, kmax, b, t, c, alpha, beta, s);
    }
    HIP_LAUNCH_CHECK();
}

template void launch_tile_phase<double>(const Phase&, const Bases&, double, double, hipStream_t);
template void launch_tile_phase<float>(const Phase&, const Bases&, float, float, hipStream_t);

} // namespace dmhm

// gfx950 tile GEMM engine: grouped output-owned small GEMMs (H*X, sub-H
// products, projections, dense-leaf accumulation, factor formation).
#include <climits>
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

// Operand bases as the tile bodies read them, by run-time slot numbers that
// are uniform across the workgroup: a plain launch's own (kernel arguments),
// or a segmented launch's shifted by the member segment's offsets (kernel
// arguments plus the segment's table entry). Both resolve through uniform
// loads, with no private array indexed at run time and no LDS copy.
struct ArgBases {
    const Bases& b;
    template <typename T> __device__ __forceinline__ T* val(int s) const { return reinterpret_cast<T*>(b.val[s]); }
    __device__ __forceinline__ const int* rank(int s) const { return b.rank[s]; }
};
struct SegBases {
    const Bases& b;
    const TileSeg* __restrict__ sg;
    template <typename T> __device__ __forceinline__ T* val(int s) const
    {
        return reinterpret_cast<T*>(b.val[s]) + sg->val[s];
    }
    __device__ __forceinline__ const int* rank(int s) const { return b.rank[s] + sg->rank[s]; }
};

template <typename BV>
__device__ __forceinline__ int contrib_q(const BV& b, const Contrib& C)
{
    int q = C.q;
    if (C.qrank >= 0) q = min(q, b.rank(C.qslot)[C.qrank]);
    return q;
}

template <typename T, int MR, int NC, int QC, int LN, int RN, typename BV>
__device__ __forceinline__ void tile_load(const BV& b, const Contrib& C, int p0, int q, int m, int k,
                                          T alpha, T beta, T (&lr)[LN], T (&rr)[RN], T& a)
{
    const int tid = threadIdx.x;
    const int qc = min(QC, q - p0);
    const T* __restrict__ L = b.template val<const T>(C.lslot) + C.loff;
    const T* __restrict__ R = b.template val<const T>(C.rslot) + C.roff;
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

template <typename T, int MR, int NC, int TM, int TN, typename BV>
__device__ __forceinline__ void tile_body(const BV& b, const Tile& tl, const Contrib* __restrict__ cs, T alpha,
                                          T beta)
{
    constexpr int QC = TileCfg<T, MR>::QC;
    constexpr int RT = MR / TM;
    static_assert(RT * (NC / TN) == 256, "the thread grid must cover 256 threads");
    constexpr int LN = (MR * QC + 255) / 256;
    constexpr int RN = (QC * NC + 255) / 256;
    constexpr int LP = MR + 4, RP = NC + 4;     // padded LDS rows (16-byte multiples)
    __shared__ T Ls[2][QC * LP];
    __shared__ T Rs[2][QC * RP];
    int m = tl.m, k = tl.k;
    if (tl.mrank >= 0) m = min(m, max(0, b.rank(tl.mslot)[tl.mrank] - tl.moff));
    if (tl.krank >= 0) k = min(k, max(0, b.rank(tl.kslot)[tl.krank] - tl.koff));
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
    T* out = b.template val<T>(tl.slot) + tl.off;
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

// Segmented launches (merged plans of independent members): the member
// segment holding workgroup g, found by bisection over the segments' first
// workgroups (one or four tiles per workgroup), and the member's bases, the
// launch's bases shifted by the member's offsets.
template <int TPB>
__device__ __forceinline__ int seg_of(const TileSeg* __restrict__ segs, int nseg, int g)
{
    if (nseg <= 64) {
        // One load per lane and a ballot: the segment is the last one whose
        // first workgroup is at or before g.
        const int l = threadIdx.x & 63;
        const int f = l < nseg ? (TPB == 1 ? segs[l].first1 : segs[l].first4) : INT_MAX;
        return (int)__popcll(__ballot(f <= g)) - 1;
    }
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int f = TPB == 1 ? segs[mid].first1 : segs[mid].first4;
        if (f <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename T, int MR, int NC, int TM, int TN, int ROLE>
__global__ void __launch_bounds__(256) tile_kernel(Bases b, const Tile* __restrict__ tiles,
                                                   const Contrib* __restrict__ cs, T alpha, T beta)
{
    tile_body<T, MR, NC, TM, TN>(ArgBases{b}, tiles[blockIdx.x], cs, alpha, beta);
}

template <typename T, int MR, int NC, int TM, int TN>
__global__ void __launch_bounds__(256) tile_seg_kernel(Bases b, const TileSeg* __restrict__ segs, int nseg, T alpha,
                                                       T beta)
{
    const TileSeg* sg = segs + seg_of<1>(segs, nseg, (int)blockIdx.x);
    const int local = (int)blockIdx.x - sg->first1;
    if (local >= sg->count) return;
    tile_body<T, MR, NC, TM, TN>(SegBases{b, sg}, sg->tiles[local], sg->cs, alpha, beta);
}

// ---------------------------------------------------------------------------
// MFMA engine: one wave per 16-row block of an output tile. A phase of 64-row
// tiles runs one workgroup of four waves per tile (wave w owns rows
// 16w..16w+15), a phase of tiles of at most 16 rows packs four tiles per
// workgroup; no LDS and no barriers, so every wave streams its contributions
// independently. A wave computes its block transposed, D^T = R^T L^T, in
// 16 x 16 MFMA blocks: the first operand is a 16-column block of R^T, the
// second the wave's 16-row block of L^T, and lane l holds
// D(row0 + l%16, cb*16 + c(l, r)) in accumulator register r, so each group of
// 16 lanes stores 16 consecutive rows of a column. The FP32 and BF16
// instructions place output row 4*(l/16) + r of the 16 x 16 block in register
// r of lane l, the FP64 one row l/16 + 4*r (measured on gfx950), hence
// c(l, r) = 4*(l/16) + r for FP32 accumulators and l/16 + 4*r for FP64. The
// inner dimension advances in chunks of 16; lane l supplies inner indices
// p0 + 4*(l/16) + e, e < 4, to both operands: four K = 4 steps of
// v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32 (MODE 0, arithmetic in T), or
// one K = 16 step of v_mfma_f32_16x16x16_bf16 on operands rounded from FP32 to
// BF16 with FP32 accumulation (MODE 1). Out-of-range entries load as zero.
// PF = 1 issues the loads of the next chunk (possibly of the next
// contribution) before the MFMAs of the current one.
// ---------------------------------------------------------------------------
typedef double mf_d4 __attribute__((ext_vector_type(4)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));
typedef __bf16 mf_b4 __attribute__((ext_vector_type(4)));

template <typename T> struct MfAcc;
template <> struct MfAcc<double> { typedef mf_d4 type; };
template <> struct MfAcc<float> { typedef mf_f4 type; };

__device__ __forceinline__ mf_d4 mf_step(double a, double b, mf_d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ mf_f4 mf_step(float a, float b, mf_f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ mf_f4 mf_step_bf16(const float (&a)[4], const float (&b)[4], mf_f4 c)
{
    mf_b4 x, y;
#pragma unroll
    for (int e = 0; e < 4; ++e) { x[e] = (__bf16)a[e]; y[e] = (__bf16)b[e]; }
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(x, y, c, 0, 0, 0);
}

// Operands of one contribution as the wave reads them.
template <typename T>
struct MfSrc {
    const T* L;
    const T* R;
    int64_t ldl, ldr;
    int q;
    T a;
    bool tL, tR;
};

template <typename T, typename BV>
__device__ __forceinline__ MfSrc<T> mf_src(const BV& b, const Contrib& C, int q, T alpha, T beta)
{
    MfSrc<T> s;
    s.L = b.template val<const T>(C.lslot) + C.loff;
    s.R = b.template val<const T>(C.rslot) + C.roff;
    s.ldl = C.ldl; s.ldr = C.ldr; s.q = q;
    s.a = (C.flags & CF_RUNTIME_ALPHA) ? alpha * T(C.alpha)
        : (C.flags & CF_RUNTIME_BETA) ? beta * T(C.alpha) : T(C.alpha);
    s.tL = C.flags & CF_TRANS_L; s.tR = C.flags & CF_TRANS_R;
    return s;
}

// Fragments of chunk p0: lf[e] = L(row, p), rf[cb][e] = a R(p, cb*16 + l%16),
// p = p0 + 4*(l/16) + e.
template <typename T, int NBMAX>
__device__ __forceinline__ void mf_load(const MfSrc<T>& s, int p0, int row, int m, int k, int NB, int li, int lk,
                                        T (&lf)[4], T (&rf)[NBMAX][4])
{
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int p = p0 + 4 * lk + e;
        const bool pin = p < s.q;
        lf[e] = (pin && row < m) ? (s.tL ? s.L[p + row * s.ldl] : s.L[row + p * s.ldl]) : T(0);
#pragma unroll
        for (int cb = 0; cb < NBMAX; ++cb) {
            const int col = cb * 16 + li;
            rf[cb][e] = (cb < NB && pin && col < k) ? s.a * (s.tR ? s.R[col + p * s.ldr] : s.R[p + col * s.ldr])
                                                    : T(0);
        }
    }
}

template <typename T, int MODE, int NBMAX, typename Acc>
__device__ __forceinline__ void mf_mma(const T (&lf)[4], const T (&rf)[NBMAX][4], int NB, Acc (&acc)[NBMAX])
{
#pragma unroll
    for (int cb = 0; cb < NBMAX; ++cb) {
        if (cb >= NB) continue;
        if constexpr (MODE == 1) {
            acc[cb] = mf_step_bf16(rf[cb], lf, acc[cb]);
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[cb] = mf_step(rf[cb][e], lf[e], acc[cb]);
        }
    }
}

// First contribution at or after c with a non-empty inner dimension.
template <typename BV>
__device__ __forceinline__ int mf_next(const BV& b, const Contrib* __restrict__ cs, int c, int ce, int& q)
{
    for (; c < ce; ++c) {
        q = contrib_q(b, cs[c]);
        if (q > 0) break;
    }
    return c;
}

template <typename T, int MODE, int MBMAX, int NBMAX, int PF, typename BV>
__device__ __forceinline__ void mfma_body(const BV& b, const Tile& tl, const Contrib* __restrict__ cs, int rb,
                                          T alpha, T beta)
{
    static_assert(MODE == 0 || sizeof(T) == 4, "the BF16 mode reads FP32 operands");
    static_assert(MBMAX == 1 || MBMAX == 4, "tiles of at most 16 or 64 rows");
    typedef typename MfAcc<T>::type Acc;
    int m = tl.m, k = tl.k;
    if (tl.mrank >= 0) m = min(m, max(0, b.rank(tl.mslot)[tl.mrank] - tl.moff));
    if (tl.krank >= 0) k = min(k, max(0, b.rank(tl.kslot)[tl.krank] - tl.koff));
    if (m <= rb * 16 || k <= 0) return;
    const int NB = (k + 15) >> 4;
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const int row = rb * 16 + li;
    Acc acc[NBMAX];
#pragma unroll
    for (int cb = 0; cb < NBMAX; ++cb)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[cb][r] = T(0);

    int q = 0;
    int c = mf_next(b, cs, tl.cb, tl.ce, q);
    if (c < tl.ce) {
        MfSrc<T> src = mf_src<T>(b, cs[c], q, alpha, beta);
        int p0 = 0;
        T lf[4], rf[NBMAX][4];
        mf_load<T, NBMAX>(src, p0, row, m, k, NB, li, lk, lf, rf);
        while (true) {
            // The next chunk: p0 + 16 of this contribution, or the next one.
            int p1 = p0 + 16, c1 = c, q1 = src.q;
            if (p1 >= src.q) { p1 = 0; c1 = mf_next(b, cs, c + 1, tl.ce, q1); }
            const bool more = c1 < tl.ce;
            MfSrc<T> src1 = src;
            if (more && c1 != c) src1 = mf_src<T>(b, cs[c1], q1, alpha, beta);
            if constexpr (PF == 1) {
                T lf1[4], rf1[NBMAX][4];
                if (more) mf_load<T, NBMAX>(src1, p1, row, m, k, NB, li, lk, lf1, rf1);
                mf_mma<T, MODE, NBMAX>(lf, rf, NB, acc);
                if (!more) break;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    lf[e] = lf1[e];
#pragma unroll
                    for (int cb = 0; cb < NBMAX; ++cb) rf[cb][e] = rf1[cb][e];
                }
            } else {
                mf_mma<T, MODE, NBMAX>(lf, rf, NB, acc);
                if (!more) break;
                mf_load<T, NBMAX>(src1, p1, row, m, k, NB, li, lk, lf, rf);
            }
            c = c1; p0 = p1; src = src1;
        }
    }
    if (row >= m) return;
    T* out = b.template val<T>(tl.slot) + tl.off;
    const bool useBeta = (tl.flags & TF_RUNTIME_BETA) && beta != T(0);
    const bool accum = (tl.flags & TF_ACCUMULATE);
    const int64_t ld = tl.ld;
#pragma unroll
    for (int cb = 0; cb < NBMAX; ++cb) {
        if (cb >= NB) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int col = cb * 16 + (sizeof(T) == 8 ? lk + 4 * r : 4 * lk + r);
            if (col < k) {
                T* o = out + row + col * ld;
                const T v = acc[cb][r];
                if (useBeta) *o = beta * *o + v;
                else if (accum) *o += v;
                else *o = v;
            }
        }
    }
}

template <typename T, int MODE, int MBMAX, int NBMAX, int PF>
__global__ void __launch_bounds__(256) tile_mfma_kernel(Bases b, const Tile* __restrict__ tiles,
                                                        const Contrib* __restrict__ cs, int ntiles, T alpha,
                                                        T beta)
{
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int id = MBMAX == 4 ? (int)blockIdx.x : (int)blockIdx.x * 4 + w;
    if (id >= ntiles) return;
    mfma_body<T, MODE, MBMAX, NBMAX, PF>(ArgBases{b}, tiles[id], cs, MBMAX == 4 ? w : 0, alpha, beta);
}

template <typename T, int MODE, int MBMAX, int NBMAX, int PF>
__global__ void __launch_bounds__(256) tile_mfma_seg_kernel(Bases b, const TileSeg* __restrict__ segs, int nseg,
                                                            T alpha, T beta)
{
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const TileSeg* sg = segs + seg_of<MBMAX == 4 ? 1 : 4>(segs, nseg, (int)blockIdx.x);
    const int local = MBMAX == 4 ? (int)blockIdx.x - sg->first1 : ((int)blockIdx.x - sg->first4) * 4 + w;
    if (local >= sg->count) return;
    mfma_body<T, MODE, MBMAX, NBMAX, PF>(SegBases{b, sg}, sg->tiles[local], sg->cs, MBMAX == 4 ? w : 0, alpha, beta);
}

// Segmented launches of a merged phase (its members' segment table).
template <typename T, int MODE, int MBMAX, int PF>
static void launch_mfma_seg_mb(const Phase& ph, const Bases& b, T alpha, T beta, hipStream_t s)
{
    const int grid = MBMAX == 4 ? ph.segBlocks1 : ph.segBlocks4;
    const int nseg = (int)ph.segSrc.size();
    const TileSeg* sg = static_cast<const TileSeg*>(ph.dSegs);
    const int kmax = ph.maxKTile;
    if (grid == 0) return;
    if (kmax <= 16) tile_mfma_seg_kernel<T, MODE, MBMAX, 1, PF><<<grid, 256, 0, s>>>(b, sg, nseg, alpha, beta);
    else if (kmax <= 32) tile_mfma_seg_kernel<T, MODE, MBMAX, 2, PF><<<grid, 256, 0, s>>>(b, sg, nseg, alpha, beta);
    else tile_mfma_seg_kernel<T, MODE, MBMAX, 4, PF><<<grid, 256, 0, s>>>(b, sg, nseg, alpha, beta);
}

template <typename T, int MR>
static void launch_tile_seg_mr(const Phase& ph, const Bases& b, T alpha, T beta, hipStream_t s)
{
    constexpr int TM = MR == 64 ? 4 : 1;
    const int grid = ph.segBlocks1;
    const int nseg = (int)ph.segSrc.size();
    const TileSeg* sg = static_cast<const TileSeg*>(ph.dSegs);
    const int kmax = ph.maxKTile;
    if (grid == 0) return;
    if (kmax <= 16) tile_seg_kernel<T, MR, 16, TM, 1><<<grid, 256, 0, s>>>(b, sg, nseg, alpha, beta);
    else if (kmax <= 32) tile_seg_kernel<T, MR, 32, TM, 2><<<grid, 256, 0, s>>>(b, sg, nseg, alpha, beta);
    else tile_seg_kernel<T, MR, 64, TM, 4><<<grid, 256, 0, s>>>(b, sg, nseg, alpha, beta);
}

template <typename T, int MODE, int PF>
static void launch_mfma_seg(const Phase& ph, const Bases& b, T alpha, T beta, hipStream_t s)
{
    if (ph.maxMTile <= 16) launch_mfma_seg_mb<T, MODE, 1, PF>(ph, b, alpha, beta, s);
    else launch_mfma_seg_mb<T, MODE, 4, PF>(ph, b, alpha, beta, s);
}

template <typename T, int MODE, int MBMAX, int PF>
static void launch_mfma_mb(int n, int kmax, const Bases& b, const Tile* t, const Contrib* c, T alpha, T beta,
                           hipStream_t s)
{
    const int grid = MBMAX == 4 ? n : (n + 3) / 4;
    if (kmax <= 16) tile_mfma_kernel<T, MODE, MBMAX, 1, PF><<<grid, 256, 0, s>>>(b, t, c, n, alpha, beta);
    else if (kmax <= 32) tile_mfma_kernel<T, MODE, MBMAX, 2, PF><<<grid, 256, 0, s>>>(b, t, c, n, alpha, beta);
    else tile_mfma_kernel<T, MODE, MBMAX, 4, PF><<<grid, 256, 0, s>>>(b, t, c, n, alpha, beta);
}

template <typename T, int MODE, int PF>
static void launch_mfma(const Phase& ph, const Bases& b, T alpha, T beta, hipStream_t s)
{
    const int n = (int)ph.tiles.size();
    const Tile* t = static_cast<const Tile*>(ph.dTiles);
    const Contrib* c = static_cast<const Contrib*>(ph.dContribs);
    if (ph.maxMTile <= 16) launch_mfma_mb<T, MODE, 1, PF>(n, ph.maxKTile, b, t, c, alpha, beta, s);
    else launch_mfma_mb<T, MODE, 4, PF>(n, ph.maxKTile, b, t, c, alpha, beta, s);
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

// Member segments of at least this many tiles launch one by one.
constexpr size_t kSegSplitTiles = 8192;

template <typename T>
void launch_tile_phase(const Phase& ph, const Bases& b, T alpha, T beta, hipStream_t s, int engine)
{
    if (!ph.segSrc.empty()) {
        if (ph.maxKTile > 64 || ph.maxMTile > 64) throw std::runtime_error("tile larger than 64 x 64");
        if (!ph.dSegs) throw std::runtime_error("segmented tile phase not uploaded");
        // Members whose own tile lists fill the chip several times over run
        // as one plain launch each, on their own bases: one launch over all
        // of them saves nothing and interleaves the members' operands in the
        // caches (measured on 12 3D members of 16^3-20^3: 30% slower).
        size_t smallest = SIZE_MAX;
        for (const auto& sg : ph.segSrc) smallest = std::min(smallest, sg.src->tiles.size());
        if (smallest >= kSegSplitTiles) {
            for (const auto& sg : ph.segSrc) {
                Bases mb;
                for (int k = 0; k < NUM_SLOTS; ++k) {
                    mb.val[k] = b.val[k] ? static_cast<void*>(static_cast<T*>(b.val[k]) + sg.val[k]) : nullptr;
                    mb.rank[k] = b.rank[k] ? b.rank[k] + sg.rank[k] : nullptr;
                }
                launch_tile_phase<T>(*sg.src, mb, alpha, beta, s, engine);
            }
            return;
        }
        if (engine != TILE_LDS) {
            const bool bf16 = (engine == TILE_MFMA_BF16 || engine == TILE_MFMA_BF16_PF) && sizeof(T) == 4;
            const bool pf = (engine == TILE_MFMA_PF || engine == TILE_MFMA_BF16_PF) && sizeof(T) == 4;
            if (bf16) {
                if constexpr (sizeof(T) == 4) {
                    if (pf) launch_mfma_seg<float, 1, 1>(ph, b, alpha, beta, s);
                    else launch_mfma_seg<float, 1, 0>(ph, b, alpha, beta, s);
                }
            } else {
                if (pf) launch_mfma_seg<T, 0, 1>(ph, b, alpha, beta, s);
                else launch_mfma_seg<T, 0, 0>(ph, b, alpha, beta, s);
            }
        } else if (ph.maxMTile <= 16) {
            launch_tile_seg_mr<T, 16>(ph, b, alpha, beta, s);
        } else {
            launch_tile_seg_mr<T, 64>(ph, b, alpha, beta, s);
        }
        HIP_LAUNCH_CHECK();
        return;
    }
    const int n = (int)ph.tiles.size();
    if (n == 0) return;
    const Tile* t = static_cast<const Tile*>(ph.dTiles);
    const Contrib* c = static_cast<const Contrib*>(ph.dContribs);
    const int kmax = ph.maxKTile;
    if (kmax > 64 || ph.maxMTile > 64) throw std::runtime_error("tile larger than 64 x 64");
    if (engine != TILE_LDS) {
        const bool bf16 = (engine == TILE_MFMA_BF16 || engine == TILE_MFMA_BF16_PF) && sizeof(T) == 4;
        // The prefetching form runs in FP32 only: in FP64 it produced wrong
        // products on gfx950 (the non-prefetching form is exact).
        const bool pf = (engine == TILE_MFMA_PF || engine == TILE_MFMA_BF16_PF) && sizeof(T) == 4;
        if (bf16) {
            if constexpr (sizeof(T) == 4) {
                if (pf) launch_mfma<float, 1, 1>(ph, b, alpha, beta, s);
                else launch_mfma<float, 1, 0>(ph, b, alpha, beta, s);
            }
        } else {
            if (pf) launch_mfma<T, 0, 1>(ph, b, alpha, beta, s);
            else launch_mfma<T, 0, 0>(ph, b, alpha, beta, s);
        }
        HIP_LAUNCH_CHECK();
        return;
    }
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

template void launch_tile_phase<double>(const Phase&, const Bases&, double, double, hipStream_t, int);
template void launch_tile_phase<float>(const Phase&, const Bases&, float, float, hipStream_t, int);

} // namespace dmhm

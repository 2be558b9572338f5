// gfx950 kernels for the H-matrix path. Wave64 throughout.
//
//  tile_kernel   grouped output-owned small GEMMs (the H*X, sub-H products,
//                projections, dense-leaf accumulation and factor formation)
//  qr_kernel     batched Householder QR of tall-skinny panels, Q formed in place
//                (geqrf + orgqr semantics, one workgroup per panel)
//  svd_kernel    batched one-sided Jacobi SVD of k x k factors (k <= 64) held
//                in LDS, FP64, with the rank decision
//  rng_kernel    Gaussian sketch panels (counter-based, order independent)
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include "kernels.hpp"

namespace dmhm {

#define HIP_LAUNCH_CHECK()                                                        \
    do {                                                                          \
        hipError_t e_ = hipGetLastError();                                        \
        if (e_ != hipSuccess)                                                     \
            throw std::runtime_error(std::string("HIP launch: ") + hipGetErrorString(e_)); \
    } while (0)

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

// ---------------------------------------------------------------------------
// Wave helpers
// ---------------------------------------------------------------------------
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ inline float wave_sum(float v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
// Broadcast lane `l` (wave-uniform) of v to the whole wave.
__device__ inline float bcast(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ inline double bcast(double v, int l)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// ---------------------------------------------------------------------------
// Streaming H*X (K right-hand sides, HBM-bound).
// mv_proj: one wave per (admissible block, <= 512-row chunk of its source
//   box): P(j, kk) (+)= sum_r V(r, j) X(r, kk). V is read with 16-byte vector
//   loads: a lane owns VW = 16 / sizeof(T) consecutive rows, LPC = 64 / VW
//   lanes cover 64 rows of one column and the CG = VW lane groups of a wave
//   walk different columns (1 KiB per load instruction); one butterfly over
//   the LPC lanes of a column group finishes each sum (atomic add when several
//   chunks share P). Panels whose offset or leading dimension is not a
//   multiple of VW (ragged grids) use the same mapping with scalar loads.
// mv_acc: one workgroup per output row atom (<= 64 rows); its 4 waves split
//   the atom's contributions (dense leaves D X, low-rank U P). Lane i owns row
//   i and issues 16 column loads per batch (each instruction moves one 64-row
//   column segment); the R column is held one value per lane and broadcast
//   with readlane; one LDS reduction across the waves at the end. (A variant
//   with 16-byte row-group loads measured slower here: 143 us against 93 us
//   on the 256^2 rank-16 operator.)
// ---------------------------------------------------------------------------
template <typename T> struct VecOf;
template <> struct VecOf<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct VecOf<double> { typedef double type __attribute__((ext_vector_type(2))); };

template <typename T, int K>
__global__ void __launch_bounds__(256) mv_proj_kernel(Bases b, const MvProjItem* __restrict__ items, int n,
                                                      int ldx)
{
    constexpr int VW = 16 / (int)sizeof(T);
    constexpr int LPC = 64 / VW;                // lanes per column
    constexpr int CG = VW;                      // column groups per wave
    constexpr int UN = 16 / CG;                 // columns per lane per pass (16 columns per pass)
    typedef typename VecOf<T>::type vec;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int id = blockIdx.x * 4 + w;
    if (id >= n) return;
    const MvProjItem it = items[id];
    const int rank = min(it.cap, b.rank[SLOT_A][it.rankIdx]);
    if (rank <= 0) return;
    const int rg = lane % LPC, cg = lane / LPC;
    const T* __restrict__ V = reinterpret_cast<const T*>(b.val[SLOT_A]) + it.voff;
    const T* __restrict__ X = reinterpret_cast<const T*>(b.val[SLOT_X]) + it.xrow;
    T* P = reinterpret_cast<T*>(b.val[SLOT_WS]) + it.poff;
    const bool vecV = (it.voff % VW == 0) && (it.ldv % VW == 0);
    const bool vecX = (it.xrow % VW == 0) && (ldx % VW == 0) &&
                      ((reinterpret_cast<uintptr_t>(b.val[SLOT_X]) & 15) == 0);
    for (int j0 = 0; j0 < rank; j0 += 16) {
        T acc[UN][K];
#pragma unroll
        for (int u = 0; u < UN; ++u)
#pragma unroll
            for (int kk = 0; kk < K; ++kk) acc[u][kk] = T(0);
        for (int r0 = rg * VW; r0 < it.rows; r0 += 64) {
            const bool full = r0 + VW <= it.rows;
            T xv[K][VW];
#pragma unroll
            for (int kk = 0; kk < K; ++kk) {
                if (full && vecX) {
                    const vec v = *reinterpret_cast<const vec*>(X + r0 + (int64_t)kk * ldx);
#pragma unroll
                    for (int e = 0; e < VW; ++e) xv[kk][e] = v[e];
                } else {
#pragma unroll
                    for (int e = 0; e < VW; ++e)
                        xv[kk][e] = (r0 + e < it.rows) ? X[r0 + e + (int64_t)kk * ldx] : T(0);
                }
            }
            vec lv[UN];
#pragma unroll
            for (int u = 0; u < UN; ++u) {
                const int j = j0 + cg + CG * u;
                if (j < rank) {
                    const T* col = V + r0 + (int64_t)j * it.ldv;
                    if (full && vecV) lv[u] = *reinterpret_cast<const vec*>(col);
                    else {
#pragma unroll
                        for (int e = 0; e < VW; ++e) lv[u][e] = (r0 + e < it.rows) ? col[e] : T(0);
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < VW; ++e) lv[u][e] = T(0);
                }
            }
#pragma unroll
            for (int u = 0; u < UN; ++u)
#pragma unroll
                for (int kk = 0; kk < K; ++kk)
#pragma unroll
                    for (int e = 0; e < VW; ++e) acc[u][kk] += lv[u][e] * xv[kk][e];
        }
        // Sum over the LPC lanes of each column group (lane bits below LPC).
#pragma unroll
        for (int u = 0; u < UN; ++u)
#pragma unroll
            for (int kk = 0; kk < K; ++kk) {
                T v = acc[u][kk];
#pragma unroll
                for (int off = LPC / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
                acc[u][kk] = v;
            }
        if (rg == 0) {
#pragma unroll
            for (int u = 0; u < UN; ++u) {
                const int j = j0 + cg + CG * u;
                if (j >= rank) continue;
#pragma unroll
                for (int kk = 0; kk < K; ++kk) {
                    T* dst = P + j + (int64_t)kk * it.cap;
                    if (it.atomic) atomicAdd(dst, acc[u][kk]);
                    else *dst = acc[u][kk];
                }
            }
        }
    }
}

template <typename T, int K>
__global__ void __launch_bounds__(256) mv_acc_kernel(Bases b, const MvAtom* __restrict__ atoms,
                                                          const MvContrib* __restrict__ cs, int ldx, int ldy,
                                                          T alpha, T beta)
{
    constexpr int UNR = 16;
    __shared__ T red[4][K][64];
    const MvAtom at = atoms[blockIdx.x];
    const int w = threadIdx.x >> 6, i = threadIdx.x & 63;
    const bool row = i < at.m;
    T acc[K];
#pragma unroll
    for (int kk = 0; kk < K; ++kk) acc[kk] = T(0);
    const T* __restrict__ H = reinterpret_cast<const T*>(b.val[SLOT_A]);
    const T* __restrict__ X = reinterpret_cast<const T*>(b.val[SLOT_X]);
    const T* __restrict__ WS = reinterpret_cast<const T*>(b.val[SLOT_WS]);
    for (int c = at.cb + w; c < at.ce; c += 4) {
        const MvContrib C = cs[c];
        int q = C.q;
        if (C.kind == 1) q = min(q, b.rank[SLOT_A][C.rankIdx]);
        const T* __restrict__ L = H + C.loff;
        const T* __restrict__ R = (C.kind == 0 ? X : WS) + C.roff;
        const int ldr = C.kind == 0 ? ldx : C.ldr;
        for (int p0 = 0; p0 < q; p0 += 64) {
            const int qc = min(64, q - p0);
            T rv[K];
#pragma unroll
            for (int kk = 0; kk < K; ++kk) rv[kk] = (i < qc) ? R[p0 + i + (int64_t)kk * ldr] : T(0);
            const T* __restrict__ Lp = L + i + (int64_t)p0 * C.ld;
            int p = 0;
            for (; p + UNR <= qc; p += UNR) {
                T l[UNR];
#pragma unroll
                for (int u = 0; u < UNR; ++u) l[u] = row ? Lp[(int64_t)(p + u) * C.ld] : T(0);
#pragma unroll
                for (int u = 0; u < UNR; ++u)
#pragma unroll
                    for (int kk = 0; kk < K; ++kk) acc[kk] += l[u] * bcast(rv[kk], p + u);
            }
            for (; p < qc; ++p) {
                const T l = row ? Lp[(int64_t)p * C.ld] : T(0);
#pragma unroll
                for (int kk = 0; kk < K; ++kk) acc[kk] += l * bcast(rv[kk], p);
            }
        }
    }
#pragma unroll
    for (int kk = 0; kk < K; ++kk) red[w][kk][i] = acc[kk];
    __syncthreads();
    if (w == 0 && row) {
#pragma unroll
        for (int kk = 0; kk < K; ++kk) {
            const T sum = red[0][kk][i] + red[1][kk][i] + red[2][kk][i] + red[3][kk][i];
            T* y = reinterpret_cast<T*>(b.val[SLOT_Y]) + at.yrow + i + (int64_t)kk * ldy;
            *y = (beta != T(0) ? beta * *y : T(0)) + alpha * sum;
        }
    }
}

template <typename T, int K>
static void launch_mv_k(const Phase& ph, const Bases& b, T alpha, T beta, hipStream_t s)
{
    if (ph.kind == PH_MV_PROJ) {
        const int n = (int)ph.mvProj.size();
        if (n) mv_proj_kernel<T, K><<<(n + 3) / 4, 256, 0, s>>>(b, static_cast<const MvProjItem*>(ph.dTasks), n,
                                                               ph.ldx);
    } else {
        const int n = (int)ph.mvAtoms.size();
        if (n) mv_acc_kernel<T, K><<<n, 256, 0, s>>>(b, static_cast<const MvAtom*>(ph.dTasks),
                                                     static_cast<const MvContrib*>(ph.dAux), ph.ldx, ph.ldy,
                                                     alpha, beta);
    }
}

template <typename T>
void launch_mv_phase(const Phase& ph, const Bases& b, T alpha, T beta, hipStream_t s)
{
    if (ph.kind == PH_MV_ZERO) {
        launch_fill<T>(reinterpret_cast<T*>(b.val[SLOT_WS]) + ph.zeroOff, ph.zeroCount, T(0), s);
        return;
    }
    switch (ph.mvK) {
    case 1: launch_mv_k<T, 1>(ph, b, alpha, beta, s); break;
    case 2: launch_mv_k<T, 2>(ph, b, alpha, beta, s); break;
    case 3: launch_mv_k<T, 3>(ph, b, alpha, beta, s); break;
    case 4: launch_mv_k<T, 4>(ph, b, alpha, beta, s); break;
    case 5: launch_mv_k<T, 5>(ph, b, alpha, beta, s); break;
    case 6: launch_mv_k<T, 6>(ph, b, alpha, beta, s); break;
    case 7: launch_mv_k<T, 7>(ph, b, alpha, beta, s); break;
    case 8: launch_mv_k<T, 8>(ph, b, alpha, beta, s); break;
    default: throw std::runtime_error("streaming H*X supports 1..8 columns");
    }
    HIP_LAUNCH_CHECK();
}

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
    const int m = tk.m, k = tk.k, kk = min(m, k);
    const int tid = threadIdx.x;
    if (m <= 0) return;
    // Factor.
    for (int j = 0; j < kk; ++j) {
        column_dots(A, m, j, j, k, dots, red);
        if (tid == 0) {
            const double alpha = (double)A[j + (int64_t)j * m];
            const double xn2 = dots[j];
            double tau = 0.0, scale = 0.0, beta = alpha;
            if (xn2 > 0.0) {
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
    // R (k x k, upper triangle; rows >= kk zero).
    if (tk.roff >= 0) {
        T* R = reinterpret_cast<T*>(b.val[SLOT_WS]) + tk.roff;
        for (int e = tid; e < k * k; e += 256) {
            const int r = e % k, c = e / k;
            R[e] = (r <= c && r < kk) ? A[r + (int64_t)c * m] : T(0);
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

// Register-resident Householder QR for panels of m <= 64*NW*RPL rows and
// k <= KMAX columns: one workgroup of NW waves per panel, thread t holding
// rows t + 64*NW*r (r < RPL) in registers (columns unrolled at compile
// time). Each reflector step needs one reduction round: the squared norm of
// the column below the diagonal and its products with the later columns are
// reduced together (wave sums, then the NW wave partials through LDS), and
// the pivot row is published through LDS in the same round, so a step costs
// one barrier. Q is formed in place (orgqr order). FP64 arithmetic; same
// reflector convention and outputs as qr_kernel; the panel and R may be
// strided (leading dimensions ld, rld), as the row blocks of a TSQR are.
template <typename T, int KMAX, int NW, int RPL>
__global__ void __launch_bounds__(64 * NW) qr_reg_kernel(Bases b, const QrTask* __restrict__ tasks)
{
    constexpr int NT = 64 * NW;
    __shared__ double part[2][NW][KMAX];
    __shared__ double rowj[2][KMAX];
    __shared__ double taus[KMAX];
    const QrTask tk = tasks[blockIdx.x];
    T* A = reinterpret_cast<T*>(b.val[tk.slot]) + tk.off;
    const int m = tk.m, k = tk.k, kk = min(m, k), ld = tk.ld;
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    double a[RPL][KMAX];
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
        const int i = tid + NT * r;
#pragma unroll
        for (int c = 0; c < KMAX; ++c) a[r][c] = (i < m && c < k) ? (double)A[i + (int64_t)c * ld] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
        if (j < kk) {
            const int buf = j & 1;
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                const int i = tid + NT * r;
                if (i == j) {
#pragma unroll
                    for (int c = j; c < KMAX; ++c) rowj[buf][c] = a[r][c];
                }
            }
            // Column products reduced one column at a time (no partial array:
            // keeps the wave within 128 registers at KMAX = 40).
#pragma unroll
            for (int c = j; c < KMAX; ++c) {
                if (c < k) {
                    double pc = 0.0;
#pragma unroll
                    for (int r = 0; r < RPL; ++r) {
                        const int i = tid + NT * r;
                        if (i > j && i < m) pc += a[r][j] * a[r][c];
                    }
                    const double sum = wave_sum(pc);
                    if (lane == 0) part[buf][w][c] = sum;
                }
            }
            __syncthreads();
            double xn2 = 0.0;
#pragma unroll
            for (int q = 0; q < NW; ++q) xn2 += part[buf][q][j];
            const double alpha = rowj[buf][j];
            double t = 0.0, scale = 0.0, beta = alpha;
            if (xn2 > 0.0) {
                beta = -copysign(sqrt(alpha * alpha + xn2), alpha);
                t = (beta - alpha) / beta;
                scale = 1.0 / (alpha - beta);
            }
#pragma unroll
            for (int c = j + 1; c < KMAX; ++c) {
                if (c < k) {
                    double d = 0.0;
#pragma unroll
                    for (int q = 0; q < NW; ++q) d += part[buf][q][c];
                    const double wc = t * (rowj[buf][c] + scale * d);
#pragma unroll
                    for (int r = 0; r < RPL; ++r) {
                        const int i = tid + NT * r;
                        if (i == j) a[r][c] -= wc;
                        else if (i > j && i < m) a[r][c] -= wc * (a[r][j] * scale);
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                const int i = tid + NT * r;
                if (i == j) a[r][j] = beta;
                else if (i > j && i < m) a[r][j] *= scale;
            }
            if (tid == 0) taus[j] = t;
        }
    }
    __syncthreads();
    if (tk.roff >= 0) {
        T* R = reinterpret_cast<T*>(b.val[SLOT_WS]) + tk.roff;
#pragma unroll
        for (int r = 0; r < RPL; ++r) {
            const int i = tid + NT * r;
            if (i < k) {
#pragma unroll
                for (int c = 0; c < KMAX; ++c)
                    if (c < k) R[i + (int64_t)c * tk.rld] = (i <= c && i < kk) ? (T)a[r][c] : T(0);
            }
        }
    }
    // Q in place: columns >= kk start at zero, then H_j for j = kk-1 .. 0.
#pragma unroll
    for (int c = 0; c < KMAX; ++c)
        if (c >= kk) {
#pragma unroll
            for (int r = 0; r < RPL; ++r) a[r][c] = 0.0;
        }
#pragma unroll
    for (int j = KMAX - 1; j >= 0; --j) {
        if (j < kk) {
            const int buf = j & 1;
            const double t = taus[j];
#pragma unroll
            for (int c = j + 1; c < KMAX; ++c) {
                if (c < k) {
                    double pc = 0.0;
#pragma unroll
                    for (int r = 0; r < RPL; ++r) {
                        const int i = tid + NT * r;
                        const double v = (i == j) ? 1.0 : ((i > j && i < m) ? a[r][j] : 0.0);
                        pc += v * a[r][c];
                    }
                    const double sum = wave_sum(pc);
                    if (lane == 0) part[buf][w][c] = sum;
                }
            }
            __syncthreads();
#pragma unroll
            for (int c = j + 1; c < KMAX; ++c) {
                if (c < k) {
                    double d = 0.0;
#pragma unroll
                    for (int q = 0; q < NW; ++q) d += part[buf][q][c];
#pragma unroll
                    for (int r = 0; r < RPL; ++r) {
                        const int i = tid + NT * r;
                        const double v = (i == j) ? 1.0 : ((i > j && i < m) ? a[r][j] : 0.0);
                        a[r][c] -= t * d * v;
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < RPL; ++r) {
                const int i = tid + NT * r;
                a[r][j] = (i == j) ? 1.0 - t : ((i > j && i < m) ? -t * a[r][j] : 0.0);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
        const int i = tid + NT * r;
        if (i < m) {
#pragma unroll
            for (int c = 0; c < KMAX; ++c)
                if (c < k) A[i + (int64_t)c * ld] = (T)a[r][c];
        }
    }
}

// Panel classes of a QR phase (tasks are sorted by class at upload):
// 0: m <= 64, 1: m <= 256, 2: m <= 512 (k <= 40), 3: m <= 1024 with k <= 24
// (16 waves of at most 128 registers fill a CU), 4: the rest (qr_kernel).
template <typename T, int KMAX>
static void launch_qr_classes(const QrTask* tasks, const int* classEnd, const Bases& b, hipStream_t s)
{
    int begin = 0;
    for (int c = 0; c < 4; ++c) {
        const int n = classEnd[c] - begin;
        if (n > 0) {
            const QrTask* t = tasks + begin;
            if (c == 0) qr_reg_kernel<T, KMAX, 1, 1><<<n, 64, 0, s>>>(b, t);
            else if (c == 1) qr_reg_kernel<T, KMAX, 4, 1><<<n, 256, 0, s>>>(b, t);
            else if (c == 2) qr_reg_kernel<T, KMAX, 8, 1><<<n, 512, 0, s>>>(b, t);
            else qr_reg_kernel<T, 24, 16, 1><<<n, 1024, 0, s>>>(b, t);    // class 3 has k <= 24
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

// ---------------------------------------------------------------------------
// Batched one-sided Jacobi SVD of k x k (k <= 64), FP64 in LDS.
// R = U S V^T; outputs G1 = V S (first `keep` columns, rest zero) and
// G2 = U (same), both k x cap, and the kept rank.
// ---------------------------------------------------------------------------
// KMAX bounds k (LDS sized for it); NT threads, 8 per column pair.
template <typename T, int KMAX, int NT>
__global__ void __launch_bounds__(NT) svd_kernel(Bases b, const SvdTask* __restrict__ tasks,
                                                 int tolMode, double relTol)
{
    static_assert(NT >= 8 * (KMAX / 2), "one 8-thread group per column pair");
    __shared__ double As[KMAX][KMAX + 1];
    __shared__ double Vs[KMAX][KMAX + 1];
    __shared__ double sig[KMAX];
    __shared__ int order[KMAX];
    __shared__ int rotated;
    __shared__ int keep_s;
    const SvdTask tk = tasks[blockIdx.x];
    const int k = tk.k, tid = threadIdx.x;
    T* ws = reinterpret_cast<T*>(b.val[SLOT_WS]);
    for (int e = tid; e < k * k; e += NT) {
        const int r = e % k, c = e / k;
        As[c][r] = (double)ws[tk.roff + e];
        Vs[c][r] = (r == c) ? 1.0 : 0.0;
    }
    const int kp = k + (k & 1);
    const int pi = tid >> 3, sub = tid & 7;
    for (int sweep = 0; sweep < 60; ++sweep) {
        __syncthreads();
        if (tid == 0) rotated = 0;
        __syncthreads();
        for (int round = 0; round < kp - 1; ++round) {
            if (pi < kp / 2) {
                auto pos = [&](int x) { return x == 0 ? 0 : 1 + ((x - 1 + round) % (kp - 1)); };
                const int a0 = pos(pi), a1 = pos(kp - 1 - pi);
                const int p = min(a0, a1), q = max(a0, a1);
                double al = 0, be = 0, ga = 0;
                if (q < k)
                    for (int r = sub; r < k; r += 8) {
                        const double x = As[p][r], y = As[q][r];
                        al += x * x; be += y * y; ga += x * y;
                    }
#pragma unroll
                for (int off = 4; off >= 1; off >>= 1) {
                    al += __shfl_xor(al, off, 64);
                    be += __shfl_xor(be, off, 64);
                    ga += __shfl_xor(ga, off, 64);
                }
                if (q < k && ga != 0.0 && fabs(ga) > 1e-15 * sqrt(al * be)) {
                    const double zeta = (be - al) / (2.0 * ga);
                    const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                    const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
                    for (int r = sub; r < k; r += 8) {
                        const double x = As[p][r], y = As[q][r];
                        As[p][r] = cs * x - sn * y; As[q][r] = sn * x + cs * y;
                        const double vx = Vs[p][r], vy = Vs[q][r];
                        Vs[p][r] = cs * vx - sn * vy; Vs[q][r] = sn * vx + cs * vy;
                    }
                    if (sub == 0) rotated = 1;
                }
            }
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
        int keep = tk.maxKeep;
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
    }
    __syncthreads();
    const int keep = keep_s;
    for (int e = tid; e < k * tk.cap; e += NT) {
        const int r = e % k, j = e / k;
        double g1 = 0.0, g2 = 0.0;
        if (j < keep) {
            const int c = order[j];
            g1 = Vs[c][r] * sig[c];
            g2 = sig[c] > 0 ? As[c][r] / sig[c] : 0.0;
        }
        ws[tk.g1off + e] = (T)g1;
        ws[tk.g2off + e] = (T)g2;
    }
}

template <typename T>
void launch_svd_phase(const Phase& ph, const Bases& b, TruncPolicy pol, hipStream_t s)
{
    const int n = (int)ph.svd.size();
    if (n == 0) return;
    const SvdTask* t = static_cast<const SvdTask*>(ph.dTasks);
    if (ph.svdK > 64) throw std::runtime_error("SVD factor wider than 64");
    if (ph.svdK <= 24) svd_kernel<T, 24, 128><<<n, 128, 0, s>>>(b, t, pol.tolMode, pol.relTol);
    else if (ph.svdK <= 40) svd_kernel<T, 40, 192><<<n, 192, 0, s>>>(b, t, pol.tolMode, pol.relTol);
    else svd_kernel<T, 64, 256><<<n, 256, 0, s>>>(b, t, pol.tolMode, pol.relTol);
    HIP_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Counter-based Gaussian generator (splitmix64 + Box-Muller).
// ---------------------------------------------------------------------------
__device__ inline uint64_t mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

__device__ inline double gauss_at(uint64_t seed, uint64_t stream, uint64_t idx)
{
    const uint64_t h = mix64(seed ^ mix64(stream * 0xD1B54A32D192ED03ULL + mix64(idx)));
    const double u1 = ((double)(h >> 32) + 1.0) / 4294967296.0;
    const double u2 = (double)(h & 0xFFFFFFFFULL) / 4294967296.0;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

template <typename T>
__global__ void __launch_bounds__(256) rng_kernel(Bases b, const RngTask* __restrict__ tasks,
                                                  uint64_t seed)
{
    const RngTask tk = tasks[blockIdx.x];
    T* x = reinterpret_cast<T*>(b.val[SLOT_WS]) + tk.off;
    const int64_t n = (int64_t)tk.m * tk.k;
    for (int64_t e = threadIdx.x; e < n; e += 256) x[e] = (T)gauss_at(seed, tk.stream, (uint64_t)e);
}

template <typename T>
void launch_rng_phase(const Phase& ph, const Bases& b, uint64_t seed, hipStream_t s)
{
    const int n = (int)ph.rng.size();
    if (n == 0) return;
    rng_kernel<T><<<n, 256, 0, s>>>(b, static_cast<const RngTask*>(ph.dTasks), seed);
    HIP_LAUNCH_CHECK();
}

template <typename T>
__global__ void gaussian_kernel(T* x, int64_t n, uint64_t seed, uint64_t stream)
{
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        x[e] = (T)gauss_at(seed, stream, (uint64_t)e);
}

static int grid_for(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 8192); }

template <typename T> void launch_gaussian(T* x, int64_t n, uint64_t seed, uint64_t stream, hipStream_t s)
{
    if (n <= 0) return;
    gaussian_kernel<T><<<grid_for(n), 256, 0, s>>>(x, n, seed, stream);
    HIP_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Elementwise and per-leaf helpers.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void scale_kernel(T* x, int64_t n, T a)
{
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        x[e] = a * x[e];
}
template <typename T> void launch_scale(T* x, int64_t n, T alpha, hipStream_t s)
{
    if (n <= 0) return;
    scale_kernel<T><<<grid_for(n), 256, 0, s>>>(x, n, alpha);
    HIP_LAUNCH_CHECK();
}

template <typename T>
__global__ void axpby_kernel(const T* a, const T* b, T alpha, T* out, int64_t n)
{
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        out[e] = a[e] + alpha * b[e];
}
template <typename T> void launch_axpby(const T* a, const T* b, T alpha, T* out, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    axpby_kernel<T><<<grid_for(n), 256, 0, s>>>(a, b, alpha, out, n);
    HIP_LAUNCH_CHECK();
}

template <typename T>
__global__ void fill_kernel(T* x, int64_t n, T v)
{
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        x[e] = v;
}
template <typename T>
__global__ void eye_kernel(T* x, int n)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n * n) x[e] = (e % n == e / n) ? T(1) : T(0);
}

template <typename T> void launch_eye(T* x, int n, hipStream_t s)
{
    if (n <= 0) return;
    eye_kernel<T><<<(n * n + 255) / 256, 256, 0, s>>>(x, n);
    HIP_LAUNCH_CHECK();
}

template <typename T> void launch_fill(T* x, int64_t n, T v, hipStream_t s)
{
    if (n <= 0) return;
    fill_kernel<T><<<grid_for(n), 256, 0, s>>>(x, n, v);
    HIP_LAUNCH_CHECK();
}
void launch_fill_int(int* x, int64_t n, int v, hipStream_t s)
{
    if (n <= 0) return;
    fill_kernel<int><<<grid_for(n), 256, 0, s>>>(x, n, v);
    HIP_LAUNCH_CHECK();
}

template <typename Ts, typename Td>
__global__ void convert_kernel(const Ts* x, Td* y, int64_t n)
{
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        y[e] = (Td)x[e];
}
template <typename Ts, typename Td> void launch_convert(const Ts* x, Td* y, int64_t n, hipStream_t s)
{
    if (n <= 0) return;
    convert_kernel<Ts, Td><<<grid_for(n), 256, 0, s>>>(x, y, n);
    HIP_LAUNCH_CHECK();
}

template <typename T>
__global__ void add_diag_kernel(T* vals, const int64_t* offs, const int* sizes, int count, T a)
{
    const int blk = blockIdx.x;
    if (blk >= count) return;
    T* D = vals + offs[blk];
    const int n = sizes[blk];
    for (int i = threadIdx.x; i < n; i += blockDim.x) D[i + (int64_t)i * n] += a;
}
template <typename T> void launch_add_diag(T* vals, const int64_t* offs, const int* sizes, int count,
                                           T alpha, hipStream_t s)
{
    if (count <= 0) return;
    add_diag_kernel<T><<<count, 64, 0, s>>>(vals, offs, sizes, count, alpha);
    HIP_LAUNCH_CHECK();
}

template <typename T>
__global__ void __launch_bounds__(256) transpose_dense_kernel(const T* src, T* dst, const int64_t* srcOff,
                                                              const int64_t* dstOff, const int* mm,
                                                              const int* nn)
{
    __shared__ T tile[64][65];
    const int blk = blockIdx.x;
    const T* S = src + srcOff[blk];
    T* D = dst + dstOff[blk];
    const int m = mm[blk], n = nn[blk];          // source is m x n, destination n x m
    for (int r0 = 0; r0 < m; r0 += 64)
        for (int c0 = 0; c0 < n; c0 += 64) {
            for (int e = threadIdx.x; e < 64 * 64; e += 256) {
                const int r = e & 63, c = e >> 6;
                if (r0 + r < m && c0 + c < n) tile[c][r] = S[(r0 + r) + (int64_t)(c0 + c) * m];
            }
            __syncthreads();
            for (int e = threadIdx.x; e < 64 * 64; e += 256) {
                const int c = e & 63, r = e >> 6;        // destination row c, column r
                if (r0 + r < m && c0 + c < n) D[(c0 + c) + (int64_t)(r0 + r) * n] = tile[c][r];
            }
            __syncthreads();
        }
}
template <typename T> void launch_transpose_dense(const T* src, T* dst, const int64_t* srcOff,
                                                  const int64_t* dstOff, const int* m, const int* n,
                                                  int count, hipStream_t s)
{
    if (count <= 0) return;
    transpose_dense_kernel<T><<<count, 256, 0, s>>>(src, dst, srcOff, dstOff, m, n);
    HIP_LAUNCH_CHECK();
}

template <typename T>
__global__ void swap_lowrank_kernel(const T* src, T* dst, const int64_t* srcU, const int64_t* srcV,
                                    const int64_t* dstU, const int64_t* dstV, const int* mrows,
                                    const int* nrows, const int* cap)
{
    const int blk = blockIdx.x;
    // destination U (mrows x cap) <- source V; destination V (nrows x cap) <- source U
    const int64_t nu = (int64_t)mrows[blk] * cap[blk], nv = (int64_t)nrows[blk] * cap[blk];
    for (int64_t e = threadIdx.x; e < nu; e += blockDim.x) dst[dstU[blk] + e] = src[srcV[blk] + e];
    for (int64_t e = threadIdx.x; e < nv; e += blockDim.x) dst[dstV[blk] + e] = src[srcU[blk] + e];
}
template <typename T> void launch_swap_lowrank(const T* src, T* dst, const int64_t* srcU,
                                               const int64_t* srcV, const int64_t* dstU,
                                               const int64_t* dstV, const int* mrows,
                                               const int* nrows, const int* cap, int count,
                                               hipStream_t s)
{
    if (count <= 0) return;
    swap_lowrank_kernel<T><<<count, 256, 0, s>>>(src, dst, srcU, srcV, dstU, dstV, mrows, nrows, cap);
    HIP_LAUNCH_CHECK();
}

__global__ void permute_int_kernel(const int* src, int* dst, const int* perm, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = src[perm[i]];
}
void launch_permute_int(const int* src, int* dst, const int* perm, int count, hipStream_t s)
{
    if (count <= 0) return;
    permute_int_kernel<<<(count + 255) / 256, 256, 0, s>>>(src, dst, perm, count);
    HIP_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Power-iteration helpers.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) sumsq_kernel(const T* x, int64_t n, double* out)
{
    __shared__ double red[4];
    double s = 0;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const double v = (double)x[e];
        s += v * v;
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(out, red[0] + red[1] + red[2] + red[3]);
}
template <typename T> void launch_sumsq(const T* x, int64_t n, double* out, hipStream_t s)
{
    sumsq_kernel<T><<<std::min(grid_for(n), 1024), 256, 0, s>>>(x, n, out);
    HIP_LAUNCH_CHECK();
}

template <typename T>
__global__ void normalize_kernel(T* x, int64_t n, const double* sumsq, double* logAcc, double root)
{
    const double ss = *sumsq;
    const double inv = ss > 0 ? 1.0 / sqrt(ss) : 0.0;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        x[e] = (T)((double)x[e] * inv);
    if (logAcc && blockIdx.x == 0 && threadIdx.x == 0) *logAcc += root * 0.5 * log(ss);
}
template <typename T> void launch_normalize(T* x, int64_t n, const double* sumsq, double* logAcc,
                                            double root, hipStream_t s)
{
    normalize_kernel<T><<<grid_for(n), 256, 0, s>>>(x, n, sumsq, logAcc, root);
    HIP_LAUNCH_CHECK();
}

__global__ void zero_double_kernel(double* x, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}
void launch_zero_double(double* x, int n, hipStream_t s)
{
    zero_double_kernel<<<(n + 63) / 64, 64, 0, s>>>(x, n);
    HIP_LAUNCH_CHECK();
}

// Explicit instantiations.
#define INST(T)                                                                                       \
    template void launch_tile_phase<T>(const Phase&, const Bases&, T, T, hipStream_t);              \
    template void launch_qr_phase<T>(const Phase&, const Bases&, hipStream_t);                      \
    template void launch_svd_phase<T>(const Phase&, const Bases&, TruncPolicy, hipStream_t);        \
    template void launch_rng_phase<T>(const Phase&, const Bases&, uint64_t, hipStream_t);           \
    template void launch_mv_phase<T>(const Phase&, const Bases&, T, T, hipStream_t);                \
    template void launch_scale<T>(T*, int64_t, T, hipStream_t);                                     \
    template void launch_fill<T>(T*, int64_t, T, hipStream_t);                                      \
    template void launch_eye<T>(T*, int, hipStream_t);                                              \
    template void launch_axpby<T>(const T*, const T*, T, T*, int64_t, hipStream_t);                 \
    template void launch_gaussian<T>(T*, int64_t, uint64_t, uint64_t, hipStream_t);                 \
    template void launch_add_diag<T>(T*, const int64_t*, const int*, int, T, hipStream_t);          \
    template void launch_transpose_dense<T>(const T*, T*, const int64_t*, const int64_t*, const int*, \
                                            const int*, int, hipStream_t);                          \
    template void launch_swap_lowrank<T>(const T*, T*, const int64_t*, const int64_t*, const int64_t*, \
                                         const int64_t*, const int*, const int*, const int*, int,   \
                                         hipStream_t);                                              \
    template void launch_sumsq<T>(const T*, int64_t, double*, hipStream_t);                         \
    template void launch_normalize<T>(T*, int64_t, const double*, double*, double, hipStream_t);
INST(double)
INST(float)
template void launch_convert<double, float>(const double*, float*, int64_t, hipStream_t);
template void launch_convert<float, double>(const float*, double*, int64_t, hipStream_t);

} // namespace dmhm

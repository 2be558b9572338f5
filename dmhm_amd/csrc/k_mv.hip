// gfx950 streaming H*X kernels for few right-hand sides (HBM-bound).
#include "kcommon.hpp"

namespace dmhm {

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

template void launch_mv_phase<double>(const Phase&, const Bases&, double, double, hipStream_t);
template void launch_mv_phase<float>(const Phase&, const Bases&, float, float, hipStream_t);

} // namespace dmhm

// gfx950 kernels of the recursive direct inversion (HMat2d/HMat3d::
// DirectInvert, src/HMat2d/Invert-incl.hpp:17-125) and of merged plans: the
// dense leaf inverse (hmat_tools::Invert on a Dense, src/hmat_tools/
// Invert.cpp:14-42, there LAPACK getrf + getri), copies and fills of index
// ranges inside one allocation (sub-matrix copies: a subtree of the block
// tree is one contiguous range of each storage pool), and identities at
// several workspace offsets.
#include "kcommon.hpp"

namespace dmhm {

// dst[d + e] = src[s + e] (copy) or dst[d + e] = 0 (fill, src unused) for
// every range (s, d, len); blockIdx.y selects the range.
template <typename T, bool FILL>
__global__ void ranges_kernel(const T* __restrict__ src, T* __restrict__ dst, const CopyRange* __restrict__ rs)
{
    const CopyRange r = rs[blockIdx.y];
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < r.len; e += (int64_t)gridDim.x * blockDim.x)
        dst[r.dst + e] = FILL ? T(0) : src[r.src + e];
}

// The value ranges (blockIdx.y < nVal) and the rank ranges of one step in
// one launch.
template <typename T, bool FILL>
__global__ void ranges2_kernel(T* __restrict__ vals, int* __restrict__ ranks, const CopyRange* __restrict__ rs, int nVal)
{
    const int y = blockIdx.y;
    const CopyRange r = rs[y];
    if (y < nVal) {
        for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < r.len; e += (int64_t)gridDim.x * blockDim.x)
            vals[r.dst + e] = FILL ? T(0) : vals[r.src + e];
    } else {
        for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < r.len; e += (int64_t)gridDim.x * blockDim.x)
            ranks[r.dst + e] = FILL ? 0 : ranks[r.src + e];
    }
}

template <typename T>
void launch_arena_ranges(T* vals, int* ranks, const CopyRange* ranges, int nVal, int nRank, int64_t maxLen, bool fill,
                         hipStream_t s)
{
    const int count = nVal + nRank;
    if (count <= 0 || maxLen <= 0) return;
    if (count > 65535) throw std::invalid_argument("too many ranges in one launch");
    const dim3 grid((unsigned)std::min<int64_t>((maxLen + 255) / 256, 2048), (unsigned)count);
    if (fill) ranges2_kernel<T, true><<<grid, 256, 0, s>>>(vals, ranks, ranges, nVal);
    else ranges2_kernel<T, false><<<grid, 256, 0, s>>>(vals, ranks, ranges, nVal);
    HIP_LAUNCH_CHECK();
}

template <typename T>
void launch_copy_ranges(const T* src, T* dst, const CopyRange* ranges, int count, int64_t maxLen, bool fill,
                        hipStream_t s)
{
    if (count <= 0 || maxLen <= 0) return;
    if (count > 65535) throw std::invalid_argument("too many ranges in one launch");
    const dim3 grid((unsigned)std::min<int64_t>((maxLen + 255) / 256, 2048), (unsigned)count);
    if (fill) ranges_kernel<T, true><<<grid, 256, 0, s>>>(src, dst, ranges);
    else ranges_kernel<T, false><<<grid, 256, 0, s>>>(src, dst, ranges);
    HIP_LAUNCH_CHECK();
}

// n x n identities at base + offs[i]; blockIdx.y selects the identity.
template <typename T>
__global__ void eye_multi_kernel(T* base, const int64_t* __restrict__ offs, int n)
{
    T* x = base + offs[blockIdx.y];
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += gridDim.x * blockDim.x)
        x[e] = (e % n == e / n) ? T(1) : T(0);
}

template <typename T>
void launch_eye_multi(T* base, const int64_t* offs, int count, int n, hipStream_t s)
{
    if (count <= 0 || n <= 0) return;
    const dim3 grid((unsigned)std::min((n * n + 255) / 256, 64), (unsigned)count);
    eye_multi_kernel<T><<<grid, 256, 0, s>>>(base, offs, n);
    HIP_LAUNCH_CHECK();
}

// In-place Gauss-Jordan inversion with partial pivoting of one m x m
// column-major matrix (leading dimension m) by one workgroup of NT threads:
// at step k the pivot row p (largest |a(i,k)|, i >= k) is swapped into row k,
// then every entry is updated from the saved column k and the scaled row k;
// the row swaps are undone as column swaps in reverse order at the end. The
// matrix is held in LDS when `inLds` (loaded from src, stored to dst), else
// worked on in dst (copied from src first). A zero pivot sets *singular.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void dense_inverse_kernel(const T* __restrict__ src, T* __restrict__ dst, int m,
                                                           bool inLds, int* __restrict__ singular)
{
    extern __shared__ __align__(16) unsigned char smem[];
    T* col = reinterpret_cast<T*>(smem);              // saved column k (m)
    T* row = col + m;                                 // scaled row k (m)
    int* piv = reinterpret_cast<int*>(row + m);       // pivot rows (m)
    __shared__ int pivot;
    T* a = inLds ? reinterpret_cast<T*>(piv + ((m + 3) & ~3)) : dst;
    const int tid = threadIdx.x;
    const int mm = m * m;
    for (int e = tid; e < mm; e += NT) a[e] = src[e];
    __syncthreads();
    for (int k = 0; k < m; ++k) {
        if (tid < 64) {
            // Pivot search by the first wave: largest magnitude, lowest row on ties.
            double best = -1.0;
            int bi = k;
            for (int i = k + tid; i < m; i += 64) {
                const double v = fabs((double)a[i + (int64_t)k * m]);
                if (v > best) { best = v; bi = i; }
            }
            for (int o = 32; o > 0; o >>= 1) {
                const double ov = __shfl_xor(best, o, 64);
                const int oi = __shfl_xor(bi, o, 64);
                if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
            }
            if (tid == 0) { pivot = bi; piv[k] = bi; }
        }
        __syncthreads();
        const int p = pivot;
        if (p != k)
            for (int j = tid; j < m; j += NT) {
                const T t = a[k + (int64_t)j * m];
                a[k + (int64_t)j * m] = a[p + (int64_t)j * m];
                a[p + (int64_t)j * m] = t;
            }
        __syncthreads();
        const T d = a[k + (int64_t)k * m];
        if (d == T(0) && tid == 0) atomicExch(singular, 1);
        const T inv = d == T(0) ? T(0) : T(1) / d;
        for (int i = tid; i < m; i += NT) {
            col[i] = a[i + (int64_t)k * m];
            row[i] = a[k + (int64_t)i * m] * inv;
        }
        __syncthreads();
        for (int e = tid; e < mm; e += NT) {
            const int i = e % m, j = e / m;
            T v;
            if (i == k) v = j == k ? inv : row[j];
            else v = j == k ? -col[i] * inv : a[e] - col[i] * row[j];
            a[e] = v;
        }
        __syncthreads();
    }
    for (int k = m - 1; k >= 0; --k) {
        const int p = piv[k];
        if (p != k)
            for (int i = tid; i < m; i += NT) {
                const T t = a[i + (int64_t)k * m];
                a[i + (int64_t)k * m] = a[i + (int64_t)p * m];
                a[i + (int64_t)p * m] = t;
            }
        __syncthreads();
    }
    if (inLds)
        for (int e = tid; e < mm; e += NT) dst[e] = a[e];
}

// The same inversion for m <= 64 with the matrix in registers, by 64/RW
// waves: lane c holds column c, wave g rows RW*g .. RW*g+RW-1. Step k: every
// lane takes the largest magnitude of its own column over its rows >= k and
// the wave's candidate is lane k's (readlane); the waves' candidates meet in
// LDS and every thread picks the same p (largest magnitude, lowest row on
// ties); the owners of rows p and k publish them in LDS, and each thread
// updates its entries from the published pivot row, reading the multipliers
// (column k) from lane k and exchanging rows k and p on the fly. Two barriers
// per step; more waves with fewer rows each hide the latency of the step's
// dependent arithmetic. The row exchanges are undone as column exchanges by
// composing them into one column permutation (sigma(c) = s_m-1(...s_0(c)),
// s_k swapping k and p_k), applied while the result is staged through LDS for
// a coalesced store.
template <typename T, int RW>
__global__ void __launch_bounds__(64 * (64 / RW)) dense_inverse_reg_kernel(const T* __restrict__ src,
                                                                          T* __restrict__ dst, int m,
                                                                          int* __restrict__ singular)
{
    constexpr int NWV = 64 / RW, LD = 65;
    __shared__ double tile[64 * LD];          // tile[j * LD + i] = a(i, j)
    __shared__ double prow[64], krow[64];
    __shared__ double cval[NWV];
    __shared__ int cidx[NWV];
    __shared__ int sigma[64];
    const int tid = threadIdx.x, c = tid & 63, g = tid >> 6, r0 = g * RW;
    for (int j = g; j < 64; j += NWV)
        tile[j * LD + c] = (j < m && c < m) ? (double)src[c + (int64_t)j * m] : 0.0;
    __syncthreads();
    double a[RW];
#pragma unroll
    for (int r = 0; r < RW; ++r) a[r] = tile[c * LD + r0 + r];
    int x = c;                                // source column of result column c
#pragma unroll 1
    for (int k = 0; k < m; ++k) {
        double best = -1.0;
        int bi = 0;
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            const int i = r0 + r;
            const double v = (i >= k && i < m) ? fabs(a[r]) : -1.0;
            if (v > best) { best = v; bi = i; }
        }
        best = bcast(best, k);
        bi = __builtin_amdgcn_readlane(bi, k);
        if (c == 0) { cval[g] = best; cidx[g] = bi; }
        __syncthreads();
        double pb = cval[0];
        int p = cidx[0];
#pragma unroll
        for (int q = 1; q < NWV; ++q)
            if (cval[q] > pb) { pb = cval[q]; p = cidx[q]; }
        p = __builtin_amdgcn_readfirstlane(p);
        const int gk = k / RW, rk = k % RW, gp = p / RW, rp = p % RW;
        if (g == gp) {
            double v = 0.0;
#pragma unroll
            for (int r = 0; r < RW; ++r) v = (r == rp) ? a[r] : v;
            prow[c] = v;
        }
        if (g == gk) {
            double v = 0.0;
#pragma unroll
            for (int r = 0; r < RW; ++r) v = (r == rk) ? a[r] : v;
            krow[c] = v;
        }
        __syncthreads();
        const double d = prow[k];             // the pivot
        const double fp = krow[k];            // row p's new entry in column k (old a(k,k))
        const double oldk = krow[c];          // row p's new entry in column c
        if (d == 0.0 && tid == 0) atomicExch(singular, 1);
        const double inv = d == 0.0 ? 0.0 : 1.0 / d;
        const double rkc = prow[c] * inv;     // pivot row, scaled
#pragma unroll
        for (int r = 0; r < RW; ++r) {
            const int i = r0 + r;
            const double ck = bcast(a[r], k);  // this row's entry in column k
            const bool swapped = (i == p) && (p != k);
            const double f = swapped ? fp : ck;
            const double cur = swapped ? oldk : a[r];
            double v;
            if (i == k) v = (c == k) ? inv : rkc;
            else v = (c == k) ? -f * inv : cur - f * rkc;
            a[r] = v;
        }
        x = (x == k) ? p : ((x == p) ? k : x);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RW; ++r) tile[c * LD + r0 + r] = a[r];
    sigma[c] = x;
    __syncthreads();
    for (int j = g; j < m; j += NWV)
        if (c < m) dst[c + (int64_t)j * m] = (T)tile[sigma[j] * LD + c];
}

// Rows per wave of the register inversion (8 waves of 8 rows: 108 us for a
// 64 x 64 leaf against 122 us with 4 waves, 166 us with 16 and 178 us for
// the LDS kernel with 512 threads; publishing the candidate rows to save the
// second barrier gained 3%).
constexpr int kInvRows = 8;

template <typename T>
void launch_dense_inverse(const T* src, T* dst, int m, int* singular, hipStream_t s)
{
    if (m <= 0) return;
    if (m <= 64) {
        dense_inverse_reg_kernel<T, kInvRows><<<1, 64 * (64 / kInvRows), 0, s>>>(src, dst, m, singular);
        HIP_LAUNCH_CHECK();
        return;
    }
    const size_t vecs = 2 * (size_t)m * sizeof(T) + (((size_t)m + 3) & ~size_t(3)) * sizeof(int);
    const size_t full = vecs + (size_t)m * m * sizeof(T);
    const bool inLds = full <= 64 * 1024;
    if (vecs > 64 * 1024) throw std::invalid_argument("dense leaf too large to invert");
    dense_inverse_kernel<T, 512><<<1, 512, inLds ? full : vecs, s>>>(src, dst, m, inLds, singular);
    HIP_LAUNCH_CHECK();
}

#define INST(T)                                                                                          \
    template void launch_copy_ranges<T>(const T*, T*, const CopyRange*, int, int64_t, bool, hipStream_t); \
    template void launch_arena_ranges<T>(T*, int*, const CopyRange*, int, int, int64_t, bool, hipStream_t); \
    template void launch_eye_multi<T>(T*, const int64_t*, int, int, hipStream_t);                      \
    template void launch_dense_inverse<T>(const T*, T*, int, int*, hipStream_t);
INST(double)
INST(float)
template void launch_copy_ranges<int>(const int*, int*, const CopyRange*, int, int64_t, bool, hipStream_t);

} // namespace dmhm

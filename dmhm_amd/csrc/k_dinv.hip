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

// Largest unsigned value over the wave (DPP within rows of 16, then the four
// row maxima through scalar registers); the result is uniform.
__device__ inline uint32_t wave_max_u32(uint32_t x)
{
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kDppXor1, 0xF, 0xF, false));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kDppXor2, 0xF, 0xF, false));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kDppMirror8, 0xF, 0xF, false));
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, kDppMirror16, 0xF, 0xF, false));
    const uint32_t a = __builtin_amdgcn_readlane(x, 0), b = __builtin_amdgcn_readlane(x, 16);
    const uint32_t c = __builtin_amdgcn_readlane(x, 32), e = __builtin_amdgcn_readlane(x, 48);
    return max(max(a, b), max(c, e));
}

// Inverse of one m x m leaf (m <= 64): Gauss-Jordan with partial pivoting,
// lane = row, the 64 columns split over NWV waves (wave w holds columns
// [w CW, (w + 1) CW) of every row in registers). Three departures from the
// textbook form keep each step to one barrier and one LDS exchange:
// - rows are not exchanged: step k takes the largest |u(i,k)| over the rows
//   not yet used as pivots (compared on the high word of the double: sign
//   cleared, exponent and 20 mantissa bits, ties to the lowest row) and
//   records the pivot row p_k. The elimination then leaves
//   inv(A)(k, p_j) = T(p_k, j) for the final array T, applied at the store;
// - the pivot row is not scaled by 1/u(p,k) at its step: the factor of every
//   other row is u(i,k) / u(p,k) either way, so the scale is kept per row and
//   applied at the store (the row's entries stay, its column-k entry is 1);
// - the wave that owns column k chooses the pivot and publishes the pivot
//   row's index, 1/u(p,k) and every row's factor through LDS; every wave then
//   updates only its own columns, reading its share of the pivot row from its
//   own lane p (v_readlane). Buffers alternate by step parity, so a step's
//   writes never meet the previous step's reads.
// Padding rows and columns beyond m hold the identity; their steps (up to
// CW - 1 in the last panel) change nothing.
template <typename T, int NWV>
__global__ void __launch_bounds__(64 * NWV) dense_inverse_split_kernel(const T* __restrict__ src,
                                                                      T* __restrict__ dst, int m,
                                                                      int* __restrict__ singular)
{
    constexpr int CW = 64 / NWV;
    static_assert(CW % 2 == 0, "step parity follows the column within a wave");
    __shared__ double gbuf[2][64], ibuf[2];
    __shared__ int pbuf[2];
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double u[CW];
#pragma unroll
    for (int j = 0; j < CW; ++j) {
        const int c = w * CW + j;
        u[j] = (c < m && lane < m) ? (double)src[lane + (int64_t)c * m] : (c == lane ? 1.0 : 0.0);
    }
    bool used = false, sing = false;
    double scale = 1.0;
    int stepOf = 0, pivOf = 0;      // this row's step; the pivot row of step `lane`
    const int panels = (m + CW - 1) / CW;
#pragma unroll 1
    for (int kb = 0; kb < panels; ++kb) {
        const bool owner = w == kb;
#pragma unroll
        for (int kk = 0; kk < CW; ++kk) {
            const int k = kb * CW + kk, par = kk & 1;
            if (owner) {
                const double v = u[kk];
                const uint32_t hi = used ? 0u : ((uint32_t)__double2hiint(v) & 0x7fffffffu);
                const uint32_t mx = wave_max_u32(hi);
                const unsigned long long cand = __ballot(!used && hi == mx);
                const int p = __builtin_amdgcn_readfirstlane(__ffsll(cand) - 1);
                const double dd = bcast(v, p);
                const double inv = dd == 0.0 ? 0.0 : 1.0 / dd;
                gbuf[par][lane] = lane == p ? 0.0 : v * inv;
                if (lane == 0) { pbuf[par] = p; ibuf[par] = inv; }
                sing = sing || dd == 0.0;
            }
            __syncthreads();
            const int p = __builtin_amdgcn_readfirstlane(pbuf[par]);
            const double inv = ibuf[par];
            const double g = gbuf[par][lane];
            const bool me = lane == p;
#pragma unroll
            for (int j = 0; j < CW; ++j) {
                if (j == kk && owner) u[j] = me ? 1.0 : -g;
                else u[j] = fma(-g, bcast(u[j], p), u[j]);
            }
            scale = me ? inv : scale;
            used = used || me;
            stepOf = me ? k : stepOf;
            pivOf = lane == k ? p : pivOf;
        }
    }
    if (sing && lane == 0) atomicExch(singular, 1);
    const bool row = lane < m && stepOf < m;
#pragma unroll
    for (int j = 0; j < CW; ++j) {
        const int c = w * CW + j;
        const int pc = __builtin_amdgcn_readlane(pivOf, c);
        if (row && c < m && pc < m) dst[stepOf + (int64_t)pc * m] = (T)(u[j] * scale);
    }
}

// Waves of the leaf inversion (a 64 x 64 FP64 leaf: 28 us with 8 waves or
// 16, 32 us with 4, 66 us with one wave that broadcasts the whole pivot row
// by lane reads, against 108 us for lane = column with rows split over 8
// waves and two barriers per step).
constexpr int kInvWaves = 8;

template <typename T>
void launch_dense_inverse(const T* src, T* dst, int m, int* singular, hipStream_t s)
{
    if (m <= 0) return;
    if (m <= 64) {
        dense_inverse_split_kernel<T, kInvWaves><<<1, 64 * kInvWaves, 0, s>>>(src, dst, m, singular);
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

// gfx950 Gaussian sketch generator, elementwise and per-leaf helpers, and
// the power-iteration reductions.
#include "kcommon.hpp"

namespace dmhm {

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

// Seeds may carry a device-resident offset (`seedOff`, or null): an iteration
// counter that a captured HIP graph advances on every replay, so replays draw
// fresh sketches exactly as the eager loop does. The counter is spread by an
// odd 64-bit multiplier so that seeds of different roles (base + 0, 1, 2)
// never meet across iterations.
__device__ inline uint64_t seed_with(uint64_t seed0, const uint64_t* seedOff)
{
    return seed0 + (seedOff ? *seedOff * 0x9E3779B97F4A7C15ULL : 0);
}
template <typename T>
__global__ void __launch_bounds__(256) rng_kernel(Bases b, const RngTask* __restrict__ tasks,
                                                  uint64_t seed0, const uint64_t* seedOff)
{
    const uint64_t seed = seed_with(seed0, seedOff);
    const RngTask tk = tasks[blockIdx.x];
    T* x = reinterpret_cast<T*>(b.val[SLOT_WS]) + tk.off;
    const int64_t n = (int64_t)tk.m * tk.k;
    for (int64_t e = threadIdx.x; e < n; e += 256) x[e] = (T)gauss_at(seed, tk.stream, (uint64_t)e);
}

template <typename T>
void launch_rng_phase(const Phase& ph, const Bases& b, uint64_t seed, hipStream_t s, const uint64_t* seedOff)
{
    const int n = (int)ph.rng.size();
    if (n == 0) return;
    rng_kernel<T><<<n, 256, 0, s>>>(b, static_cast<const RngTask*>(ph.dTasks), seed, seedOff);
    HIP_LAUNCH_CHECK();
}

template <typename T>
__global__ void gaussian_kernel(T* x, int64_t n, uint64_t seed0, uint64_t stream, const uint64_t* seedOff)
{
    const uint64_t seed = seed_with(seed0, seedOff);
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
        x[e] = (T)gauss_at(seed, stream, (uint64_t)e);
}

template <typename T> void launch_gaussian(T* x, int64_t n, uint64_t seed, uint64_t stream, hipStream_t s,
                                           const uint64_t* seedOff)
{
    if (n <= 0) return;
    gaussian_kernel<T><<<grid_for(n), 256, 0, s>>>(x, n, seed, stream, seedOff);
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

// Power iteration over a batch of vectors, one per member (segment): the
// start vectors (element e of member m drawn as element e of a lone vector,
// so a member's draws do not depend on the batch), and the normalisation with
// the log of the norm accumulated per member over its first `steps` steps.
template <typename T>
__global__ void seg_gaussian_kernel(T* x, const VecSegment* __restrict__ segs, uint64_t seed0, uint64_t stream,
                                    const uint64_t* seedOff)
{
    const VecSegment sg = segs[blockIdx.y];
    const uint64_t seed = seed_with(seed0, seedOff);
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < sg.n; e += (int64_t)gridDim.x * blockDim.x)
        x[sg.off + e] = (T)gauss_at(seed, stream, (uint64_t)e);
}
template <typename T>
void launch_seg_gaussian(T* x, const VecSegment* segs, int count, int64_t maxN, uint64_t seed, uint64_t stream,
                         hipStream_t s, const uint64_t* seedOff)
{
    if (count <= 0 || maxN <= 0) return;
    const dim3 grid((unsigned)std::min<int64_t>((maxN + 255) / 256, 256), (unsigned)count);
    seg_gaussian_kernel<T><<<grid, 256, 0, s>>>(x, segs, seed, stream, seedOff);
    HIP_LAUNCH_CHECK();
}

template <typename T>
__global__ void __launch_bounds__(1024) seg_normalize_kernel(T* x, const VecSegment* __restrict__ segs,
                                                             double* logAcc, int step)
{
    __shared__ double red[16];
    __shared__ double total;
    const VecSegment sg = segs[blockIdx.x];
    T* v = x + sg.off;
    double s = 0;
    for (int64_t e = threadIdx.x; e < sg.n; e += 1024) {
        const double a = (double)v[e];
        s += a * a;
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int w = 0; w < 16; ++w) t += red[w];
        total = t;
        if (logAcc && step >= 0 && step < sg.steps) logAcc[blockIdx.x] += 0.5 * log(t) / sg.steps;
    }
    __syncthreads();
    const double inv = total > 0 ? 1.0 / sqrt(total) : 0.0;
    for (int64_t e = threadIdx.x; e < sg.n; e += 1024) v[e] = (T)((double)v[e] * inv);
}
template <typename T>
void launch_seg_normalize(T* x, const VecSegment* segs, int count, double* logAcc, int step, hipStream_t s)
{
    if (count <= 0) return;
    seg_normalize_kernel<T><<<count, 1024, 0, s>>>(x, segs, logAcc, step);
    HIP_LAUNCH_CHECK();
}

__global__ void bump_kernel(uint64_t* c) { if (threadIdx.x == 0) *c += 1; }
void launch_bump(uint64_t* counter, hipStream_t s)
{
    bump_kernel<<<1, 64, 0, s>>>(counter);
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
    template void launch_rng_phase<T>(const Phase&, const Bases&, uint64_t, hipStream_t, const uint64_t*); \
    template void launch_scale<T>(T*, int64_t, T, hipStream_t);                                     \
    template void launch_fill<T>(T*, int64_t, T, hipStream_t);                                      \
    template void launch_eye<T>(T*, int, hipStream_t);                                              \
    template void launch_axpby<T>(const T*, const T*, T, T*, int64_t, hipStream_t);                 \
    template void launch_gaussian<T>(T*, int64_t, uint64_t, uint64_t, hipStream_t, const uint64_t*); \
    template void launch_add_diag<T>(T*, const int64_t*, const int*, int, T, hipStream_t);          \
    template void launch_transpose_dense<T>(const T*, T*, const int64_t*, const int64_t*, const int*, \
                                            const int*, int, hipStream_t);                          \
    template void launch_swap_lowrank<T>(const T*, T*, const int64_t*, const int64_t*, const int64_t*, \
                                         const int64_t*, const int*, const int*, const int*, int,   \
                                         hipStream_t);                                              \
    template void launch_sumsq<T>(const T*, int64_t, double*, hipStream_t);                         \
    template void launch_normalize<T>(T*, int64_t, const double*, double*, double, hipStream_t);          \
    template void launch_seg_gaussian<T>(T*, const VecSegment*, int, int64_t, uint64_t, uint64_t, hipStream_t, \
                                         const uint64_t*);                                           \
    template void launch_seg_normalize<T>(T*, const VecSegment*, int, double*, int, hipStream_t);
INST(double)
INST(float)
template void launch_convert<double, float>(const double*, float*, int64_t, hipStream_t);
template void launch_convert<float, double>(const float*, double*, int64_t, hipStream_t);


} // namespace dmhm

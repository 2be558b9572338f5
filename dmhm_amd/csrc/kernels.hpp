// Host-side launchers for the gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "plan.hpp"

namespace dmhm {

struct Bases {
    void* val[NUM_SLOTS];
    int* rank[NUM_SLOTS];
};

// Truncation policy of the SVD phase.
struct TruncPolicy {
    int tolMode;        // 0 fixed rank, 1 relative tolerance
    double relTol;
};

// Tile engines: TILE_LDS, the workgroup-per-tile kernel staging operands
// through LDS (vector FMA); TILE_MFMA, the wave-per-row-block MFMA kernel in
// the storage precision; TILE_MFMA_BF16, the same with FP32 operands rounded
// to BF16 and FP32 accumulation (FP64 storage falls back to TILE_MFMA); the
// _PF forms prefetch the next chunk before the current one's MFMAs.
enum TileEngine { TILE_LDS = 0, TILE_MFMA = 1, TILE_MFMA_PF = 2, TILE_MFMA_BF16 = 3, TILE_MFMA_BF16_PF = 4,
                  TILE_ENGINES = 5 };
template <typename T>
void launch_tile_phase(const Phase& ph, const Bases& b, T alpha, T beta, hipStream_t s, int engine);
template <typename T>
void launch_qr_phase(const Phase& ph, const Bases& b, hipStream_t s);
template <typename T>
void launch_svd_phase(const Phase& ph, const Bases& b, TruncPolicy pol, hipStream_t s);
template <typename T>
void launch_mv_phase(const Phase& ph, const Bases& b, T alpha, T beta, hipStream_t s);
template <typename T>
void launch_rng_phase(const Phase& ph, const Bases& b, uint64_t seed, hipStream_t s,
                      const uint64_t* seedOff = nullptr);

// Elementwise helpers over raw device arrays.
template <typename T> void launch_scale(T* x, int64_t n, T alpha, hipStream_t s);
template <typename T> void launch_fill(T* x, int64_t n, T v, hipStream_t s);
// out = a + alpha * b (elementwise)
template <typename T> void launch_axpby(const T* a, const T* b, T alpha, T* out, int64_t n, hipStream_t s);
void launch_fill_int(int* x, int64_t n, int v, hipStream_t s);
// x (n x n, ld n) = identity
template <typename T> void launch_eye(T* x, int n, hipStream_t s);
template <typename Ts, typename Td> void launch_convert(const Ts* x, Td* y, int64_t n, hipStream_t s);
// Seeds add the device word *seedOff when it is given (graph-replayed loops).
template <typename T> void launch_gaussian(T* x, int64_t n, uint64_t seed, uint64_t stream, hipStream_t s,
                                           const uint64_t* seedOff = nullptr);
// *counter += 1 on the device (the iteration counter of a replayed loop).
void launch_bump(uint64_t* counter, hipStream_t s);

// Per-leaf operations driven by small index tables on the device.
// add alpha to the diagonal of square dense leaves: offsets + sizes.
template <typename T> void launch_add_diag(T* vals, const int64_t* offs, const int* sizes, int count,
                                           T alpha, hipStream_t s);
// dst dense leaf i = transpose of src dense leaf perm[i] (m x n -> n x m).
template <typename T> void launch_transpose_dense(const T* src, T* dst, const int64_t* srcOff,
                                                  const int64_t* dstOff, const int* m, const int* n,
                                                  int count, hipStream_t s);
// dst U/V of low-rank block i = src V/U of block perm[i]; ranks copied.
template <typename T> void launch_swap_lowrank(const T* src, T* dst, const int64_t* srcU,
                                               const int64_t* srcV, const int64_t* dstU,
                                               const int64_t* dstV, const int* mrows,
                                               const int* nrows, const int* cap, int count,
                                               hipStream_t s);
void launch_permute_int(const int* src, int* dst, const int* perm, int count, hipStream_t s);

// Direct inversion and merged plans (k_dinv.hip). Ranges are element ranges
// of one allocation: dst[d + e] = src[s + e] for e < len.
struct CopyRange { int64_t src, dst, len; };
template <typename T> void launch_copy_ranges(const T* src, T* dst, const CopyRange* ranges, int count,
                                              int64_t maxLen, bool fill, hipStream_t s);
// Value ranges (the first nVal) and rank ranges of one allocation pair in one
// launch; maxLen bounds every range's length.
template <typename T> void launch_arena_ranges(T* vals, int* ranks, const CopyRange* ranges, int nVal, int nRank,
                                               int64_t maxLen, bool fill, hipStream_t s);
// count n x n identities at base + offs[i] (offs on the device)
template <typename T> void launch_eye_multi(T* base, const int64_t* offs, int count, int n, hipStream_t s);
// dst (m x m, ld m) = inverse of src (Gauss-Jordan, partial pivoting); a zero
// pivot sets *singular (device int).
template <typename T> void launch_dense_inverse(const T* src, T* dst, int m, int* singular, hipStream_t s);

// Vector utilities for the power iteration (EstimateTwoNorm).
template <typename T> void launch_sumsq(const T* x, int64_t n, double* out, hipStream_t s);
template <typename T> void launch_normalize(T* x, int64_t n, const double* sumsq, double* logAcc,
                                            double root, hipStream_t s);
void launch_zero_double(double* x, int n, hipStream_t s);
// The same over a batch: member m's vector is x[off, off + n) and takes
// `steps` power steps (its EstimateTwoNorm count); logAcc[m] accumulates.
struct VecSegment { int64_t off; int n; int steps; };
template <typename T> void launch_seg_gaussian(T* x, const VecSegment* segs, int count, int64_t maxN, uint64_t seed,
                                               uint64_t stream, hipStream_t s, const uint64_t* seedOff = nullptr);
template <typename T> void launch_seg_normalize(T* x, const VecSegment* segs, int count, double* logAcc, int step,
                                                hipStream_t s);

// Device copy of a block tree for the CSR import (k_build.hip). Boxes are
// flattened level by level (box (l, id) at lvlBase[l] + id); boxChild holds
// flattened child indices, nc per box; block arrays follow Geometry::blocks.
struct GeoDev {
    int dim = 2, nc = 4, numLevels = 1, n = 0;
    const int *lvlBase = nullptr, *boxOff = nullptr, *boxSize = nullptr, *boxChild = nullptr;
    const int *blkType = nullptr, *blkLevel = nullptr, *blkT = nullptr, *blkS = nullptr;
    const int *blkChildBegin = nullptr, *blkLr = nullptr, *blkDense = nullptr, *children = nullptr;
    const int* lrBlock = nullptr;
};
struct LayDev {
    const int64_t *denseOff = nullptr, *uOff = nullptr, *vOff = nullptr;
};
void launch_csr_locate(const GeoDev& g, const int* rowptr, const int* colind, int n, int* nzBlock, int* lrCount,
                       int* bad, hipStream_t s);
template <typename T>
void launch_csr_fill(const GeoDev& g, const LayDev& lay, const int* rowptr, const int* colind, const double* vals,
                     int n, const int* nzBlock, const int* lrBase, int* lrFill, const int* lrCount, int* stageK,
                     int* rowOf, int nLr, T* dst, hipStream_t s);
void launch_natural_map(const GeoDev& g, int nx, int ny, int nz, int* map, hipStream_t s);

// Rank model (k_mlp.hip): 3-layer perceptron, NF inputs, 64 ReLU units per
// hidden layer; mlp_param_count(NF) floats of parameters.
int mlp_param_count(int nf);
void launch_mlp_infer(const float* params, int nf, const float* X, int n, float* y, int* widths, float margin,
                      int extra, int wmin, int wmax, hipStream_t s);
void launch_mlp_grad(const float* params, int nf, const float* X, const float* t, int n, float* grads,
                     float* loss, hipStream_t s);
void launch_mlp_adam(float* p, float* g, float* m, float* v, int np, float lr, int step, hipStream_t s);
void launch_rank_features(const float* stat, int nstat, const int* prev, float invMax, float* out, int n,
                          hipStream_t s);
void launch_widen_flagged(int* widths, int* flags, int n, int wmax, int* count, long long* wsum, hipStream_t s);

// Device scratch of the allocation-free H*X (noalloc.hip): scratch_setup takes
// the array's address on `device` (which must be current) once;
// scratch_address returns it (nullptr before set-up) and its size in bytes.
void scratch_setup(int device);
void* scratch_address(int device, size_t* bytes);

} // namespace dmhm

/*
 * C-ABI of the MI355X-native H-matrix library (libdmhm_amd.so).
 *
 * Plain pointers and sizes only. Each entry point replaces a member or free
 * function of the reference's sequential classes; the reference interface it
 * stands for is cited beside it (paths relative to the reference tree). The
 * header-only C++ facade include/dmhm_amd/hmat.hpp re-exposes these with the
 * reference's class and method names (dmhm::HMat2d<Scalar>, HMat3d<Scalar>,
 * hmat_tools::...).
 *
 * Conventions
 *  - Dense operands are column-major with a leading dimension, as
 *    Dense<Scalar> (include/dmhm/core/dense.hpp:24-88).
 *  - Unknowns are in hierarchical ordering (dmhm_build_natural_to_hierarchical_map).
 *  - Low-rank blocks are U V^T (transpose convention, core/low_rank.hpp:23-35).
 *  - Every function returns DMHM_OK or an error code; dmhm_last_error()
 *    gives the text. The reference throws only in debug builds
 *    (std::logic_error / std::runtime_error); the facade maps codes back to
 *    those exception types.
 *  - Real scalars only (float and double storage, selected per handle).
 *  - Pointer arguments flagged DMHM_DEVICE are device (HBM) pointers; with
 *    DMHM_HOST the library copies through pinned staging buffers.
 *  - One handle is used by one host thread at a time; calls are ordered on
 *    the library stream and return after the GPU work has finished.
 */
#ifndef DMHM_AMD_H
#define DMHM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum dmhm_status {
    DMHM_OK = 0,
    DMHM_EINVAL = 1,      /* bad argument (reference: std::logic_error)          */
    DMHM_ENONCONF = 2,    /* nonconformal operands (MultiplyHMat-incl.hpp:19-25) */
    DMHM_ESINGULAR = 3,   /* singular dense leaf                                 */
    DMHM_ENOCONV = 4,     /* Schulz did not reach the stop tolerance             */
    DMHM_EHIP = 5,        /* HIP runtime error                                   */
    DMHM_ENOMEM = 6,      /* device allocation failed                            */
    DMHM_ENODEV = 7       /* no GPU / library not initialised                    */
};

enum dmhm_precision { DMHM_F64 = 8, DMHM_F32 = 4 };
enum dmhm_memory { DMHM_HOST = 0, DMHM_DEVICE = 1 };

typedef struct dmhm_hmat_s* dmhm_hmat;
typedef struct dmhm_hbatch_s* dmhm_hbatch;

/* Library set-up: binds device `device`, creates the library stream.
 * Reference: dmhm::Initialize (src/core/global.cpp:55-94; MPI and the global
 * RNG seeding are replaced by per-handle seeds). */
int dmhm_init(int device);
int dmhm_finalize(void);
const char* dmhm_last_error(void);
/* Device synchronisation point used by benchmarks. */
int dmhm_synchronize(void);

/* HMat2d::BuildNaturalToHierarchicalMap (src/HMat2d/HMat2d.cpp:69-87),
 * HMat3d::BuildNaturalToHierarchicalMap (src/HMat3d/HMat3d.cpp:93-111).
 * dim = 2 or 3, sizes[dim]; map[natural] = hierarchical. Host only. */
int dmhm_build_natural_to_hierarchical_map(int dim, const int* sizes, int numLevels, int* map);
/* The same map computed on the GPU (one thread per grid point descends the
 * box tree); map is a host array. Needs dmhm_init. */
int dmhm_build_natural_to_hierarchical_map_device(int dim, const int* sizes, int numLevels, int* map);

/* Block counts per level of the tree HMat2d/HMat3d builds for a geometry
 * (ImportSparse recursion, src/HMat2d/HMat2d.cpp:961-1033; admissibility
 * :790-809). Host only; no device needed. */
int dmhm_geometry_block_counts(int dim, const int* sizes, int numLevels, int strong,
                               long long* dense, long long* lowRank, int maxLevels);

/* Leaf blocks of the tree in DFS order, 7 ints each: type (0 low rank,
 * 1 dense), level, row offset, rows, column offset, columns, index within
 * its type (the order of dmhm_hmat_ranks). Returns the count through *count;
 * writes at most maxLeaves records. Host only. */
int dmhm_geometry_leaves(int dim, const int* sizes, int numLevels, int strong, int* out,
                         int maxLeaves, int* count);

/* Term counts of one H*H product C = A B on a geometry, from the product
 * recursion (src/HMat2d/MultiplyHMat-incl.hpp:12-380): out[0] NODE*NODE terms
 * into admissible blocks (the randomized-sketch case), out[1] dense*dense
 * terms into admissible blocks, out[2] dense*dense terms into dense leaves,
 * out[3] low-rank terms into admissible or dense blocks, out[4] low-rank
 * terms pushed into NODE blocks. Host only. */
int dmhm_geometry_product_terms(int dim, const int* sizes, int numLevels, int strong,
                                long long* out);

/* HMat2d(const Sparse<Scalar>&, numLevels, maxRank, strong, xSize, ySize)
 * (src/HMat2d/HMat2d.cpp:144-162 + ImportSparse :961-1033) and the HMat3d
 * counterpart. CSR (host arrays) with global hierarchical indices, imported
 * on the GPU. Admissible blocks are stored exactly (rank = number of nonzeros
 * in the block, one unit column of U and one value column of V per nonzero
 * in CSR order, ConvertSubmatrix.cpp:57-119). */
int dmhm_hmat_create_from_csr(int dim, const int* sizes, int numLevels, int maxRank, int strong,
                              int precision, const int* rowOffsets, const int* columnIndices,
                              const double* nonzeros, dmhm_hmat* out);

/* HMat2d(numLevels, maxRank, symmetric=false, strong, xSize, ySize)
 * (src/HMat2d/HMat2d.cpp:104-117): a zero matrix with the full block tree
 * (low-rank blocks at rank 0, dense leaves zero). */
int dmhm_hmat_create_zero(int dim, const int* sizes, int numLevels, int maxRank, int strong,
                          int precision, dmhm_hmat* out);

int dmhm_hmat_destroy(dmhm_hmat h);

typedef struct {
    int dim, n, numLevels, maxRank, strong, precision;
    long long numDense, numLowRank;
    long long storageBytes;      /* bytes of dense + U + V pools         */
    long long sumRank;           /* sum of current low-rank block ranks  */
    int maxBlockRank;
} dmhm_info;
/* Height/Width/NumLevels/MaxRank/StronglyAdmissible (hmat2d.hpp:96-104). */
int dmhm_hmat_info(dmhm_hmat h, dmhm_info* info);
/* Dense/low-rank counts per level (levels <= maxLevels). */
int dmhm_hmat_block_counts(dmhm_hmat h, long long* dense, long long* lowRank, int maxLevels);

/* HMat2d::SetToRandom (src/HMat2d/SetToRandom-incl.hpp:12-53): dense leaves
 * and U, V of every admissible block at rank = maxRank filled with N(0,1),
 * from a counter-based generator keyed by `seed`. */
int dmhm_hmat_set_to_random(dmhm_hmat h, uint64_t seed);

/* Y := alpha op(H) X + beta Y, op = 'N' or 'T' ('C' == 'T' for real data).
 * HMat2d::Multiply / TransposeMultiply / AdjointMultiply for Dense and
 * Vector (src/HMat2d/MultiplyDense-incl.hpp:12-262,
 * MultiplyVector-incl.hpp:12-245). X: n x k (ldx), Y: n x k (ldy), in the
 * handle's precision, host or device memory as `memory` says. */
int dmhm_hmat_multiply_dense(dmhm_hmat h, char op, double alpha, const void* X, int ldx, int k,
                             double beta, void* Y, int ldy, int memory);

/* Y := alpha H X + beta Y for device X, Y (n x k, 1 <= k <= 8) without any
 * device allocation: no counterpart in the reference; for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops). The
 * projections P = V^T X live in a 64 MiB device array per GPU held by the
 * library until the process ends; a workspace larger than that is rejected
 * (DMHM_EINVAL). The plan must have been built by one dmhm_hmat_multiply_dense
 * call with op 'N' and the same k, ldx, ldy (else DMHM_EINVAL). Calls that
 * overlap on one GPU race on the shared array. Returns after the GPU work. */
int dmhm_hmat_multiply_dense_noalloc(dmhm_hmat h, double alpha, const void* X, int ldx, int k,
                                     double beta, void* Y, int ldy);

/* Rank model: a 3-layer perceptron (DMHM_RANK_FEATURES inputs, two hidden
 * layers of 64 ReLU units, one linear output) predicting the numerical rank of
 * an admissible block of a product from the block's features (level, box
 * sizes and offsets, coefficient statistics of the operator, the block's rank
 * in the left operand). No counterpart in the reference, whose sketch width is
 * fixed at SampleRank(maxRank) = maxRank + Oversample()
 * (include/dmhm/hmat2d.hpp:32-33). */
typedef struct dmhm_rank_model_s* dmhm_rank_model;
#define DMHM_RANK_FEATURES 12

typedef struct {
    int oversample;      /* sketch width = maxRank + oversample (Oversample(), global.cpp:30) */
    int tolMode;         /* 0: fixed rank maxRank (reference); 1: keep s_i > relTol*s_0      */
    double relTol;       /* relative tolerance for tolMode 1 (Truncation-incl.hpp:37-63)     */
    uint64_t seed;       /* sketch seed                                                      */
    dmhm_rank_model rankModel; /* tolMode 1 only (else ignored): per-block sketch width =
                            predicted rank + rankMargin + oversample (at most maxRank +
                            oversample); blocks whose SVD keeps more than width -
                            oversample values are recomputed at full width. NULL: the
                            fixed width maxRank + oversample                              */
    double rankMargin;
} dmhm_hh_opts;

/* C := alpha A B, HMat2d::Multiply(alpha, B, C) (src/HMat2d/MultiplyHMat-incl.hpp:12-187).
 * Every admissible block of C is the rank-truncated randomized range
 * approximation of the exact restricted product (Gaussian sketch, Householder
 * QR, Jacobi SVD); dense leaves are exact. *C is created (same geometry). */
int dmhm_hmat_multiply_hmat(double alpha, dmhm_hmat A, dmhm_hmat B, dmhm_hmat* C,
                            const dmhm_hh_opts* opts);

/* C := alpha A B + beta C, HMat2d::Multiply(alpha, B, beta, C)
 * (src/HMat2d/MultiplyHMat-incl.hpp:189-380). C keeps its geometry; every
 * admissible block of C is recompressed from a sketch of the exact
 * alpha A B + beta C restricted to it. */
int dmhm_hmat_multiply_hmat_update(double alpha, dmhm_hmat A, dmhm_hmat B, double beta, dmhm_hmat C,
                                   const dmhm_hh_opts* opts);

/* A :~= alpha B + A, HMat2d::UpdateWith (src/HMat2d/HMat2d.cpp:746-784):
 * dense leaves summed exactly, admissible blocks recompressed to maxRank
 * (RoundedUpdate semantics, src/hmat_tools/RoundedUpdate.cpp:14-162). */
int dmhm_hmat_update_with(dmhm_hmat A, double alpha, dmhm_hmat B, const dmhm_hh_opts* opts);

/* HMat2d(const LowRank<Scalar>& F, numLevels, maxRank, strong, xSize, ySize)
 * (src/HMat2d/HMat2d.cpp:124-142 + ImportLowRank :811-904): admissible blocks
 * take the row slices of U (n x r) and V (n x r) exactly (rank r, not
 * truncated), dense leaves U_t V_s^T. Host arrays, column-major, ld = n. */
int dmhm_hmat_create_from_lowrank(int dim, const int* sizes, int numLevels, int maxRank, int strong,
                                  int precision, int r, const double* U, const double* V, dmhm_hmat* out);

/* hmat_tools::RoundedUpdate (src/hmat_tools/RoundedUpdate.cpp:14-162):
 * B :~= alpha A + beta B for m x n low-rank A = AU AV^T (rank rA) and
 * B = BU BV^T (rank rB). If rA + rB <= min(maxRank, m, n) the result is the
 * exact concatenation [beta BU, alpha AU], [BV, AV]; otherwise QR of both
 * stacked factors, SVD of R1 R2^T on the GPU, truncation to
 * min(maxRank, m, n) (rA + rB <= 64). outU (m x .), outV (n x .) receive the
 * result (capacity rA + rB columns), *outRank its rank. Host arrays. */
int dmhm_rounded_update(int maxRank, double alpha, int m, int n, int rA, const double* AU,
                        const double* AV, double beta, int rB, const double* BU, const double* BV,
                        double* outU, double* outV, int* outRank);

/* hmat_tools::Compress (src/hmat_tools/Compress.cpp:14-53): best rank
 * min(maxRank, m, n) approximation D ~= U V^T of an m x n dense matrix
 * (Householder QR + Jacobi SVD on the GPU; min(m, n) <= 64). Host arrays;
 * outU (m x min(m,n)), outV (n x min(m,n)) capacity. */
int dmhm_compress_dense(int maxRank, int m, int n, const double* D, double* outU, double* outV,
                        int* outRank);

/* HMat2d::CopyFrom / TransposeFrom / AdjointFrom / Scale / SetToIdentity /
 * AddConstantToDiagonal (src/HMat2d/HMat2d.cpp:331-744). */
int dmhm_hmat_copy(dmhm_hmat src, dmhm_hmat* dst);
int dmhm_hmat_transpose(dmhm_hmat src, dmhm_hmat* dst);
int dmhm_hmat_scale(dmhm_hmat h, double alpha);
int dmhm_hmat_set_to_identity(dmhm_hmat h);
int dmhm_hmat_add_constant_to_diagonal(dmhm_hmat h, double alpha);
/* Precision conversion (FP64 <-> FP32 copies of the pools). */
int dmhm_hmat_convert(dmhm_hmat src, int precision, dmhm_hmat* dst);

/* hmat_tools::EstimateTwoNorm (include/dmhm/hmat_tools.hpp:1504-1544):
 * k = ceil(log(0.8 sqrt(n) 10^confidence) / log(theta)) power steps on H from
 * a Gaussian start; estimate = theta * prod ||H x_i||^(1/k). */
int dmhm_hmat_estimate_two_norm(dmhm_hmat h, double theta, double confidence, uint64_t seed,
                                double* estimate);

typedef struct {
    int numIterations;   /* > 0: fixed count; <= 0: stop on the residual estimate */
    int maxIterations;   /* cap in stop mode (reference: 30, Invert-incl.hpp:145-149) */
    double theta;        /* 1.5 */
    double confidence;   /* 6   */
    double stopTol;      /* stop when the estimate of ||I - X A||_2 < stopTol */
    int mixed;           /* 0: all FP64; 1: FP32 iterations until the estimate
                            drops below switchTol, then FP64 for the rest */
    double switchTol;
    dmhm_hh_opts hh;
    int stallStop;       /* 1: also stop (or leave low precision) once the estimate
                            is below 0.1 and stays above 0.9 of the previous one or
                            above half of the one two steps back: the truncation
                            floor. Such a step ends after Z = I - X A and keeps X,
                            skipping the product Z X */
    int graph;           /* 1: once the plans of a precision are built and their
                            tile engines tuned by eager steps, run each step as a
                            replay of a captured HIP graph (same launches, same
                            seeds; one readback per step for the stop decision) */
    int bf16;            /* mixed mode: the tile GEMMs of the FP32 steps take BF16
                            operands (FP32 accumulation) while the residual
                            estimate is >= bf16Tol; QR, SVD and the estimate
                            stay in FP64 arithmetic. From X_0 = alpha A^T on an
                            ill-conditioned operator BF16 rounding can turn the
                            smallest eigenvalues of X A negative, from where the
                            iteration diverges (the 256^2 lognormal operator
                            does): use it on well-conditioned operators only */
    double bf16Tol;
    int start;           /* 0: X_0 = alpha A^T (the reference's start); 1: X_0 =
                            the recursive direct inverse of A (dmhm_hmat_direct_invert
                            with opts->hh, in A's precision), so the iteration
                            only refines the truncation error; every step then
                            runs in A's precision (mixed has no effect) and, in
                            stop mode, begins with the estimate of ||I - X A||_2
                            applied through H*x products, so a step at the
                            truncation floor ends before Z = I - X A is formed */
} dmhm_schulz_opts;

typedef struct {
    int iterations;
    int iterationsLowPrecision;
    double residualEstimate;     /* last estimate of ||I - X A||_2 */
    double normEstimate;         /* estimate of ||A||_2 used for X_0 */
    double seconds;              /* wall time of the whole inversion */
    double history[128];         /* residual estimate per iteration */
    int graphIterations;         /* iterations run as HIP graph replays */
    int iterationsBf16;          /* iterations with BF16 tile operands */
} dmhm_schulz_report;

/* A := inv(A) by Schulz iteration X_{k+1} = (2I - X_k A) X_k from
 * X_0 = alpha A^T, alpha = 2/est^2 (HMat2d::SchulzInvert,
 * src/HMat2d/Invert-incl.hpp:128-196). Replaces h's contents. The iterates
 * ping-pong between two preallocated buffers; no allocation after the first
 * two steps of each precision. */
int dmhm_hmat_schulz_invert(dmhm_hmat h, const dmhm_schulz_opts* opts, dmhm_schulz_report* report);

/* A := inv(A) by recursive block Gauss-Jordan elimination over the children
 * of every diagonal NODE block (HMat2d::DirectInvert / HMat3d::DirectInvert,
 * src/HMat2d/Invert-incl.hpp:17-125, src/HMat3d/Invert-incl.hpp; dense
 * leaves: hmat_tools::Invert, src/hmat_tools/Invert.cpp:14-42, Gauss-Jordan
 * with partial pivoting here). Every sub-matrix product is an H*H product
 * with randomized compression (opts: oversample, truncation, seed; NULL:
 * oversample 4, fixed rank maxRank). The elimination is lowered once per
 * geometry and rank into a program of merged launches (each step's
 * independent sub-matrix products in one launch per phase) and replayed.
 * Replaces h's contents (standard layout, rank <= maxRank). DMHM_ESINGULAR if
 * a pivot block has a zero pivot. */
int dmhm_hmat_direct_invert(dmhm_hmat h, const dmhm_hh_opts* opts);
/* The cached direct-inversion program of h's geometry and rank: number of
 * sub-matrix products, of merged launches sequences (plans), and the host
 * seconds it took to build (builds it if needed). No reference counterpart. */
/* Host only (no GPU needed): lowers the direct inversion of a geometry at
 * rank maxRank and checks that every task of the program stays inside its
 * buffers. stats[7]: sub-matrix products, merged plans, steps, dense leaf
 * inversions, workspace elements, plan phases (launches), identity phases.
 * DMHM_EINVAL names the first violation. No reference counterpart. */
int dmhm_check_direct_invert_program(int dim, const int* sizes, int numLevels, int maxRank, int strong,
                                     int oversample, long long* stats);
int dmhm_direct_invert_program_stats(dmhm_hmat h, int oversample, long long* products, long long* plans,
                                     double* seconds);

/* Batches of independent H-matrices (no reference counterpart: the reference
 * inverts such operators one after the other, HMat3d::SchulzInvert per
 * operator, src/HMat3d/Invert-incl.hpp:127-188). Members may differ in
 * dimension, grid, levels and rank but share a precision; each batch
 * operation runs the members of one class (same geometry, layouts and
 * operand forms) in one launch per phase, chunked by the workspace budget,
 * while members whose products fill the chip alone (a tile phase of 65,536
 * tiles or more) run one after another. A member's result is the result of
 * the same operation on that member alone (same seeds), up to the tile engine
 * each phase is tuned to. */
int dmhm_hbatch_create(int count, const dmhm_hmat* members, dmhm_hbatch* out);   /* copies the members */
int dmhm_hbatch_destroy(dmhm_hbatch b);
int dmhm_hbatch_count(dmhm_hbatch b, int* count);
/* A new handle holding a copy of member `index`. */
int dmhm_hbatch_member(dmhm_hbatch b, int index, dmhm_hmat* out);
/* C_i := alpha A_i B_i for every member (dmhm_hmat_multiply_hmat per member). */
int dmhm_hbatch_multiply_hmat(double alpha, dmhm_hbatch A, dmhm_hbatch B, const dmhm_hh_opts* opts, dmhm_hbatch* C);
/* EstimateTwoNorm of every member (estimates: count doubles). */
int dmhm_hbatch_estimate_two_norm(dmhm_hbatch b, double theta, double confidence, uint64_t seed, double* estimates);
/* SchulzInvert of every member in place (dmhm_hmat_schulz_invert per member;
 * each member stops on its own estimate; FP32 or FP64 as stored, start
 * alpha A^T, no rank model; reports: count entries). DMHM_ENOCONV if a
 * member did not reach stopTol. */
int dmhm_hbatch_schulz_invert(dmhm_hbatch b, const dmhm_schulz_opts* opts, dmhm_schulz_report* reports);

/* Packed byte image of an H-matrix (HMat2d::PackedSize / Pack / Unpack and
 * the HMat2d(const Vector<byte>&) constructor; include/dmhm/hmat2d.hpp:89,108-112,
 * src/HMat2d/Pack-incl.hpp:27-160). Own format, not byte-compatible with the
 * reference's: a header (magic "DMHMAMD1", geometry, maxRank, precision),
 * the per-block capacities and ranks, the dense leaves, then each admissible
 * block's U and V at its current rank. Unpack creates a new handle on the
 * current GPU; pack(unpack(x)) reproduces x's image byte for byte. */
int dmhm_hmat_packed_size(dmhm_hmat h, uint64_t* bytes);
int dmhm_hmat_pack(dmhm_hmat h, void* buf, uint64_t capacity, uint64_t* written);
int dmhm_hmat_unpack(const void* buf, uint64_t size, dmhm_hmat* out);
/* The same from an image of implicit length (HMat2d::Unpack(const byte*),
 * include/dmhm/hmat2d.hpp:111): the length follows from the image's header
 * and rank table; *consumed (if not NULL) receives it. */
int dmhm_hmat_unpack_image(const void* buf, uint64_t* consumed, dmhm_hmat* out);

/* Expand to a dense n x n column-major host matrix (HMat2d::Print multiplies
 * by the identity, HMat2d.cpp:233-248). For tests on small matrices. */
int dmhm_hmat_to_dense(dmhm_hmat h, double* out, int ld);

/* Read the per-block ranks (one int per admissible block, tree DFS order). */
int dmhm_hmat_ranks(dmhm_hmat h, int* ranks, int count);

/* Bytes streamed by one H*X with k right-hand sides (algorithmic: every
 * dense and low-rank factor entry at its current rank, plus X and Y). */
int dmhm_hmat_matvec_bytes(dmhm_hmat h, int k, double* bytes, double* flops);

/* Measurement support (no counterpart in the reference; used by bench.py).
 * Runs `reps` products H*X with k right-hand sides on library-owned device
 * buffers and times each launch with HIP events on the library stream:
 *   ms[0] mean time per product, ms[1] mean projection launch, ms[2] mean
 *   accumulation launch;
 *   bytes[0] algorithmic bytes of the projection launch (V factors at their
 *   ranks + X), bytes[1] of the accumulation launch (dense leaves + U factors
 *   + Y). */
int dmhm_profile_matvec(dmhm_hmat h, int k, int reps, double* ms, double* bytes);

/* Times one H*H product alpha A B (result discarded) launch by launch:
 *   ms[0] total, ms[1] all tile launches, ms[2] QR, ms[3] SVD, ms[4] RNG,
 *   ms[5] the longest tile launch; flops[0] algorithmic flops of all tile
 *   launches, flops[1] of the longest tile launch. */
int dmhm_profile_hh(double alpha, dmhm_hmat A, dmhm_hmat B, const dmhm_hh_opts* opts, double* ms,
                    double* flops);

/* Rank model (no counterpart in the reference; see dmhm_hh_opts). Created
 * with He-initialised weights from `seed`; parameters are FP32 on the device
 * (dmhm_rank_model_num_params floats, layout W1[NF x 64], b1[64], W2[64 x 64],
 * b2[64], w3[64], b3). X: n x DMHM_RANK_FEATURES row-major host array, t: n
 * target ranks. gradient: d/dparams of mean (y - t)^2 (host output) and the
 * loss; train: `steps` Adam steps (learning rate lr) over consecutive
 * mini-batches of `batch` samples of a seeded shuffle of the data, on the GPU;
 * losses (if non-null) receives each step's mini-batch loss. */
int dmhm_rank_model_create(uint64_t seed, dmhm_rank_model* out);
int dmhm_rank_model_destroy(dmhm_rank_model m);
int dmhm_rank_model_num_params(void);
int dmhm_rank_model_get_params(dmhm_rank_model m, float* params, int count);
int dmhm_rank_model_set_params(dmhm_rank_model m, const float* params, int count);
int dmhm_rank_model_gradient(dmhm_rank_model m, const float* X, const float* t, int n, float* grads,
                             float* loss);
int dmhm_rank_model_predict(dmhm_rank_model m, const float* X, int n, float* y);
int dmhm_rank_model_train(dmhm_rank_model m, const float* X, const float* t, int n, int steps, int batch,
                          float lr, uint64_t seed, float* losses);
/* Features (count = numLowRank x DMHM_RANK_FEATURES floats, host) of the
 * admissible blocks of a product whose left operand is A, in the order of
 * dmhm_hmat_ranks: the input the rank model sees in that product. */
int dmhm_hmat_rank_features(dmhm_hmat A, float* out, int count);
/* The last product with a rank model: out[0] admissible blocks, out[1] sum of
 * the predicted sketch widths, out[2] blocks recomputed at full width,
 * out[3] blocks x full width. */
int dmhm_last_product_stats(long long* out);

/* Tile engine of the grouped small GEMMs (measurement and test support, no
 * counterpart in the reference): -1 autotunes every tile phase (each phase
 * times the candidate engines on its first runs and keeps the fastest; the
 * default), 0 the LDS-staged vector-FMA kernel, 1 the MFMA kernel, 2 the
 * MFMA kernel with one-chunk prefetch, 3 / 4 the same with BF16 operands
 * where BF16 is allowed (the FP32 steps of a mixed Schulz inversion).
 * DMHM_TILE_ENGINE=auto|lds|mfma|mfma_pf|bf16|bf16_pf sets it at dmhm_init. */
int dmhm_set_tile_engine(int engine);
int dmhm_get_tile_engine(void);

/* Per-phase timings of the last dmhm_profile_hh call (measurement support,
 * no counterpart in the reference): phase kind (0 tile, 1 QR, 2 SVD, 3 RNG),
 * milliseconds and algorithmic flops, in plan order; *count = number of
 * phases (at most `cap` entries written). */
int dmhm_profile_hh_phases(int cap, double* ms, double* flops, int* kinds, int* count);

#ifdef __cplusplus
}
#endif
#endif /* DMHM_AMD_H */

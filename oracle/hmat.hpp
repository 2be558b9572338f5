// ORACLE — test infrastructure only. Nothing in the product links, loads or
// calls this code; it is the FP64 CPU checker and the CPU baseline.
//
// Restatement of the reference's sequential H-matrix classes HMat2d / HMat3d
// (include/dmhm/hmat2d.hpp:22-393, include/dmhm/hmat3d.hpp, src/HMat2d/*,
// src/HMat3d/*) and the hmat_tools routines they call
// (include/dmhm/hmat_tools.hpp:1504-1746, src/hmat_tools/*.cpp), written for
// real double precision and one class for both dimensions (dim = 2: quadtree,
// 16 children per node; dim = 3: octree, 64 children).
//
// Parity status (each is pinned by tests/test_oracle_*.py):
//  tree/map/admissibility ... block counts vs the probe of the reference
//                              recursion (tests/golden/tree_counts.json),
//                              permutation checks
//  ImportSparse, H*X, H^T*X .. dense brute force from the CSR / expansion
//  QR / SVD substitutes ...... numpy.linalg (LAPACK)
//  RoundedUpdate, Compress ... Eckart-Young optimality (numpy SVD)
//  randomized H*H sketch ..... exact when the product rank <= r; HMT bound
//  H*H product ............... dense product of the expanded operands
//  EstimateTwoNorm, Schulz ... dense 2-norms, quadratic convergence
//  DirectInvert .............. the dense inverse (numpy.linalg.inv) when no
//                              block is truncated; the residual otherwise
//  dense invert .............. numpy.linalg.inv
//  LCG / Box-Muller .......... Python big-integer closed form
// Parity unpinned (no reference build: MPI is absent; the reference holds no
// golden vectors for these): bitwise identity with the reference of the
// sketch outputs (Gaussian draw order, pivot order, SVD signs) and of the
// Schulz iterate sequence. Only the properties above are checked for them.
#pragma once
#include "la.hpp"
#include "rng.hpp"
#include <memory>

namespace orc {

// Per-call options that the reference keeps in globals (global.cpp:30-34).
struct Options {
    int oversample = 4;          // Oversample() default (global.cpp:30)
    int tolMode = 0;             // 0: fixed rank (reference); 1: SVDTrunc rule
    double relTol = 1e-10;       // relative tolerance for tolMode = 1
    bool refQuirk = true;        // reproduce the random.hpp:441,443 row quirk
};
extern Options g_opts;
extern Streams g_rng;

// Rank kept out of singular values s[0..k) under the current options:
// fixed mode keeps min(k, cap); tolerance mode follows DistHMat2d::SVDTrunc
// (src/DistHMat2d/Truncation-incl.hpp:37-63): keep s_i > relTol*s_0 among the
// first min(k,cap), or one value if s_0 < relTol^2.
int truncated_rank(const double* s, int k, int cap);

class HMat {
public:
    enum Type { NODE, LOW_RANK, DENSE };
    int dim = 2, numLevels = 0, maxRank = 0;
    bool strong = false;
    int sSize[3] = {1, 1, 1}, tSize[3] = {1, 1, 1};
    int sBox[3] = {0, 0, 0}, tBox[3] = {0, 0, 0};
    int sOff = 0, tOff = 0;
    Type type = NODE;
    std::vector<std::unique_ptr<HMat>> ch;   // child(t,s) = ch[s + nc*t]
    LowRank F;
    Dense D;

    int nc() const { return 1 << dim; }
    int height() const { return tSize[0] * tSize[1] * (dim == 3 ? tSize[2] : 1); }
    int width() const { return sSize[0] * sSize[1] * (dim == 3 ? sSize[2] : 1); }
    HMat& child(int t, int s) { return *ch[s + nc() * t]; }
    const HMat& child(int t, int s) const { return *ch[s + nc() * t]; }

    // Top-level square constructors (hmat2d.hpp:44-62).
    static std::unique_ptr<HMat> top(int dim, int numLevels, int maxRank, bool strong,
                                     const int* sizes);
    // Geometry of child c of a box with the given per-axis sizes.
    static void child_sizes(int dim, const int* size, int c, int* out);
    static int child_count(int dim, const int* size, int c);

    bool admissible() const;   // HMat2d.cpp:790-809, HMat3d.cpp:851-875
    void copy_geometry_child(const HMat& parent, int t, int s, int tOffRel, int sOffRel);

    void import_sparse(const int* rowptr, const int* colind, const double* vals);
    void import_low_rank(const LowRank& F);
    void update_with_low_rank(double alpha, const LowRank& F);
    void set_to_random();

    // Y := alpha op(H) X + beta Y, op = N or T (MultiplyDense-incl.hpp:12-160)
    void multiply_dense(bool trans, double alpha, const double* X, int ldx, int k,
                        double beta, double* Y, int ldy) const;

    void copy_from(const HMat& B);
    void transpose_from(const HMat& B);
    void scale(double alpha);
    void set_to_identity();
    void add_constant_to_diagonal(double alpha);

    // C := alpha A B and C := alpha A B + beta C (MultiplyHMat-incl.hpp)
    static void multiply(double alpha, const HMat& A, const HMat& B, HMat& C);
    static void multiply_update(double alpha, const HMat& A, const HMat& B,
                                double beta, HMat& C);
    // Only the children C(t, 0..nc-1) of C := alpha A B, for two NODE operands
    // with a NODE product (the rest of C stays empty): a bounded sample of the
    // product's work, used to time the CPU baseline.
    static void multiply_child_row(double alpha, const HMat& A, const HMat& B, HMat& C, int t);
    // Only the child C(t, s) of C := alpha A B (same conditions; timing sample).
    static void multiply_child(double alpha, const HMat& A, const HMat& B, HMat& C, int t, int s);

    // A := inv(A) by recursive block Gauss-Jordan elimination over the
    // children (DirectInvert, Invert-incl.hpp:17-125).
    void direct_invert();
    // Schulz inversion of *this (Invert-incl.hpp:128-196). history receives the
    // per-iteration residual estimates; returns the number of iterations run.
    int schulz_invert(int numIterations, double theta, double confidence,
                      double stopTol, int maxIter, std::vector<double>* history);

    void to_dense(double* out, int ld) const;   // expand into a dense matrix
    void count_blocks(std::vector<long>& denseByLevel, std::vector<long>& lowRankByLevel,
                      int level) const;
    void rank_stats(long& sumRank, int& maxR, long& storage) const;
};

// hmat_tools pieces used by the products (src/hmat_tools/*.cpp).
void rounded_update(int maxRank, double alpha, const LowRank& A, double beta, LowRank& B);
void compress_dense(int maxRank, Dense& D, LowRank& F);
void randomized_multiply(int sampleRank, double alpha, const HMat& A, const HMat& B,
                         LowRank& F);
double estimate_two_norm(const HMat& A, double theta, double confidence);

} // namespace orc

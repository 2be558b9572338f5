// ORACLE — test infrastructure only. C entry points used by tests/ (ctypes),
// __graft_entry__.smoke() and bench.py's cpu_baseline leg. Never linked into
// the product.
#include "hmat.hpp"
#include <cstdint>

using namespace orc;

namespace {
// BuildMapOnQuadrant (HMat2d.cpp:27-67, HMat3d.cpp:28-90): natural index
// x + nx*y (+ nx*ny*z) -> hierarchical index, recursive lower-left,
// lower-right, upper-left, upper-right (then the back octants in 3D).
void map_rec(int* map, int& index, int level, int numLevels, int dim, const int* full,
             const int* box)
{
    if (level == numLevels - 1) {
        for (int k = 0; k < (dim == 3 ? box[2] : 1); ++k)
            for (int j = 0; j < box[1]; ++j) {
                int* row = map + j * full[0] + k * full[0] * full[1];
                for (int i = 0; i < box[0]; ++i) row[i] = index++;
            }
        return;
    }
    const int nc = 1 << dim;
    for (int c = 0; c < nc; ++c) {
        int sub[3] = {1, 1, 1}, off[3] = {0, 0, 0};
        for (int a = 0; a < dim; ++a) {
            const int half = box[a] / 2;
            const bool hi = (c >> a) & 1;
            sub[a] = hi ? box[a] - half : half;
            off[a] = hi ? half : 0;
        }
        int* base = map + off[0] + off[1] * full[0] + (dim == 3 ? off[2] * full[0] * full[1] : 0);
        map_rec(base, index, level + 1, numLevels, dim, full, sub);
    }
}
} // namespace

extern "C" {

void orc_set_options(int oversample, int tolMode, double relTol, int refQuirk)
{
    g_opts.oversample = oversample; g_opts.tolMode = tolMode;
    g_opts.relTol = relTol; g_opts.refQuirk = refQuirk != 0;
}

void orc_rng_initialize() { g_rng.initialize(); }
void orc_rng_set_state(uint64_t serial, uint64_t parallel)
{ g_rng.serial.seed(serial); g_rng.parallel.seed(parallel); }
void orc_rng_serial_raw(int n, uint64_t* out) { for (int i = 0; i < n; ++i) out[i] = g_rng.serial.next(); }
void orc_rng_serial_uniform(int n, double* out) { for (int i = 0; i < n; ++i) out[i] = g_rng.serial.uniform(); }
void orc_rng_serial_gaussian_vector(int n, double* x) { g_rng.serial.gaussian_vector(n, x); }
void orc_rng_serial_gaussian_vectors(int m, int n, double* A)
{ g_rng.serial.gaussian_vectors(m, n, A, g_opts.refQuirk); }

void orc_gemm(char ta, char tb, int m, int n, int k, double alpha, const double* A, int lda,
              const double* B, int ldb, double beta, double* C, int ldc)
{ gemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc); }
void orc_qr(int m, int n, double* A, double* tau) { qr(m, n, A, m, tau); }
void orc_pivoted_qr(int m, int n, double* A, int* jpvt, double* tau) { pivoted_qr(m, n, A, m, jpvt, tau); }
void orc_form_q(int m, int n, int k, double* A, const double* tau) { form_q(m, n, k, A, m, tau); }
void orc_apply_q(int m, int n, int k, const double* A, const double* tau, double* C)
{ apply_q(m, n, k, A, m, tau, C, m); }
void orc_svd(int m, int n, const double* A, double* U, double* s, double* VT) { svd(m, n, A, m, U, s, VT); }

void orc_build_map(int dim, const int* sizes, int numLevels, int* map)
{
    int full[3] = {sizes[0], sizes[1], dim == 3 ? sizes[2] : 1};
    int index = 0;
    map_rec(map, index, 0, numLevels, dim, full, full);
}

void* orc_hmat_create(int dim, const int* sizes, int numLevels, int maxRank, int strong,
                      const int* rowptr, const int* colind, const double* vals)
{
    auto h = HMat::top(dim, numLevels, maxRank, strong != 0, sizes);
    h->import_sparse(rowptr, colind, vals);
    return h.release();
}
void orc_hmat_free(void* h) { delete static_cast<HMat*>(h); }
// Child C(t, s) of C := alpha A B (timing sample of a product).
void* orc_hmat_multiply_child(double alpha, void* A, void* B, int t, int s)
{
    auto C = std::make_unique<HMat>();
    C->type = HMat::LOW_RANK;
    HMat::multiply_child(alpha, *static_cast<HMat*>(A), *static_cast<HMat*>(B), *C, t, s);
    return C.release();
}
// Children C(t, *) of C := alpha A B (timing sample of a product).
void* orc_hmat_multiply_child_row(double alpha, void* A, void* B, int t)
{
    auto C = std::make_unique<HMat>();
    HMat::multiply_child_row(alpha, *static_cast<HMat*>(A), *static_cast<HMat*>(B), *C, t);
    return C.release();
}
int orc_hmat_height(void* h) { return static_cast<HMat*>(h)->height(); }
void orc_hmat_set_to_random(void* h) { static_cast<HMat*>(h)->set_to_random(); }
void orc_hmat_multiply_dense(void* h, int trans, double alpha, const double* X, int ldx, int k,
                             double beta, double* Y, int ldy)
{ static_cast<HMat*>(h)->multiply_dense(trans != 0, alpha, X, ldx, k, beta, Y, ldy); }
void orc_hmat_to_dense(void* h, double* out)
{ auto* H = static_cast<HMat*>(h); H->to_dense(out, H->height()); }
void orc_hmat_count_blocks(void* h, long* dense, long* lowrank, int maxLevels)
{
    std::vector<long> d, f;
    static_cast<HMat*>(h)->count_blocks(d, f, 0);
    for (int l = 0; l < maxLevels; ++l) {
        dense[l] = l < (int)d.size() ? d[l] : 0;
        lowrank[l] = l < (int)f.size() ? f[l] : 0;
    }
}
void orc_hmat_rank_stats(void* h, long* out)
{
    long sr = 0, st = 0; int mr = 0;
    static_cast<HMat*>(h)->rank_stats(sr, mr, st);
    out[0] = sr; out[1] = mr; out[2] = st;
}
void* orc_hmat_multiply(double alpha, void* A, void* B)
{
    auto* C = new HMat();
    HMat::multiply(alpha, *static_cast<HMat*>(A), *static_cast<HMat*>(B), *C);
    return C;
}
void orc_hmat_multiply_update(double alpha, void* A, void* B, double beta, void* C)
{ HMat::multiply_update(alpha, *static_cast<HMat*>(A), *static_cast<HMat*>(B), beta, *static_cast<HMat*>(C)); }
void orc_hmat_direct_invert(void* h) { static_cast<HMat*>(h)->direct_invert(); }
int orc_invert(int n, double* A)
{
    try {
        invert(n, A, n);
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}
void* orc_hmat_copy(void* h) { auto* C = new HMat(); C->copy_from(*static_cast<HMat*>(h)); return C; }
void* orc_hmat_transpose(void* h) { auto* C = new HMat(); C->transpose_from(*static_cast<HMat*>(h)); return C; }
void orc_hmat_scale(void* h, double a) { static_cast<HMat*>(h)->scale(a); }
void orc_hmat_add_diag(void* h, double a) { static_cast<HMat*>(h)->add_constant_to_diagonal(a); }
void orc_hmat_set_identity(void* h) { static_cast<HMat*>(h)->set_to_identity(); }
double orc_estimate_two_norm(void* h, double theta, double conf)
{ return estimate_two_norm(*static_cast<HMat*>(h), theta, conf); }
int orc_hmat_schulz(void* h, int numIter, double theta, double conf, double stopTol, int maxIter,
                    double* history, int histCap)
{
    std::vector<double> hist;
    const int it = static_cast<HMat*>(h)->schulz_invert(numIter, theta, conf, stopTol, maxIter, &hist);
    for (int i = 0; i < histCap; ++i) history[i] = i < (int)hist.size() ? hist[i] : -1.0;
    return it;
}
void orc_rounded_update(int maxRank, double alpha, int m, int n, int rA, const double* AU,
                        const double* AV, double beta, int rB, const double* BU, const double* BV,
                        double* outU, double* outV, int* outR)
{
    LowRank A, B;
    A.U.resize(m, rA); A.V.resize(n, rA); B.U.resize(m, rB); B.V.resize(n, rB);
    std::copy(AU, AU + (size_t)m * rA, A.U.a.begin()); std::copy(AV, AV + (size_t)n * rA, A.V.a.begin());
    std::copy(BU, BU + (size_t)m * rB, B.U.a.begin()); std::copy(BV, BV + (size_t)n * rB, B.V.a.begin());
    rounded_update(maxRank, alpha, A, beta, B);
    *outR = B.rank();
    std::copy(B.U.a.begin(), B.U.a.end(), outU);
    std::copy(B.V.a.begin(), B.V.a.end(), outV);
}
void orc_compress_dense(int maxRank, int m, int n, const double* Din, double* U, double* V, int* r)
{
    Dense D(m, n);
    std::copy(Din, Din + (size_t)m * n, D.a.begin());
    LowRank F;
    compress_dense(maxRank, D, F);
    *r = F.rank();
    std::copy(F.U.a.begin(), F.U.a.end(), U);
    std::copy(F.V.a.begin(), F.V.a.end(), V);
}
void orc_randomized_multiply(int sampleRank, double alpha, void* A, void* B, int capU, double* U,
                             double* V, int* r)
{
    LowRank F;
    randomized_multiply(sampleRank, alpha, *static_cast<HMat*>(A), *static_cast<HMat*>(B), F);
    *r = std::min(F.rank(), capU);
    std::copy(F.U.a.begin(), F.U.a.begin() + (size_t)F.U.m * *r, U);
    std::copy(F.V.a.begin(), F.V.a.begin() + (size_t)F.V.m * *r, V);
}

} // extern "C"

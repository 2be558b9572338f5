// ORACLE — test infrastructure only (see hmat.hpp header).
#include "hmat.hpp"
#include <cstdio>

namespace orc {

Options g_opts;
Streams g_rng;

int truncated_rank(const double* s, int k, int cap)
{
    const int lim = std::min(k, cap);
    if (g_opts.tolMode == 0) return std::max(lim, 0);
    if (k == 0) return 0;
    const double relTol = g_opts.relTol;
    if (s[0] < relTol * relTol) return std::min(1, lim);
    const double tol = relTol * s[0];
    int r = 0;
    while (r < lim && s[r] > tol) ++r;
    return r;
}

// --------------------------------------------------------------------------
// Small dense helpers
// --------------------------------------------------------------------------
static void scale_dense(double alpha, Dense& D)
{   // hmat_tools::Scale (Dense), hmat_tools.hpp:1606-1636
    if (alpha == 1.0) return;
    for (auto& v : D.a) v = (alpha == 0.0) ? 0.0 : v * alpha;
}

static void scale_lowrank(double alpha, LowRank& F)
{   // hmat_tools::Scale (LowRank), hmat_tools.hpp:1638-1650
    if (alpha == 0.0) { F.U.resize(F.U.m, 0); F.V.resize(F.V.m, 0); }
    else scale_dense(alpha, F.U);
}

static LowRank slice(const LowRank& F, int r0, int nr, int c0, int ncols)
{
    LowRank S;
    const int r = F.rank();
    S.U.resize(nr, r); S.V.resize(ncols, r);
    for (int j = 0; j < r; ++j) {
        for (int i = 0; i < nr; ++i) S.U(i, j) = F.U(r0 + i, j);
        for (int i = 0; i < ncols; ++i) S.V(i, j) = F.V(c0 + i, j);
    }
    return S;
}

// Dense D := alpha U V^T + beta D
static void lowrank_to_dense(double alpha, const LowRank& F, double beta, Dense& D)
{
    gemm('N', 'T', D.m, D.n, F.rank(), alpha, F.U.a.data(), F.U.m,
         F.V.a.data(), std::max(F.V.m, 1), beta, D.a.data(), std::max(D.m, 1));
}

// F := alpha A B for low-rank A, B (MultiplyMatrix.cpp:243-311)
static void mult_ff(double alpha, const LowRank& A, const LowRank& B, LowRank& C)
{
    const int m = A.height(), n = B.width(), Ar = A.rank(), Br = B.rank();
    if (Ar <= Br) {
        C.U = A.U;
        Dense W(Br, Ar);
        gemm('T', 'N', Br, Ar, B.height(), 1.0, B.U.a.data(), std::max(B.U.m, 1),
             A.V.a.data(), std::max(A.V.m, 1), 0.0, W.a.data(), std::max(Br, 1));
        C.V.resize(n, Ar);
        gemm('N', 'N', n, Ar, Br, alpha, B.V.a.data(), std::max(B.V.m, 1),
             W.a.data(), std::max(Br, 1), 0.0, C.V.a.data(), std::max(n, 1));
    } else {
        Dense W(Ar, Br);
        gemm('T', 'N', Ar, Br, A.width(), 1.0, A.V.a.data(), std::max(A.V.m, 1),
             B.U.a.data(), std::max(B.U.m, 1), 0.0, W.a.data(), std::max(Ar, 1));
        C.U.resize(m, Br);
        gemm('N', 'N', m, Br, Ar, alpha, A.U.a.data(), std::max(A.U.m, 1),
             W.a.data(), std::max(Ar, 1), 0.0, C.U.a.data(), std::max(m, 1));
        C.V = B.V;
    }
}

// F := alpha D F (MultiplyMatrix.cpp:313-352)
static void mult_df(double alpha, const Dense& A, const LowRank& B, LowRank& C)
{
    C.U.resize(A.m, B.rank());
    gemm('N', 'N', A.m, B.rank(), A.n, alpha, A.a.data(), std::max(A.m, 1),
         B.U.a.data(), std::max(B.U.m, 1), 0.0, C.U.a.data(), std::max(A.m, 1));
    C.V = B.V;
}

// F := alpha F D (MultiplyMatrix.cpp:354-403)
static void mult_fd(double alpha, const LowRank& A, const Dense& B, LowRank& C)
{
    C.U = A.U;
    C.V.resize(B.n, A.rank());
    gemm('T', 'N', B.n, A.rank(), A.width(), alpha, B.a.data(), std::max(B.m, 1),
         A.V.a.data(), std::max(A.V.m, 1), 0.0, C.V.a.data(), std::max(B.n, 1));
}

// SVD-based truncation of a dense matrix held in U (MultiplyMatrix.cpp:405-449
// and Compress.cpp:14-53): F.U = left vectors, F.V = V Sigma.
void compress_dense(int maxRank, Dense& D, LowRank& F)
{
    const int m = D.m, n = D.n, k = std::min(m, n);
    Dense U(m, k), VT(k, n);
    std::vector<double> s(k);
    if (k > 0) svd(m, n, D.a.data(), m, U.a.data(), s.data(), VT.a.data());
    const int r = truncated_rank(s.data(), k, std::min(k, maxRank));
    F.U.resize(m, r); F.V.resize(n, r);
    for (int j = 0; j < r; ++j) {
        for (int i = 0; i < m; ++i) F.U(i, j) = U(i, j);
        for (int i = 0; i < n; ++i) F.V(i, j) = s[j] * VT(j, i);
    }
}

// F := alpha D D truncated to maxRank (MultiplyMatrix.cpp:405-449)
static void mult_dd_trunc(int maxRank, double alpha, const Dense& A, const Dense& B,
                          LowRank& C)
{
    Dense P(A.m, B.n);
    gemm('N', 'N', A.m, B.n, A.n, alpha, A.a.data(), std::max(A.m, 1), B.a.data(),
         std::max(B.m, 1), 0.0, P.a.data(), std::max(A.m, 1));
    compress_dense(maxRank, P, C);
}

// F := alpha D D + beta F, then Compress (MultiplyMatrix.cpp:502-524)
static void mult_dd_update(int maxRank, double alpha, const Dense& A, const Dense& B,
                           double beta, LowRank& C)
{
    Dense P(A.m, B.n);
    gemm('N', 'N', A.m, B.n, A.n, alpha, A.a.data(), std::max(A.m, 1), B.a.data(),
         std::max(B.m, 1), 0.0, P.a.data(), std::max(A.m, 1));
    lowrank_to_dense(beta, C, 1.0, P);    // Update(beta, C, 1, D), Update.cpp
    compress_dense(maxRank, P, C);
}

// B :~= alpha A + beta B (RoundedUpdate.cpp:14-162)
void rounded_update(int maxRank, double alpha, const LowRank& A, double beta, LowRank& B)
{
    const int m = A.height(), n = A.width(), minDim = std::min(m, n);
    const int Ar = A.rank(), Br = B.rank(), r = Ar + Br;
    const int capRank = std::min(std::min(r, minDim), maxRank);
    if (capRank == r) {
        // Early exit: concatenate [beta B.U, alpha A.U], [B.V, A.V].
        LowRank out;
        out.U.resize(m, r); out.V.resize(n, r);
        for (int j = 0; j < Br; ++j) {
            for (int i = 0; i < m; ++i) out.U(i, j) = (beta == 1.0) ? B.U(i, j) : beta * B.U(i, j);
            for (int i = 0; i < n; ++i) out.V(i, j) = B.V(i, j);
        }
        for (int j = 0; j < Ar; ++j) {
            for (int i = 0; i < m; ++i) out.U(i, Br + j) = alpha * A.U(i, j);
            for (int i = 0; i < n; ++i) out.V(i, Br + j) = A.V(i, j);
        }
        B = std::move(out);
        return;
    }
    Dense U(m, r), V(n, r);
    for (int j = 0; j < Ar; ++j) {
        for (int i = 0; i < m; ++i) U(i, j) = alpha * A.U(i, j);
        for (int i = 0; i < n; ++i) V(i, j) = A.V(i, j);
    }
    for (int j = 0; j < Br; ++j) {
        for (int i = 0; i < m; ++i) U(i, Ar + j) = beta * B.U(i, j);
        for (int i = 0; i < n; ++i) V(i, Ar + j) = B.V(i, j);
    }
    const int minDimU = std::min(m, r), minDimV = std::min(n, r);
    std::vector<double> tauU(minDimU), tauV(minDimV);
    qr(m, r, U.a.data(), m, tauU.data());
    qr(n, r, V.a.data(), n, tauV.data());
    // W := R1 R2^T with R1 (minDimU x r), R2 (minDimV x r) upper trapezoidal.
    Dense R1(minDimU, r), R2(minDimV, r);
    for (int j = 0; j < r; ++j) {
        for (int i = 0; i <= std::min(j, minDimU - 1); ++i) R1(i, j) = U(i, j);
        for (int i = 0; i <= std::min(j, minDimV - 1); ++i) R2(i, j) = V(i, j);
    }
    Dense W(minDimU, minDimV);
    gemm('N', 'T', minDimU, minDimV, r, 1.0, R1.a.data(), minDimU, R2.a.data(),
         minDimV, 0.0, W.a.data(), minDimU);
    const int k = std::min(minDimU, minDimV);
    Dense WU(minDimU, k), VT(k, minDimV);
    std::vector<double> s(k);
    svd(minDimU, minDimV, W.a.data(), minDimU, WU.a.data(), s.data(), VT.a.data());
    const int rr = truncated_rank(s.data(), k, capRank);
    LowRank out;
    out.U.resize(m, rr); out.V.resize(n, rr);
    for (int j = 0; j < rr; ++j) {
        for (int i = 0; i < minDimU; ++i) out.U(i, j) = s[j] * WU(i, j);
        for (int i = 0; i < minDimV; ++i) out.V(i, j) = VT(j, i);
    }
    apply_q(m, rr, minDimU, U.a.data(), m, tauU.data(), out.U.a.data(), m);
    apply_q(n, rr, minDimV, V.a.data(), n, tauV.data(), out.V.a.data(), n);
    B = std::move(out);
}

// --------------------------------------------------------------------------
// Tree geometry
// --------------------------------------------------------------------------
void HMat::child_sizes(int dim, const int* size, int c, int* out)
{
    for (int a = 0; a < 3; ++a) {
        if (a >= dim) { out[a] = 1; continue; }
        const int half = size[a] / 2;
        out[a] = ((c >> a) & 1) ? size[a] - half : half;
    }
}

int HMat::child_count(int dim, const int* size, int c)
{
    int s[3];
    child_sizes(dim, size, c, s);
    return s[0] * s[1] * s[2];
}

std::unique_ptr<HMat> HMat::top(int dim, int numLevels, int maxRank, bool strong,
                                const int* sizes)
{
    auto h = std::make_unique<HMat>();
    h->dim = dim; h->numLevels = numLevels; h->maxRank = maxRank; h->strong = strong;
    for (int a = 0; a < 3; ++a) {
        h->sSize[a] = h->tSize[a] = (a < dim) ? sizes[a] : 1;
        h->sBox[a] = h->tBox[a] = 0;
    }
    return h;
}

bool HMat::admissible() const
{
    if (strong) {
        int d = 0;
        for (int a = 0; a < dim; ++a) d = std::max(d, std::abs(sBox[a] - tBox[a]));
        return d > 1;
    }
    for (int a = 0; a < dim; ++a) if (sBox[a] != tBox[a]) return true;
    return false;
}

void HMat::copy_geometry_child(const HMat& p, int t, int s, int tOffRel, int sOffRel)
{
    dim = p.dim; numLevels = p.numLevels - 1; maxRank = p.maxRank; strong = p.strong;
    child_sizes(dim, p.tSize, t, tSize);
    child_sizes(dim, p.sSize, s, sSize);
    for (int a = 0; a < 3; ++a) {
        tBox[a] = (a < dim) ? 2 * p.tBox[a] + ((t >> a) & 1) : 0;
        sBox[a] = (a < dim) ? 2 * p.sBox[a] + ((s >> a) & 1) : 0;
    }
    tOff = p.tOff + tOffRel;
    sOff = p.sOff + sOffRel;
}

// Iterate the (t,s) children with their relative offsets.
template <class Fn>
static void for_children(const HMat& h, Fn fn)
{
    const int nc = h.nc();
    for (int t = 0, to = 0; t < nc; to += HMat::child_count(h.dim, h.tSize, t), ++t)
        for (int s = 0, so = 0; s < nc; so += HMat::child_count(h.dim, h.sSize, s), ++s)
            fn(t, s, to, so);
}

static void make_children(HMat& h)
{
    const int nc = h.nc();
    h.ch.clear();
    h.ch.resize(nc * nc);
    for_children(h, [&](int t, int s, int to, int so) {
        h.ch[s + nc * t] = std::make_unique<HMat>();
        h.ch[s + nc * t]->copy_geometry_child(h, t, s, to, so);
    });
}

// ImportSparse (HMat2d.cpp:961-1033) with ConvertSubmatrix
// (ConvertSubmatrix.cpp:14-119). CSR holds global hierarchical indices.
void HMat::import_sparse(const int* rowptr, const int* colind, const double* vals)
{
    ch.clear();
    const int m = height(), n = width();
    if (admissible()) {
        type = LOW_RANK;
        int r = 0;
        for (int i = 0; i < m; ++i)
            for (int k = rowptr[tOff + i]; k < rowptr[tOff + i + 1]; ++k)
                if (colind[k] >= sOff && colind[k] < sOff + n) ++r;
        F.U.resize(m, r); F.V.resize(n, r);
        r = 0;
        for (int i = 0; i < m; ++i)
            for (int k = rowptr[tOff + i]; k < rowptr[tOff + i + 1]; ++k)
                if (colind[k] >= sOff && colind[k] < sOff + n) {
                    F.U(i, r) = 1.0; F.V(colind[k] - sOff, r) = vals[k]; ++r;
                }
    } else if (numLevels > 1) {
        type = NODE;
        make_children(*this);
        for (auto& c : ch) c->import_sparse(rowptr, colind, vals);
    } else {
        type = DENSE;
        D.resize(m, n);
        for (int i = 0; i < m; ++i)
            for (int k = rowptr[tOff + i]; k < rowptr[tOff + i + 1]; ++k)
                if (colind[k] >= sOff && colind[k] < sOff + n)
                    D(i, colind[k] - sOff) = vals[k];
    }
}

// ImportLowRank (HMat2d.cpp:811-904)
void HMat::import_low_rank(const LowRank& W)
{
    ch.clear();
    if (admissible()) {
        type = LOW_RANK; F = W;
    } else if (numLevels > 1) {
        type = NODE;
        make_children(*this);
        for_children(*this, [&](int t, int s, int to, int so) {
            HMat& c = child(t, s);
            c.import_low_rank(slice(W, to, c.height(), so, c.width()));
        });
    } else {
        type = DENSE;
        D.resize(height(), width());
        lowrank_to_dense(1.0, W, 0.0, D);
    }
}

// UpdateWithLowRank (HMat2d.cpp:906-959)
void HMat::update_with_low_rank(double alpha, const LowRank& W)
{
    if (admissible()) rounded_update(maxRank, alpha, W, 1.0, F);
    else if (numLevels > 1) {
        for_children(*this, [&](int t, int s, int to, int so) {
            HMat& c = child(t, s);
            c.update_with_low_rank(alpha, slice(W, to, c.height(), so, c.width()));
        });
    } else lowrank_to_dense(alpha, W, 1.0, D);
}

// SetToRandom (SetToRandom-incl.hpp:12-53) on the "parallel" stream.
void HMat::set_to_random()
{
    switch (type) {
    case NODE:
        for (auto& c : ch) c->set_to_random();
        break;
    case DENSE:
        g_rng.parallel.gaussian_vectors(D.m, D.n, D.a.data(), g_opts.refQuirk);
        break;
    case LOW_RANK: {
        // Resize keeps the old flat contents (dense.hpp:256-282), then fill.
        auto grow = [](Dense& M, int cols) {
            std::vector<double> nb((size_t)M.m * cols, 0.0);
            std::copy(M.a.begin(), M.a.begin() + std::min(M.a.size(), nb.size()), nb.begin());
            M.a.swap(nb); M.n = cols;
        };
        grow(F.U, maxRank); grow(F.V, maxRank);
        g_rng.parallel.gaussian_vectors(F.U.m, F.U.n, F.U.a.data(), g_opts.refQuirk);
        g_rng.parallel.gaussian_vectors(F.V.m, F.V.n, F.V.a.data(), g_opts.refQuirk);
        break;
    }
    }
}

// Y := alpha op(H) X + beta Y (MultiplyDense-incl.hpp:12-160,
// MultiplyMatrix.cpp:142-197, TransposeMultiplyMatrix.cpp)
void HMat::multiply_dense(bool trans, double alpha, const double* X, int ldx, int k,
                          double beta, double* Y, int ldy) const
{
    const int m = trans ? width() : height();
    if (beta != 1.0)
        for (int j = 0; j < k; ++j)
            for (int i = 0; i < m; ++i)
                Y[i + (size_t)j * ldy] = (beta == 0.0) ? 0.0 : beta * Y[i + (size_t)j * ldy];
    switch (type) {
    case NODE:
        for_children(*this, [&](int t, int s, int to, int so) {
            if (!trans) child(t, s).multiply_dense(false, alpha, X + so, ldx, k, 1.0, Y + to, ldy);
            else child(t, s).multiply_dense(true, alpha, X + to, ldx, k, 1.0, Y + so, ldy);
        });
        break;
    case LOW_RANK: {
        const Dense& L = trans ? F.V : F.U;   // output side
        const Dense& R = trans ? F.U : F.V;   // input side
        const int r = F.rank();
        if (r == 0) break;
        Dense W(r, k);
        gemm('T', 'N', r, k, R.m, 1.0, R.a.data(), std::max(R.m, 1), X, ldx, 0.0,
             W.a.data(), r);
        gemm('N', 'N', L.m, k, r, alpha, L.a.data(), std::max(L.m, 1), W.a.data(), r,
             1.0, Y, ldy);
        break;
    }
    case DENSE:
        gemm(trans ? 'T' : 'N', 'N', m, k, trans ? D.m : D.n, alpha, D.a.data(),
             std::max(D.m, 1), X, ldx, 1.0, Y, ldy);
        break;
    }
}

void HMat::copy_from(const HMat& B)
{
    dim = B.dim; numLevels = B.numLevels; maxRank = B.maxRank; strong = B.strong;
    for (int a = 0; a < 3; ++a) {
        sSize[a] = B.sSize[a]; tSize[a] = B.tSize[a]; sBox[a] = B.sBox[a]; tBox[a] = B.tBox[a];
    }
    sOff = B.sOff; tOff = B.tOff; type = B.type;
    ch.clear(); F = LowRank(); D = Dense();
    if (type == NODE) {
        ch.resize(B.ch.size());
        for (size_t i = 0; i < B.ch.size(); ++i) {
            ch[i] = std::make_unique<HMat>();
            ch[i]->copy_from(*B.ch[i]);
        }
    } else if (type == LOW_RANK) F = B.F;
    else D = B.D;
}

// TransposeFrom / AdjointFrom for real scalars (HMat2d.cpp:467-647)
void HMat::transpose_from(const HMat& B)
{
    dim = B.dim; numLevels = B.numLevels; maxRank = B.maxRank; strong = B.strong;
    for (int a = 0; a < 3; ++a) {
        sSize[a] = B.tSize[a]; tSize[a] = B.sSize[a]; sBox[a] = B.tBox[a]; tBox[a] = B.sBox[a];
    }
    sOff = B.tOff; tOff = B.sOff; type = B.type;
    ch.clear(); F = LowRank(); D = Dense();
    if (type == NODE) {
        const int nc = B.nc();
        ch.resize(nc * nc);
        for (int t = 0; t < nc; ++t)
            for (int s = 0; s < nc; ++s) {
                ch[s + nc * t] = std::make_unique<HMat>();
                ch[s + nc * t]->transpose_from(B.child(s, t));
            }
    } else if (type == LOW_RANK) { F.U = B.F.V; F.V = B.F.U; }
    else {
        D.resize(B.D.n, B.D.m);
        for (int j = 0; j < B.D.n; ++j)
            for (int i = 0; i < B.D.m; ++i) D(j, i) = B.D(i, j);
    }
}

void HMat::scale(double alpha)
{
    switch (type) {
    case NODE: for (auto& c : ch) c->scale(alpha); break;
    case LOW_RANK: scale_lowrank(alpha, F); break;
    case DENSE: scale_dense(alpha, D); break;
    }
}

void HMat::set_to_identity()
{   // HMat2d.cpp:649-702
    if (type == NODE) {
        const int nc = this->nc();
        for (int t = 0; t < nc; ++t)
            for (int s = 0; s < nc; ++s) {
                if (s == t) child(t, s).set_to_identity();
                else child(t, s).scale(0.0);
            }
    } else if (type == DENSE) {
        scale_dense(0.0, D);
        for (int j = 0; j < std::min(D.m, D.n); ++j) D(j, j) = 1.0;
    }
}

void HMat::add_constant_to_diagonal(double alpha)
{   // HMat2d.cpp:704-744
    if (type == NODE) {
        for (int i = 0; i < nc(); ++i) child(i, i).add_constant_to_diagonal(alpha);
    } else if (type == DENSE) {
        for (int j = 0; j < std::min(D.m, D.n); ++j) D(j, j) += alpha;
    }
}

// --------------------------------------------------------------------------
// Products
// --------------------------------------------------------------------------
static void set_product_geometry(const HMat& A, const HMat& B, HMat& C)
{   // MultiplyHMat-incl.hpp:27-43
    C.dim = A.dim; C.numLevels = A.numLevels; C.maxRank = A.maxRank;
    C.tOff = A.tOff; C.sOff = B.sOff;
    C.strong = A.strong || B.strong;
    for (int a = 0; a < 3; ++a) {
        C.tSize[a] = A.tSize[a]; C.sSize[a] = B.sSize[a];
        C.tBox[a] = A.tBox[a]; C.sBox[a] = B.sBox[a];
    }
}

// Randomized F := alpha A B (hmat_tools.hpp:1657-1746)
void randomized_multiply(int sampleRank, double alpha, const HMat& A, const HMat& B,
                         LowRank& F)
{
    const int maxRankA = std::min(A.height(), A.width());
    const int maxRankB = std::min(B.height(), B.width());
    const int r = std::min(std::min(A.maxRank, B.maxRank), std::min(maxRankA, maxRankB));
    const int nB = B.width(), mA = A.height();
    Dense Omega(nB, sampleRank);
    g_rng.serial.gaussian_vectors(nB, sampleRank, Omega.a.data(), g_opts.refQuirk);
    Dense X(B.height(), sampleRank), Y(mA, sampleRank);
    B.multiply_dense(false, alpha, Omega.a.data(), nB, sampleRank, 0.0, X.a.data(), X.m);
    A.multiply_dense(false, 1.0, X.a.data(), X.m, sampleRank, 0.0, Y.a.data(), mA);
    const int minDim = std::min(mA, sampleRank);
    {
        std::vector<int> jpvt(sampleRank);
        std::vector<double> tau(minDim);
        pivoted_qr(mA, sampleRank, Y.a.data(), mA, jpvt.data(), tau.data());
        form_q(mA, minDim, minDim, Y.a.data(), mA, tau.data());
        Y.keep_cols(minDim);
    }
    Dense X2(A.width(), minDim), FV(nB, minDim);
    A.multiply_dense(true, alpha, Y.a.data(), mA, minDim, 0.0, X2.a.data(), X2.m);
    B.multiply_dense(true, 1.0, X2.a.data(), X2.m, minDim, 0.0, FV.a.data(), nB);
    const int k = std::min(nB, minDim);
    Dense Us(nB, k), VT(k, minDim);
    std::vector<double> s(k);
    svd(nB, minDim, FV.a.data(), nB, Us.a.data(), s.data(), VT.a.data());
    const int rr = truncated_rank(s.data(), k, std::min(r, k));
    F.V.resize(nB, rr);
    for (int j = 0; j < rr; ++j)
        for (int i = 0; i < nB; ++i) F.V(i, j) = Us(i, j);
    // F.U := Q (Sigma VT)^T
    Dense SVT(rr, minDim);
    for (int j = 0; j < minDim; ++j)
        for (int i = 0; i < rr; ++i) SVT(i, j) = s[i] * VT(i, j);
    F.U.resize(mA, rr);
    gemm('N', 'T', mA, rr, minDim, 1.0, Y.a.data(), mA, SVT.a.data(), std::max(rr, 1), 0.0,
         F.U.a.data(), mA);
}

// W := alpha A B for the cases where A or B is low rank.
static void lowrank_product(double alpha, const HMat& A, const HMat& B, LowRank& W)
{
    if (A.type == HMat::LOW_RANK && B.type == HMat::LOW_RANK) mult_ff(alpha, A.F, B.F, W);
    else if (A.type == HMat::LOW_RANK) {
        W.U = A.F.U;
        W.V.resize(B.width(), A.F.rank());
        if (B.type == HMat::DENSE) { mult_fd(alpha, A.F, B.D, W); }
        else B.multiply_dense(true, alpha, A.F.V.a.data(), std::max(A.F.V.m, 1), A.F.rank(),
                              0.0, W.V.a.data(), std::max(W.V.m, 1));
    } else {
        W.V = B.F.V;
        W.U.resize(A.height(), B.F.rank());
        if (A.type == HMat::DENSE) { mult_df(alpha, A.D, B.F, W); }
        else A.multiply_dense(false, alpha, B.F.U.a.data(), std::max(B.F.U.m, 1), B.F.rank(),
                              0.0, W.U.a.data(), std::max(W.U.m, 1));
    }
}

// Dense C := alpha A B + beta C for leaf operands (MultiplyMatrix.cpp:14-241)
static void dense_product(double alpha, const HMat& A, const HMat& B, double beta, Dense& C)
{
    const int m = A.height(), n = B.width();
    if (A.type == HMat::DENSE && B.type == HMat::DENSE) {
        gemm('N', 'N', m, n, A.width(), alpha, A.D.a.data(), std::max(A.D.m, 1), B.D.a.data(),
             std::max(B.D.m, 1), beta, C.a.data(), std::max(m, 1));
        return;
    }
    LowRank W;
    lowrank_product(alpha, A, B, W);
    lowrank_to_dense(1.0, W, beta, C);
}

void HMat::multiply(double alpha, const HMat& A, const HMat& B, HMat& C)
{
    if (A.width() != B.height() || A.numLevels != B.numLevels)
        throw std::logic_error("Attempted nonconformal matrix-matrix multiply");
    set_product_geometry(A, B, C);
    C.ch.clear(); C.F = LowRank(); C.D = Dense();
    const bool aH = (A.type == NODE), bH = (B.type == NODE);
    if (C.admissible()) {
        C.type = LOW_RANK;
        if (aH && bH) {
            randomized_multiply(C.maxRank + g_opts.oversample, alpha, A, B, C.F);
        } else if (A.type == DENSE && B.type == DENSE) {
            mult_dd_trunc(C.maxRank, alpha, A.D, B.D, C.F);
        } else {
            lowrank_product(alpha, A, B, C.F);
        }
    } else if (C.numLevels > 1) {
        C.type = NODE;
        if (!aH || !bH) {
            LowRank W;
            lowrank_product(alpha, A, B, W);
            C.import_low_rank(W);
        } else {
            make_children(C);
            const int nc = C.nc();
            for (int t = 0; t < nc; ++t)
                for (int s = 0; s < nc; ++s) {
                    multiply(alpha, A.child(t, 0), B.child(0, s), C.child(t, s));
                    for (int u = 1; u < nc; ++u)
                        multiply_update(alpha, A.child(t, u), B.child(u, s), 1.0, C.child(t, s));
                }
        }
    } else {
        C.type = DENSE;
        C.D.resize(C.height(), C.width());
        dense_product(alpha, A, B, 0.0, C.D);
    }
}

void HMat::multiply_child(double alpha, const HMat& A, const HMat& B, HMat& C, int t, int s)
{
    if (A.width() != B.height() || A.numLevels != B.numLevels)
        throw std::logic_error("Attempted nonconformal matrix-matrix multiply");
    if (C.type != NODE || C.ch.empty()) {
        set_product_geometry(A, B, C);
        C.ch.clear(); C.F = LowRank(); C.D = Dense();
        if (A.type != NODE || B.type != NODE || C.admissible() || C.numLevels <= 1)
            throw std::logic_error("a child product needs NODE operands and a NODE product");
        C.type = NODE;
        make_children(C);
    }
    const int nc = C.nc();
    multiply(alpha, A.child(t, 0), B.child(0, s), C.child(t, s));
    for (int u = 1; u < nc; ++u)
        multiply_update(alpha, A.child(t, u), B.child(u, s), 1.0, C.child(t, s));
}

void HMat::multiply_child_row(double alpha, const HMat& A, const HMat& B, HMat& C, int t)
{
    C.ch.clear();
    C.type = LOW_RANK;                  // forces multiply_child to set up C's children
    for (int s = 0; s < (1 << A.dim); ++s) multiply_child(alpha, A, B, C, t, s);
}

void HMat::multiply_update(double alpha, const HMat& A, const HMat& B, double beta, HMat& C)
{
    const bool aH = (A.type == NODE), bH = (B.type == NODE);
    if (C.admissible()) {
        if (aH && bH) {
            LowRank W;
            randomized_multiply(C.maxRank + g_opts.oversample, alpha, A, B, W);
            rounded_update(C.maxRank, 1.0, W, beta, C.F);
        } else if (A.type == DENSE && B.type == DENSE) {
            mult_dd_update(C.maxRank, alpha, A.D, B.D, beta, C.F);
        } else {
            LowRank W;
            lowrank_product(alpha, A, B, W);
            rounded_update(C.maxRank, 1.0, W, beta, C.F);
        }
    } else if (C.numLevels > 1) {
        if (!aH || !bH) {
            LowRank W;
            lowrank_product(alpha, A, B, W);
            C.scale(beta);
            C.update_with_low_rank(1.0, W);
        } else {
            const int nc = C.nc();
            for (int t = 0; t < nc; ++t)
                for (int s = 0; s < nc; ++s) {
                    multiply_update(alpha, A.child(t, 0), B.child(0, s), beta, C.child(t, s));
                    for (int u = 1; u < nc; ++u)
                        multiply_update(alpha, A.child(t, u), B.child(u, s), 1.0, C.child(t, s));
                }
        }
    } else {
        dense_product(alpha, A, B, beta, C.D);
    }
}

// EstimateTwoNorm (hmat_tools.hpp:1504-1544), serial stream.
double estimate_two_norm(const HMat& A, double theta, double confidence)
{
    const int n = A.height();
    const int k = (int)std::ceil(std::log(0.8 * std::sqrt((double)n) * std::pow(10.0, confidence)) /
                                 std::log(theta));
    std::vector<double> x(n), y(n);
    g_rng.serial.gaussian_vector(n, x.data());
    auto nrm = [&](const std::vector<double>& v) {
        double s = 0; for (double e : v) s += e * e; return std::sqrt(s);
    };
    double t = nrm(x);
    for (double& e : x) e /= t;
    double estimate = theta;
    const double root = 1.0 / k;
    for (int i = 0; i < k; ++i) {
        A.multiply_dense(false, 1.0, x.data(), n, 1, 0.0, y.data(), n);
        x = y;
        t = nrm(x);
        for (double& e : x) e /= t;
        estimate *= std::pow(t, root);
    }
    return estimate;
}

void HMat::direct_invert()
{   // Invert-incl.hpp:17-125: Gauss-Jordan over the children of a NODE, the
    // inverse formed in *this (starting as the identity), B a working copy.
    if (height() != width()) throw std::logic_error("Cannot invert non-square matrices");
    switch (type) {
    case NODE: {
        HMat B;
        B.copy_from(*this);
        set_to_identity();
        const int nc = this->nc();
        for (int l = 0; l < nc; ++l) {
            child(l, l).copy_from(B.child(l, l));
            child(l, l).direct_invert();
            for (int j = 0; j < l; ++j) {          // A_lj := A_ll A_lj
                HMat C;
                C.copy_from(child(l, j));
                multiply(1.0, child(l, l), C, child(l, j));
            }
            for (int j = l + 1; j < nc; ++j) {     // B_lj := A_ll B_lj
                HMat C;
                C.copy_from(B.child(l, j));
                multiply(1.0, child(l, l), C, B.child(l, j));
            }
            for (int i = l + 1; i < nc; ++i) {
                for (int j = 0; j <= l; ++j)       // A_ij -= B_il A_lj
                    multiply_update(-1.0, B.child(i, l), child(l, j), 1.0, child(i, j));
                for (int j = l + 1; j < nc; ++j)   // B_ij -= B_il B_lj
                    multiply_update(-1.0, B.child(i, l), B.child(l, j), 1.0, B.child(i, j));
            }
        }
        for (int l = nc - 1; l >= 0; --l)
            for (int i = l - 1; i >= 0; --i)
                for (int j = 0; j < nc; ++j)       // A_ij -= B_il A_lj
                    multiply_update(-1.0, B.child(i, l), child(l, j), 1.0, child(i, j));
        break;
    }
    case DENSE:
        invert(D.m, D.a.data(), D.m);
        break;
    case LOW_RANK:
        throw std::logic_error("Cannot invert low-rank matrices");
    }
}

int HMat::schulz_invert(int numIterations, double theta, double confidence, double stopTol,
                        int maxIter, std::vector<double>* history)
{   // Invert-incl.hpp:128-196
    bool stopflag = false;
    if (numIterations <= 0) { numIterations = maxIter > 0 ? maxIter : 30; stopflag = true; }
    const double est = estimate_two_norm(*this, theta, confidence);
    const double alpha = 2.0 / (est * est);
    HMat X;
    X.transpose_from(*this);
    X.scale(alpha);
    int it = 0;
    for (int k = 0; k < numIterations; ++k) {
        it = k + 1;
        HMat Z;
        multiply(-1.0, X, *this, Z);
        if (stopflag) {
            Z.add_constant_to_diagonal(1.0);
            const double e = estimate_two_norm(Z, theta, confidence);
            if (history) history->push_back(e);
            if (std::fabs(e) < stopTol) {
                Z.add_constant_to_diagonal(1.0);
                HMat XCopy; XCopy.copy_from(X);
                multiply(1.0, Z, XCopy, X);
                break;
            }
            Z.add_constant_to_diagonal(1.0);
        } else Z.add_constant_to_diagonal(2.0);
        HMat XCopy; XCopy.copy_from(X);
        multiply(1.0, Z, XCopy, X);
    }
    copy_from(X);
    return it;
}

void HMat::to_dense(double* out, int ld) const
{
    switch (type) {
    case NODE:
        for_children(*this, [&](int t, int s, int to, int so) {
            child(t, s).to_dense(out + to + (size_t)so * ld, ld);
        });
        break;
    case LOW_RANK: {
        const int m = height(), n = width();
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) out[i + (size_t)j * ld] = 0.0;
        gemm('N', 'T', m, n, F.rank(), 1.0, F.U.a.data(), std::max(m, 1), F.V.a.data(),
             std::max(n, 1), 0.0, out, ld);
        break;
    }
    case DENSE:
        for (int j = 0; j < D.n; ++j)
            for (int i = 0; i < D.m; ++i) out[i + (size_t)j * ld] = D(i, j);
        break;
    }
}

void HMat::count_blocks(std::vector<long>& dl, std::vector<long>& fl, int level) const
{
    if ((int)dl.size() <= level) { dl.resize(level + 1, 0); fl.resize(level + 1, 0); }
    if (type == NODE) for (auto& c : ch) c->count_blocks(dl, fl, level + 1);
    else if (type == DENSE) dl[level]++;
    else fl[level]++;
}

void HMat::rank_stats(long& sumRank, int& maxR, long& storage) const
{
    if (type == NODE) for (auto& c : ch) c->rank_stats(sumRank, maxR, storage);
    else if (type == DENSE) storage += (long)D.m * D.n;
    else {
        sumRank += F.rank(); maxR = std::max(maxR, F.rank());
        storage += (long)(F.U.m + F.V.m) * F.rank();
    }
}

} // namespace orc

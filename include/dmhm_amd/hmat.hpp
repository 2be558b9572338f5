// Header-only C++ facade over the dmhm_amd C-ABI (include/dmhm_amd.h).
//
// It re-exposes the sequential H-matrix interface of the reference library
// with the same class names, method names, argument order and meaning:
//   dmhm::HMat2d<Scalar>   include/dmhm/hmat2d.hpp:22-275
//   dmhm::HMat3d<Scalar>   include/dmhm/hmat3d.hpp:22-290
//   dmhm::Dense / Vector / LowRank / Sparse
//                          include/dmhm/core/{dense,vector,low_rank,sparse}.hpp
//   dmhm::hmat_tools::EstimateTwoNorm / RoundedUpdate / Compress
//                          include/dmhm/hmat_tools.hpp:44-160,1504-1544
//   dmhm::Initialize / Finalize / Oversample / SetOversample
//                          include/dmhm/core/environment.hpp:38-39,82-83
// so a sequential caller written against those headers compiles against this
// one. Every computation runs in libdmhm_amd.so on the GPU; this header only
// converts containers and turns status codes into exceptions.
//
// Differences a caller can see:
//  * Scalar is float or double. The complex instantiations of the reference
//    (HMat2d.cpp:1321-1324) are not provided.
//  * Symmetric (NODE_SYMMETRIC) storage is not provided: the constructors that
//    take `symmetric` reject `true` (the reference's H*H and Schulz throw for
//    it as well, MultiplyHMat-incl.hpp:139-140).
//  * Sub-matrix constructors (explicit offsets / source-target sizes) are not
//    provided; every HMat2d/HMat3d is a square top-level matrix.
//  * Errors raise std::logic_error (bad arguments, nonconformal operands) or
//    std::runtime_error (device failures) in every build mode; the reference
//    checks only when RELEASE is off.
//  * The random streams differ from the reference's serial LCG: SetToRandom
//    and the sketches draw from a counter-based generator keyed by a seed.
//  * Pack / Unpack use this library's own byte format (not the reference's).
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../dmhm_amd.h"

namespace dmhm {

namespace detail {
inline void check(int rc)
{
    if (rc == DMHM_OK) return;
    const char* msg = dmhm_last_error();
    const std::string text = msg ? msg : "dmhm_amd error";
    if (rc == DMHM_EINVAL || rc == DMHM_ENONCONF) throw std::logic_error(text);
    throw std::runtime_error(text);
}
inline int& oversample_setting()
{
    static int value = 4;   // Oversample() default (src/core/global.cpp:30)
    return value;
}
inline uint64_t next_seed()
{
    static uint64_t counter = 1;
    return counter++;
}
template <typename Scalar> struct precision_of;
template <> struct precision_of<double> { static constexpr int value = DMHM_F64; };
template <> struct precision_of<float> { static constexpr int value = DMHM_F32; };
} // namespace detail

typedef unsigned char byte;

// Environment (include/dmhm/core/environment.hpp:38-39): binds GPU 0.
inline void Initialize(int& /*argc*/, char**& /*argv*/) { detail::check(dmhm_init(0)); }
inline void Initialize(int device = 0) { detail::check(dmhm_init(device)); }
inline void Finalize() { detail::check(dmhm_finalize()); }
inline int Oversample() { return detail::oversample_setting(); }
inline void SetOversample(int oversample) { detail::oversample_setting() = oversample; }

// ---------------------------------------------------------------------------
// Containers (owning, column-major; the reference's views are not needed by
// callers of the sequential interface and are left out).
// ---------------------------------------------------------------------------
template <typename T>
class Vector {
public:
    Vector() = default;
    explicit Vector(int height) : data_(height > 0 ? height : 0) {}
    int Height() const { return (int)data_.size(); }
    int Size() const { return (int)data_.size(); }
    void Resize(int height) { data_.resize(height > 0 ? height : 0); }
    void Clear() { data_.clear(); }
    void PushBack(T value) { data_.push_back(value); }
    T Get(int i) const { return data_[i]; }
    void Set(int i, T value) { data_[i] = value; }
    T& operator[](int i) { return data_[i]; }
    const T& operator[](int i) const { return data_[i]; }
    T* Buffer(int i = 0) { return data_.data() + i; }
    const T* LockedBuffer(int i = 0) const { return data_.data() + i; }
private:
    std::vector<T> data_;
};

enum MatrixType { GENERAL, SYMMETRIC };

template <typename T>
class Dense {
public:
    Dense() = default;
    Dense(int height, int width, MatrixType type = GENERAL) { Resize(height, width); type_ = type; }
    int Height() const { return height_; }
    int Width() const { return width_; }
    int LDim() const { return ldim_; }
    MatrixType Type() const { return type_; }
    void SetType(MatrixType type) { type_ = type; }
    // Resize keeps the leading columns when only the width shrinks (in-place
    // truncation, include/dmhm/core/dense.hpp:256-282).
    void Resize(int height, int width)
    {
        if (height == height_) {
            data_.resize((size_t)height * width);
            width_ = width;
            return;
        }
        height_ = height;
        width_ = width;
        ldim_ = height > 1 ? height : 1;
        data_.assign((size_t)ldim_ * width, T(0));
    }
    void Clear() { height_ = width_ = 0; ldim_ = 1; data_.clear(); }
    T Get(int i, int j) const { return data_[i + (size_t)j * ldim_]; }
    void Set(int i, int j, T value) { data_[i + (size_t)j * ldim_] = value; }
    T* Buffer(int i = 0, int j = 0) { return data_.data() + i + (size_t)j * ldim_; }
    const T* LockedBuffer(int i = 0, int j = 0) const { return data_.data() + i + (size_t)j * ldim_; }
private:
    int height_ = 0, width_ = 0, ldim_ = 1;
    MatrixType type_ = GENERAL;
    std::vector<T> data_;
};

// A = U V^T (transpose, not adjoint; include/dmhm/core/low_rank.hpp:23-35).
template <typename T>
struct LowRank {
    Dense<T> U, V;
    int Height() const { return U.Height(); }
    int Width() const { return V.Height(); }
    int Rank() const { return U.Width(); }
};

// CSR in hierarchical ordering (include/dmhm/core/sparse.hpp:17-26).
template <typename T>
struct Sparse {
    bool symmetric = false;
    int height = 0, width = 0;
    Vector<T> nonzeros;
    Vector<int> columnIndices;
    Vector<int> rowOffsets;
};

// ---------------------------------------------------------------------------
// H-matrices
// ---------------------------------------------------------------------------
namespace detail {

template <typename Scalar, int Dim>
class HMatImpl {
    static_assert(std::is_same<Scalar, double>::value || std::is_same<Scalar, float>::value,
                  "dmhm_amd provides real float and double H-matrices");
public:
    typedef Scalar Real;
    static constexpr int kPrecision = precision_of<Scalar>::value;

    static int SampleRank(int approxRank) { return approxRank + Oversample(); }

    HMatImpl() = default;
    HMatImpl(const HMatImpl&) = delete;
    HMatImpl& operator=(const HMatImpl&) = delete;
    HMatImpl(HMatImpl&& o) noexcept { std::swap(h_, o.h_); }
    HMatImpl& operator=(HMatImpl&& o) noexcept { std::swap(h_, o.h_); return *this; }
    virtual ~HMatImpl() { Clear(); }

    void Clear()
    {
        if (h_) dmhm_hmat_destroy(h_);
        h_ = nullptr;
    }
    dmhm_hmat Handle() const { return h_; }

    void SetToRandom() { detail::check(dmhm_hmat_set_to_random(need(), next_seed())); }

    int Height() const { return info().n; }
    int Width() const { return info().n; }
    int NumLevels() const { return info().numLevels; }
    int MaxRank() const { return info().maxRank; }
    int SourceOffset() const { return 0; }
    int TargetOffset() const { return 0; }
    bool Symmetric() const { return false; }
    bool StronglyAdmissible() const { return info().strong != 0; }

    // y := alpha H x + beta y  /  y := alpha H x
    void Multiply(Scalar alpha, const Vector<Scalar>& x, Scalar beta, Vector<Scalar>& y) const
    { mult('N', alpha, x.LockedBuffer(), Height(), 1, beta, y.Buffer(), Height()); }
    void Multiply(Scalar alpha, const Vector<Scalar>& x, Vector<Scalar>& y) const
    { y.Resize(Height()); Multiply(alpha, x, Scalar(0), y); }
    void TransposeMultiply(Scalar alpha, const Vector<Scalar>& x, Scalar beta, Vector<Scalar>& y) const
    { mult('T', alpha, x.LockedBuffer(), Height(), 1, beta, y.Buffer(), Height()); }
    void TransposeMultiply(Scalar alpha, const Vector<Scalar>& x, Vector<Scalar>& y) const
    { y.Resize(Height()); TransposeMultiply(alpha, x, Scalar(0), y); }
    void AdjointMultiply(Scalar alpha, const Vector<Scalar>& x, Scalar beta, Vector<Scalar>& y) const
    { TransposeMultiply(alpha, x, beta, y); }
    void AdjointMultiply(Scalar alpha, const Vector<Scalar>& x, Vector<Scalar>& y) const
    { TransposeMultiply(alpha, x, y); }

    // C := alpha H B + beta C  /  C := alpha H B
    void Multiply(Scalar alpha, const Dense<Scalar>& B, Scalar beta, Dense<Scalar>& C) const
    { mult('N', alpha, B.LockedBuffer(), B.LDim(), B.Width(), beta, C.Buffer(), C.LDim()); }
    void Multiply(Scalar alpha, const Dense<Scalar>& B, Dense<Scalar>& C) const
    { C.Resize(Height(), B.Width()); Multiply(alpha, B, Scalar(0), C); }
    void TransposeMultiply(Scalar alpha, const Dense<Scalar>& B, Scalar beta, Dense<Scalar>& C) const
    { mult('T', alpha, B.LockedBuffer(), B.LDim(), B.Width(), beta, C.Buffer(), C.LDim()); }
    void TransposeMultiply(Scalar alpha, const Dense<Scalar>& B, Dense<Scalar>& C) const
    { C.Resize(Height(), B.Width()); TransposeMultiply(alpha, B, Scalar(0), C); }
    void AdjointMultiply(Scalar alpha, const Dense<Scalar>& B, Scalar beta, Dense<Scalar>& C) const
    { TransposeMultiply(alpha, B, beta, C); }
    void AdjointMultiply(Scalar alpha, const Dense<Scalar>& B, Dense<Scalar>& C) const
    { TransposeMultiply(alpha, B, C); }

    // Whole-matrix routines (src/HMat2d/HMat2d.cpp:331-744).
    void CopyFrom(const HMatImpl& B) { replace([&](dmhm_hmat* out) { return dmhm_hmat_copy(B.need(), out); }); }
    void ConjugateFrom(const HMatImpl& B) { CopyFrom(B); }
    void Conjugate() {}
    void TransposeFrom(const HMatImpl& B)
    { replace([&](dmhm_hmat* out) { return dmhm_hmat_transpose(B.need(), out); }); }
    void AdjointFrom(const HMatImpl& B) { TransposeFrom(B); }
    void Scale(Scalar alpha) { detail::check(dmhm_hmat_scale(need(), alpha)); }
    void SetToIdentity() { detail::check(dmhm_hmat_set_to_identity(need())); }
    void AddConstantToDiagonal(Scalar alpha) { detail::check(dmhm_hmat_add_constant_to_diagonal(need(), alpha)); }
    // A :~= alpha B + A
    void UpdateWith(Scalar alpha, const HMatImpl& B)
    {
        const dmhm_hh_opts o = hh_options();
        detail::check(dmhm_hmat_update_with(need(), alpha, B.need(), &o));
    }
    // C :~= alpha A B
    void Multiply(Scalar alpha, const HMatImpl& B, HMatImpl& C) const
    {
        const dmhm_hh_opts o = hh_options();
        C.replace([&](dmhm_hmat* out) { return dmhm_hmat_multiply_hmat(alpha, need(), B.need(), out, &o); });
    }
    // C :~= alpha A B + beta C
    void Multiply(Scalar alpha, const HMatImpl& B, Scalar beta, HMatImpl& C) const
    {
        const dmhm_hh_opts o = hh_options();
        detail::check(dmhm_hmat_multiply_hmat_update(alpha, need(), B.need(), beta, C.need(), &o));
    }
    // A :~= inv(A) by Schulz iteration (src/HMat2d/Invert-incl.hpp:128-196).
    // A := inv(A) by recursive Schur complements (include/dmhm/hmat2d.hpp,
    // src/HMat2d/Invert-incl.hpp:17-125); sub-matrix products use the H*H
    // options of this handle's class (oversample, fixed rank).
    void DirectInvert()
    {
        const dmhm_hh_opts o = hh_options();
        detail::check(dmhm_hmat_direct_invert(need(), &o));
    }

    // numIterations <= 0 stops on the estimate of ||I - X A||_2 < stopTol or
    // after 30 iterations, as the reference does; not reaching stopTol is
    // not an error (LastSchulzReport() holds the residual history).
    void SchulzInvert(int numIterations, Real theta = 1.5, Real confidence = 6, Real stopTol = 1e-4)
    {
        dmhm_schulz_opts o = SchulzOptions();
        o.numIterations = numIterations;
        o.theta = theta;
        o.confidence = confidence;
        o.stopTol = stopTol;
        SchulzInvert(o);
    }
    // Extended form: mixed precision, iteration cap, H*H options.
    void SchulzInvert(const dmhm_schulz_opts& o)
    {
        const int rc = dmhm_hmat_schulz_invert(need(), &o, &report_);
        if (rc != DMHM_ENOCONV) detail::check(rc);
    }
    static dmhm_schulz_opts SchulzOptions()
    {
        dmhm_schulz_opts o;
        std::memset(&o, 0, sizeof(o));
        o.numIterations = 0;
        o.maxIterations = 30;
        o.theta = 1.5;
        o.confidence = 6;
        o.stopTol = 1e-4;
        o.mixed = 0;
        o.switchTol = 1e-3;
        o.hh = hh_options();
        o.stallStop = 0;
        o.graph = 1;
        o.bf16 = 0;
        o.bf16Tol = 0.5;
        o.start = 0;
        return o;
    }
    const dmhm_schulz_report& LastSchulzReport() const { return report_; }

    // Packed byte image (include/dmhm/hmat2d.hpp:89,108-112; own format, see
    // dmhm_hmat_pack).
    std::size_t PackedSize() const
    {
        uint64_t n = 0;
        detail::check(dmhm_hmat_packed_size(need(), &n));
        return (std::size_t)n;
    }
    std::size_t Pack(byte* packedHMat) const
    {
        uint64_t w = 0;
        detail::check(dmhm_hmat_pack(need(), packedHMat, PackedSize(), &w));
        return (std::size_t)w;
    }
    std::size_t Pack(Vector<byte>& packedHMat) const
    {
        packedHMat.Resize((int)PackedSize());
        return Pack(packedHMat.Buffer());
    }
    std::size_t Unpack(const byte* packedHMat, std::size_t size)
    {
        replace([&](dmhm_hmat* out) { return dmhm_hmat_unpack(packedHMat, size, out); });
        return size;
    }
    // The reference's form (include/dmhm/hmat2d.hpp:111): the image carries
    // its own length; returns the bytes read.
    std::size_t Unpack(const byte* packedHMat)
    {
        uint64_t n = 0;
        replace([&](dmhm_hmat* out) { return dmhm_hmat_unpack_image(packedHMat, &n, out); });
        return (std::size_t)n;
    }
    std::size_t Unpack(const Vector<byte>& packedHMat)
    {
        return Unpack(packedHMat.LockedBuffer(), (std::size_t)packedHMat.Size());
    }

protected:
    dmhm_hmat h_ = nullptr;
    dmhm_schulz_report report_{};

    dmhm_hmat need() const
    {
        if (!h_) throw std::logic_error("H-matrix is empty");
        return h_;
    }
    dmhm_info info() const
    {
        dmhm_info inf;
        detail::check(dmhm_hmat_info(need(), &inf));
        return inf;
    }
    static dmhm_hh_opts hh_options()
    {
        dmhm_hh_opts o;
        o.oversample = Oversample();
        o.tolMode = 0;
        o.relTol = 1e-10;
        o.seed = next_seed();
        o.rankModel = nullptr;
        o.rankMargin = 0;
        return o;
    }
    template <typename F> void replace(F make)
    {
        dmhm_hmat out = nullptr;
        detail::check(make(&out));
        Clear();
        h_ = out;
    }
    void mult(char op, Scalar alpha, const Scalar* X, int ldx, int k, Scalar beta, Scalar* Y, int ldy) const
    {
        detail::check(dmhm_hmat_multiply_dense(need(), op, alpha, X, ldx, k, beta, Y, ldy, DMHM_HOST));
    }
    void create_zero(const int* sizes, int numLevels, int maxRank, bool symmetric, bool strong)
    {
        if (symmetric) throw std::logic_error("symmetric H-matrix storage is not provided");
        replace([&](dmhm_hmat* out) {
            return dmhm_hmat_create_zero(Dim, sizes, numLevels, maxRank, strong, kPrecision, out);
        });
    }
    void create_sparse(const Sparse<Scalar>& S, const int* sizes, int numLevels, int maxRank, bool strong)
    {
        if (S.symmetric) throw std::logic_error("symmetric sparse input is not provided");
        std::vector<double> nz(S.nonzeros.Size());
        for (int i = 0; i < S.nonzeros.Size(); ++i) nz[i] = S.nonzeros[i];
        replace([&](dmhm_hmat* out) {
            return dmhm_hmat_create_from_csr(Dim, sizes, numLevels, maxRank, strong, kPrecision,
                                             S.rowOffsets.LockedBuffer(), S.columnIndices.LockedBuffer(),
                                             nz.data(), out);
        });
    }
    void create_lowrank(const LowRank<Scalar>& F, const int* sizes, int numLevels, int maxRank, bool strong)
    {
        const int n = F.Height(), r = F.Rank();
        std::vector<double> U((size_t)n * r), V((size_t)n * r);
        for (int j = 0; j < r; ++j)
            for (int i = 0; i < n; ++i) {
                U[i + (size_t)j * n] = F.U.Get(i, j);
                V[i + (size_t)j * n] = F.V.Get(i, j);
            }
        replace([&](dmhm_hmat* out) {
            return dmhm_hmat_create_from_lowrank(Dim, sizes, numLevels, maxRank, strong, kPrecision, r,
                                                 U.data(), V.data(), out);
        });
    }
};

} // namespace detail

template <typename Scalar>
class HMat2d : public detail::HMatImpl<Scalar, 2> {
    typedef detail::HMatImpl<Scalar, 2> Base;
public:
    // Natural x + xSize*y -> hierarchical index (HMat2d.cpp:27-87).
    static void BuildNaturalToHierarchicalMap(Vector<int>& map, int xSize, int ySize, int numLevels)
    {
        const int sizes[2] = {xSize, ySize};
        map.Resize(xSize * ySize);
        detail::check(dmhm_build_natural_to_hierarchical_map(2, sizes, numLevels, map.Buffer()));
    }
    HMat2d() = default;
    explicit HMat2d(const Vector<byte>& packedHMat) { this->Unpack(packedHMat); }
    HMat2d(int numLevels, int maxRank, bool symmetric, bool stronglyAdmissible, int xSize, int ySize)
    {
        const int sizes[2] = {xSize, ySize};
        this->create_zero(sizes, numLevels, maxRank, symmetric, stronglyAdmissible);
    }
    HMat2d(const LowRank<Scalar>& F, int numLevels, int maxRank, bool stronglyAdmissible, int xSize, int ySize)
    {
        const int sizes[2] = {xSize, ySize};
        this->create_lowrank(F, sizes, numLevels, maxRank, stronglyAdmissible);
    }
    HMat2d(const Sparse<Scalar>& S, int numLevels, int maxRank, bool stronglyAdmissible, int xSize, int ySize)
    {
        const int sizes[2] = {xSize, ySize};
        this->create_sparse(S, sizes, numLevels, maxRank, stronglyAdmissible);
    }
    using Base::Multiply;
};

template <typename Scalar>
class HMat3d : public detail::HMatImpl<Scalar, 3> {
    typedef detail::HMatImpl<Scalar, 3> Base;
public:
    // Natural x + xSize*y + xSize*ySize*z -> hierarchical index (HMat3d.cpp:28-90).
    static void BuildNaturalToHierarchicalMap(Vector<int>& map, int xSize, int ySize, int zSize, int numLevels)
    {
        const int sizes[3] = {xSize, ySize, zSize};
        map.Resize(xSize * ySize * zSize);
        detail::check(dmhm_build_natural_to_hierarchical_map(3, sizes, numLevels, map.Buffer()));
    }
    HMat3d() = default;
    explicit HMat3d(const Vector<byte>& packedHMat) { this->Unpack(packedHMat); }
    HMat3d(int numLevels, int maxRank, bool symmetric, bool stronglyAdmissible, int xSize, int ySize, int zSize)
    {
        const int sizes[3] = {xSize, ySize, zSize};
        this->create_zero(sizes, numLevels, maxRank, symmetric, stronglyAdmissible);
    }
    HMat3d(const LowRank<Scalar>& F, int numLevels, int maxRank, bool stronglyAdmissible, int xSize, int ySize,
           int zSize)
    {
        const int sizes[3] = {xSize, ySize, zSize};
        this->create_lowrank(F, sizes, numLevels, maxRank, stronglyAdmissible);
    }
    HMat3d(const Sparse<Scalar>& S, int numLevels, int maxRank, bool stronglyAdmissible, int xSize, int ySize,
           int zSize)
    {
        const int sizes[3] = {xSize, ySize, zSize};
        this->create_sparse(S, sizes, numLevels, maxRank, stronglyAdmissible);
    }
    using Base::Multiply;
};

// ---------------------------------------------------------------------------
// hmat_tools
// ---------------------------------------------------------------------------
namespace hmat_tools {

// Power-iteration estimate of ||A||_2 (include/dmhm/hmat_tools.hpp:1504-1544).
template <typename Scalar, int Dim>
Scalar EstimateTwoNorm(const detail::HMatImpl<Scalar, Dim>& A, Scalar theta, Scalar confidence)
{
    double est = 0;
    detail::check(dmhm_hmat_estimate_two_norm(A.Handle(), theta, confidence, detail::next_seed(), &est));
    return (Scalar)est;
}

// B :~= alpha A + beta B, truncated to maxRank (src/hmat_tools/RoundedUpdate.cpp:14-162).
template <typename Scalar>
void RoundedUpdate(int maxRank, Scalar alpha, const LowRank<Scalar>& A, Scalar beta, LowRank<Scalar>& B)
{
    const int m = A.Height(), n = A.Width(), rA = A.Rank(), rB = B.Rank();
    if (B.Height() != m || B.Width() != n) throw std::logic_error("nonconformal RoundedUpdate");
    auto flat = [](const Dense<Scalar>& M, int rows, int cols) {
        std::vector<double> out((size_t)rows * cols);
        for (int j = 0; j < cols; ++j)
            for (int i = 0; i < rows; ++i) out[i + (size_t)j * rows] = M.Get(i, j);
        return out;
    };
    const std::vector<double> AU = flat(A.U, m, rA), AV = flat(A.V, n, rA), BU = flat(B.U, m, rB),
                              BV = flat(B.V, n, rB);
    const int cap = rA + rB > 0 ? rA + rB : 1;
    std::vector<double> U((size_t)m * cap), V((size_t)n * cap);
    int r = 0;
    detail::check(dmhm_rounded_update(maxRank, alpha, m, n, rA, AU.data(), AV.data(), beta, rB, BU.data(),
                                      BV.data(), U.data(), V.data(), &r));
    B.U.Resize(m, r);
    B.V.Resize(n, r);
    for (int j = 0; j < r; ++j) {
        for (int i = 0; i < m; ++i) B.U.Set(i, j, (Scalar)U[i + (size_t)j * m]);
        for (int i = 0; i < n; ++i) B.V.Set(i, j, (Scalar)V[i + (size_t)j * n]);
    }
}

// F := best rank-min(maxRank, m, n) approximation of D (src/hmat_tools/Compress.cpp:14-53).
template <typename Scalar>
void Compress(int maxRank, Dense<Scalar>& D, LowRank<Scalar>& F)
{
    const int m = D.Height(), n = D.Width(), k = m < n ? m : n;
    std::vector<double> d((size_t)m * n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) d[i + (size_t)j * m] = D.Get(i, j);
    std::vector<double> U((size_t)m * (k > 0 ? k : 1)), V((size_t)n * (k > 0 ? k : 1));
    int r = 0;
    detail::check(dmhm_compress_dense(maxRank, m, n, d.data(), U.data(), V.data(), &r));
    F.U.Resize(m, r);
    F.V.Resize(n, r);
    for (int j = 0; j < r; ++j) {
        for (int i = 0; i < m; ++i) F.U.Set(i, j, (Scalar)U[i + (size_t)j * m]);
        for (int i = 0; i < n; ++i) F.V.Set(i, j, (Scalar)V[i + (size_t)j * n]);
    }
}

} // namespace hmat_tools
} // namespace dmhm

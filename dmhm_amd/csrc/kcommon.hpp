// Device helpers shared by the gfx950 kernel translation units (tile, mv,
// qr, svd, misc). Wave64 throughout.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include "kernels.hpp"

namespace dmhm {

#define HIP_LAUNCH_CHECK()                                                        \
    do {                                                                          \
        hipError_t e_ = hipGetLastError();                                        \
        if (e_ != hipSuccess)                                                     \
            throw std::runtime_error(std::string("HIP launch: ") + hipGetErrorString(e_)); \
    } while (0)

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ inline float wave_sum(float v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
// Wave sums of N values at once: each butterfly step issues the N shuffles
// together, so their latency is paid once per step instead of once per value.
template <int N>
__device__ __forceinline__ void wave_sum_all(double (&v)[N])
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        double t[N];
#pragma unroll
        for (int c = 0; c < N; ++c) t[c] = __shfl_xor(v[c], off, 64);
#pragma unroll
        for (int c = 0; c < N; ++c) v[c] += t[c];
    }
}
// Broadcast lane `l` (wave-uniform) of v to the whole wave.
__device__ inline float bcast(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ inline double bcast(double v, int l)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

template <typename T> struct VecOf;
template <> struct VecOf<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct VecOf<double> { typedef double type __attribute__((ext_vector_type(2))); };

// Grid of a grid-stride elementwise launch over n elements.
inline int grid_for(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 8192); }

} // namespace dmhm

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

// Cross-lane moves inside rows of 16 lanes by DPP (VALU, no LDS round trip):
// the xor-1 and xor-2 quad permutations and the mirrors within 8 and 16
// lanes. After an xor-1 and an xor-2 step every lane of a quad holds the
// quad's sum; a mirror then pairs each lane with one of the other half, so
// each step adds the same two partial sums on both sides (a + b and b + a)
// and every lane of a group ends with the bitwise same sum.
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppMirror8 = 0x141, kDppMirror16 = 0x140;

template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over aligned groups of G lanes (G = 1, 2, 4, 8 or 16); every lane of a
// group receives it. The whole group must be active.
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v)
{
    static_assert(G == 1 || G == 2 || G == 4 || G == 8 || G == 16, "group of 1, 2, 4, 8 or 16 lanes");
    if constexpr (G >= 2) v += dpp_mov<kDppXor1>(v);
    if constexpr (G >= 4) v += dpp_mov<kDppXor2>(v);
    if constexpr (G >= 8) v += dpp_mov<kDppMirror8>(v);
    if constexpr (G >= 16) v += dpp_mov<kDppMirror16>(v);
    return v;
}

// Wave sums (all 64 lanes active): DPP inside each row of 16, then two
// shuffles across rows.
__device__ inline double wave_sum(double v)
{
    v = group_sum<16>(v);
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}
__device__ inline float wave_sum(float v)
{
    v = group_sum<16>(v);
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
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

// Device-side restatements of the reference's per-cell arithmetic.
//
// Bit-exactness contract: every function here performs the same IEEE
// operations, in the same order and at the same precision, as the reference's
// CPU code it cites. The translation unit is compiled with -ffp-contract=off
// (no FMA formation), correctly rounded fp32 divide/sqrt and fp32 denormals
// kept, mirroring the reference's FP policy (CMakeLists.txt:76-82).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace oceanamd {

// reference src/core/trig.hpp:36-56
constexpr float kTwoOverPi = 0.636619772367581343076f;
constexpr float kPiOver2A = 1.5707963705062866f;
constexpr float kPiOver2B = -4.3711388286737929e-8f;
constexpr float kSin0 = -1.9515295891e-4f;
constexpr float kSin1 = 8.3321608736e-3f;
constexpr float kSin2 = -1.6666654611e-1f;
constexpr float kCos0 = 2.443315711809948e-5f;
constexpr float kCos1 = -1.388731625493765e-3f;
constexpr float kCos2 = 4.166664568298827e-2f;
constexpr double kTwoPiD = 6.28318530717958647692;
constexpr double kInvTwoPiD = 0.159154943091895335769;

// omega*t folded into [0, 2pi) in double, narrowed (reference trig.hpp:76-80).
__device__ __forceinline__ float reduced_phase(float omega, double time_over_two_pi)
{
    const double d = static_cast<double>(omega) * time_over_two_pi;
    return static_cast<float>((d - floor(d)) * kTwoPiD);
}

// Cody-Waite reduction + minimax polynomials + quadrant select
// (reference trig.hpp:87-129).
__device__ __forceinline__ void sincos_poly(float x, float& out_s, float& out_c)
{
    const float qf = floorf(x * kTwoOverPi + 0.5f);
    const int32_t q = static_cast<int32_t>(qf);
    float y = x - qf * kPiOver2A;
    y = y - qf * kPiOver2B;
    const float z = y * y;
    const float sin_y = y + ((kSin0 * z + kSin1) * z + kSin2) * z * y;
    const float cos_y = 1.0f - 0.5f * z + ((kCos0 * z + kCos1) * z + kCos2) * z * z;
    const int32_t quadrant = q & 3;
    const bool swap = (quadrant & 1) != 0;
    const float bs = swap ? cos_y : sin_y;
    const float bc = swap ? sin_y : cos_y;
    const bool neg_s = (quadrant & 2) != 0;
    const bool neg_c = ((quadrant + 1) & 2) != 0;
    out_s = neg_s ? -bs : bs;
    out_c = neg_c ? -bc : bc;
}

// The opt-in fast mode's phase (not the reference's rounding): omega*t/2pi
// folded into [0, 1) revolutions in double, narrowed, and the hardware's
// sine and cosine, which take revolutions (v_sin_f32 / v_cos_f32).
__device__ __forceinline__ void sincos_fast(float omega, double time_over_two_pi, float& out_s,
                                            float& out_c)
{
    const double d = static_cast<double>(omega) * time_over_two_pi;
    const float rev = static_cast<float>(d - floor(d));
    out_s = __builtin_amdgcn_sinf(rev);
    out_c = __builtin_amdgcn_cosf(rev);
}

// Signed wavenumber of grid index i (reference src/core/spectrum.hpp:85-92).
__device__ __forceinline__ float wave_component(uint32_t i, uint32_t n, float patch_length)
{
    constexpr float kTwoPiF = 6.28318530717958647692f;
    const int si = (i < n / 2) ? static_cast<int>(i) : static_cast<int>(i) - static_cast<int>(n);
    return kTwoPiF * static_cast<float>(si) / patch_length;
}

// 1/|k| with |k| in double, 0 at DC (reference src/core/spectrum.cpp:300-306).
__device__ __forceinline__ float inv_wavenumber(float kx, float kz)
{
    const double k2 = static_cast<double>(kx) * kx + static_cast<double>(kz) * kz;
    const double k = sqrt(k2);
    return (k > 0.0) ? static_cast<float>(1.0 / k) : 0.0f;
}

}  // namespace oceanamd

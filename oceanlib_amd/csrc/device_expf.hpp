// Single-precision exp with the exact results of the host C library's expf.
//
// The reference stamps its Ricker sources with std::exp(float) (reference
// src/interaction.cpp:71-75), i.e. the host libm's expf. glibc >= 2.27 ships
// the table-driven method of ARM's optimized-routines (Szabolcs Nagy, 2017):
// x*32/ln2 = k + r, 2^(k/32) from a 32-entry table with the exponent folded
// in, 2^(r/32) from a cubic in double, one final rounding to float. This is
// that published method written out for the device. On FMA-capable x86-64
// hosts glibc's build of it fuses the argument reduction (z*InvLn2N + shift
// and z*InvLn2N - kd), and so does this one; an exhaustive CPU check
// (tests/test_device_expf.py) shows bit-identical results to the host expf
// for every float in [-110, 0], the range the stamps use.
//
// TAB[i] = bits(2^(i/32)) - (i << 47), from 2^(i/32) rounded to double; the
// polynomial is the published degree-3 fit of 2^(r/32) for |r| <= 1/2.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>

namespace oceanamd {

__host__ __device__ inline float expf_libm(float x)
{
    constexpr std::uint64_t TAB[32] = {
    0x3ff0000000000000ULL, 0x3fefd9b0d3158574ULL,
    0x3fefb5586cf9890fULL, 0x3fef9301d0125b51ULL,
    0x3fef72b83c7d517bULL, 0x3fef54873168b9aaULL,
    0x3fef387a6e756238ULL, 0x3fef1e9df51fdee1ULL,
    0x3fef06fe0a31b715ULL, 0x3feef1a7373aa9cbULL,
    0x3feedea64c123422ULL, 0x3feece086061892dULL,
    0x3feebfdad5362a27ULL, 0x3feeb42b569d4f82ULL,
    0x3feeab07dd485429ULL, 0x3feea47eb03a5585ULL,
    0x3feea09e667f3bcdULL, 0x3fee9f75e8ec5f74ULL,
    0x3feea11473eb0187ULL, 0x3feea589994cce13ULL,
    0x3feeace5422aa0dbULL, 0x3feeb737b0cdc5e5ULL,
    0x3feec49182a3f090ULL, 0x3feed503b23e255dULL,
    0x3feee89f995ad3adULL, 0x3feeff76f2fb5e47ULL,
    0x3fef199bdd85529cULL, 0x3fef3720dcef9069ULL,
    0x3fef5818dcfba487ULL, 0x3fef7c97337b9b5fULL,
    0x3fefa4afa2a490daULL, 0x3fefd0765b6e4540ULL};
    constexpr double kShift = 0x1.8p+52;
    constexpr double kInvLn2N = 0x1.71547652b82fep+0 * 32;
    constexpr double kC0 = 0x1.c6af84b912394p-5 / 32 / 32 / 32;
    constexpr double kC1 = 0x1.ebfce50fac4f3p-3 / 32 / 32;
    constexpr double kC2 = 0x1.62e42ff0c52d6p-1 / 32;
    if (x < -0x1.9fe368p6f) return 0.0f;  // below log(2^-150): underflows to +0
    const double xd = static_cast<double>(x);
    double kd = fma(kInvLn2N, xd, kShift);
    std::uint64_t ki;
    std::memcpy(&ki, &kd, 8);
    kd -= kShift;
    const double r = fma(kInvLn2N, xd, -kd);
    std::uint64_t t = TAB[ki % 32];
    t += ki << 47;
    double s;
    std::memcpy(&s, &t, 8);
    const double z = kC0 * r + kC1;
    const double r2 = r * r;
    double y = kC2 * r + 1.0;
    y = z * r2 + y;
    y = y * s;
    return static_cast<float>(y);
}

}  // namespace oceanamd

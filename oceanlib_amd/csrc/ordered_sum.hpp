// A sequential double sum of floats, s = 0; s += (double)v_i in index order,
// computed by a whole workgroup with the same bits.
//
// FoamField::coverage() is that sum over the field (reference
// src/foam.cpp:307-313). Its rounding depends on the order, so a tree
// reduction is not the same number. The shortcut used here: while the running
// value S > 0 stays inside one binade [2^e, 2^(e+1)), every double it takes is
// a multiple of u = 2^(e-52), and one step of the chain is
//
//     fl(S + t) = S + r(t),   r(t) = t rounded to the nearest multiple of u,
//
// provided S + t does not leave the binade and t is not exactly halfway
// between two multiples of u (a tie, whose rounding depends on S's last bit).
// For non-negative terms the running value only grows, so a run of terms
// t_1..t_k stays in the binade exactly when S + r_1 + ... + r_k < 2^(e+1),
// i.e. when the prefix sum P_k = r_1 + ... + r_k is below room = 2^(e+1) - S
// (an exact difference). Each r_i and every P below room are multiples of u
// below 2^52 u, so they and any sum of consecutive ones are exact doubles in
// any order: a workgroup scans the r_i, finds the first index j where a term
// is flagged (negative, non-finite, a tie) or P_j >= room, adds P_(j-1) to S
// in one exact step, adds t_j the plain way (it may cross into the next
// binade), and goes on from j + 1 with the new binade. The scan's computed
// prefix sums reach room exactly when the exact ones do (each computed
// partial is exact while below room and at least room once the exact value
// is), so j is found without error. A running value that is negative or not
// finite (only possible with negative or non-finite terms) ends the shortcut:
// the rest is added term by term. Shared by the coverage kernel
// (wave_kernels.hip) and its CPU emulation test.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace oceanamd {

struct Binade {
    double inv_u; // 1 / spacing of the doubles in S's binade (a power of two)
    double u;     // the spacing
    double room;  // 2^(e+1) - S: exact
    int mode;     // 0: S > 0 in a binade; 1: S == 0; 2: term by term from here on
};

__host__ __device__ __forceinline__ Binade binade_of(double s)
{
    Binade b{0.0, 0.0, 0.0, 2};
    const uint64_t bits = __builtin_bit_cast(uint64_t, s);
    if (s == 0.0 && (bits >> 63) == 0u) {
        b.mode = 1;
        b.room = __builtin_inf();
        return b;
    }
    const int e = int((bits >> 52) & 0x7FFu);
    // Negative, NaN, infinite, below 2^-970 (u would not be a normal double;
    // a sum of floats never is) or in the top binade: term by term.
    if ((bits >> 63) != 0u || e < 53 || e >= 0x7FE) return b;
    b.mode = 0;
    b.u = __builtin_bit_cast(double, uint64_t(e - 52) << 52);        // 2^(e-1075), S's ulp
    b.inv_u = __builtin_bit_cast(double, uint64_t(2098 - e) << 52);  // 2^(1075-e) <= 2^1022
    const double top = __builtin_bit_cast(double, uint64_t(e + 1) << 52);
    b.room = top - s;
    return b;
}

// r(t) for one term, and whether it must stop the run (a term the binade
// shortcut cannot take: negative, NaN, infinite, a tie; while S == 0, any
// nonzero term, since 0 + t is simply t).
__host__ __device__ __forceinline__ double rounded_term(float t, const Binade& b, bool& stop)
{
    const uint32_t tb = __builtin_bit_cast(uint32_t, t);
    const bool nonneg_finite = (tb & 0x7F800000u) != 0x7F800000u && ((tb >> 31) == 0u || t == 0.0f);
    if (b.mode == 1) {
        stop = !nonneg_finite || t != 0.0f;
        return 0.0;
    }
    if (!nonneg_finite) {
        stop = true;
        return 0.0;
    }
    // A power-of-two scaling: exact (a float times at most 2^1022 can overflow
    // only when S < 2^-894, which a sum of floats never is; an infinite q then
    // gives an infinite r, which ends the run).
    const double q = double(t) * b.inv_u;
    const double f = __builtin_floor(q);
    const double d = q - f;            // exact
    if (d == 0.5) {
        stop = true;
        return 0.0;
    }
    stop = false;
    return (d > 0.5 ? f + 1.0 : f) * b.u;  // exact
}

// `prefix`: the computed sum of r over the run up to and including this term.
// The run ends at the first term that stops it or whose prefix reaches room.
__host__ __device__ __forceinline__ bool ends_run(bool stop, double prefix, const Binade& b)
{
    return stop || prefix >= b.room;
}

}  // namespace oceanamd

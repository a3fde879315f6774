// An exact shortcut through a sequential double-precision sum of floats.
//
// The interaction field subtracts from each stamp its mean, which the
// reference computes as one sequential double sum over the stamp's clipped box
// in z-major order (src/interaction.cpp:345-356): s = 0; s += (double)v_i.
// Its rounding depends on that order, so it cannot be reassociated, and done
// term by term it is a chain of dependent double additions (up to 161 x 161
// of them for a 2 m source at 0.1 m cells). But where no addition of the
// chain rounds, the sequential value is simply the exact sum, which any order
// gives. For a chunk of consecutive terms t_1..t_n and the running value S:
//
//   - every nonzero t_i is a multiple of 2^q (q: the lowest set bit over the
//     chunk's terms), and B >= sum |t_i|;
//   - S is a multiple of 2^lowbit(S);
//   - with g = 2^min(q, lowbit(S)), every partial sum S + t_1 + ... + t_j is a
//     multiple of g of magnitude <= |S| + B. If |S| + B < g * 2^53 each of
//     them is a double, so no addition of the chain rounds and the chain ends
//     at S + sum t_i exactly;
//   - if also B < 2^(q + 53), every partial sum of the chunk's own terms, in
//     any order, is exact as well, so a tree reduction gives sum t_i.
//
// Chunks for which either test fails are added term by term, in order, so
// the result is the sequential sum bit for bit in every case. Shared by the
// device kernel (iwave_kernels.hip) and its CPU emulation test.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace oceanamd {

constexpr int kNoBits = 1 << 20;  // lowbit of zero: above every real exponent

// Exponent of the lowest set bit of a finite nonzero float (-149 .. 104).
__host__ __device__ __forceinline__ int lowbit_exp(float t)
{
    const uint32_t b = __builtin_bit_cast(uint32_t, t) & 0x7FFFFFFFu;
    if (b == 0u) return kNoBits;
    const uint32_t e = b >> 23;
    const uint32_t m = (b & 0x007FFFFFu) | (e != 0u ? 0x00800000u : 0u);
    return int(e == 0u ? 1u : e) - 150 + __builtin_ctz(m);
}

// Exponent of the lowest set bit of a finite nonzero double (-1074 .. 971).
__host__ __device__ __forceinline__ int lowbit_exp(double s)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, s) & 0x7FFFFFFFFFFFFFFFull;
    if (b == 0u) return kNoBits;
    const uint64_t e = b >> 52;
    const uint64_t m = (b & 0x000FFFFFFFFFFFFFull) | (e != 0u ? 0x0010000000000000ull : 0u);
    return int(e == 0u ? 1u : e) - 1075 + __builtin_ctzll(m);
}

__host__ __device__ __forceinline__ double pow2(int e)
{
    // 2^e for -1074 <= e <= 1023, exactly
    if (e >= -1022) return __builtin_bit_cast(double, uint64_t(e + 1023) << 52);
    return __builtin_bit_cast(double, 1ull << (e + 1074));
}

struct ChunkStats {
    double total;  // sum of the chunk's terms; exact when `exact`
    double bound;  // >= sum |t_i|
    int q;         // lowest set bit over the nonzero terms (kNoBits if none)
    int exact;     // every term finite and bound < 2^(q + 53)
};

// Folds one term into a chunk's partial statistics (any order).
struct ChunkAcc {
    double total = 0.0, abs_sum = 0.0;
    int q = kNoBits;
    bool finite = true;

    __host__ __device__ void add(float t)
    {
        const uint32_t b = __builtin_bit_cast(uint32_t, t);
        finite = finite && (b & 0x7F800000u) != 0x7F800000u;
        total += double(t);
        abs_sum += double(t < 0.0f ? -t : t);
        const int l = lowbit_exp(t);
        q = l < q ? l : q;
    }
    __host__ __device__ void merge(double t, double a, int l, bool f)
    {
        total += t;
        abs_sum += a;
        q = l < q ? l : q;
        finite = finite && f;
    }
    __host__ __device__ ChunkStats finish() const
    {
        ChunkStats c;
        c.total = total;
        // covers the rounding of abs_sum (at most 2^-45 relative for 256 terms)
        c.bound = abs_sum * (1.0 + 0x1p-30);
        c.q = q;
        c.exact = finite && (q >= kNoBits || (q + 53 <= 1023 && c.bound < pow2(q + 53))) ? 1 : 0;
        return c;
    }
};

// True when the chain of additions S + t_1 + ... over the chunk is exact, so
// that it ends at S + c.total.
__host__ __device__ __forceinline__ bool chunk_is_exact(double s, const ChunkStats& c)
{
    if (!c.exact) return false;
    const int ls = lowbit_exp(s);
    const int g = c.q < ls ? c.q : ls;
    if (g >= kNoBits) return true;  // S == 0 and every term zero
    if (g + 53 > 1023) return false;
    const double lhs = ((s < 0.0 ? -s : s) + c.bound) * (1.0 + 0x1p-40);
    return lhs < pow2(g + 53);
}

}  // namespace oceanamd

// Radix-2 Stockham inverse FFT executed by a "team" of lanes, 8 complex
// points per lane, in registers, with LDS exchanges between passes.
//
// The butterfly graph is exactly the reference's (src/core/fft.cpp:36-82 and
// src/core/fft_kernel_scalar.cpp:5-35): at stage t (s = 2^t) element j < N/2
// pairs with j + N/2, the sum goes to ((j>>t)<<(t+1)) | (j & (s-1)) and the
// difference, rotated by w[j & ~(s-1)], to that index | s. Every butterfly
// is computed as (a+b, (a-b)*w) with the complex product formed as
// (dr*wr - di*wi, dr*wi + di*wr) and no FMA, so each output is bit-identical
// to the reference's, including the trivial twiddle w = (1, 0).
//
// What changes is where the data lives. A lane holds registers a = 0..7 at
// indices J | a*T (T = N/8 lanes per sequence), i.e. the 8 points that differ
// only in the top 3 index bits. Those are exactly the points the next three
// stages combine with each other, so three stages run in registers before the
// team swaps data through LDS (log2 N stages = ceil(log2 N / 3) passes).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace oceanamd {

constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v >> 1); }

template <int N>
struct FftShape {
    static constexpr int R = 8;                 // points per lane
    static constexpr int T = N / R;             // lanes per sequence
    static constexpr int LOGN = ilog2(N);
    static constexpr int PASSES = (LOGN + 2) / 3;
    static_assert(N >= 16 && (N & (N - 1)) == 0, "N must be a power of two >= 16");
};

// LDS addressing with one float2 of padding per 16, to spread the power-of-two
// strides the passes write with over more banks.
__host__ __device__ __forceinline__ uint32_t lds_pad(uint32_t i) { return i + (i >> 4); }
template <int N>
constexpr int lds_span() { return N + N / 16; }
// Distance between consecutive teams' exchange areas: lds_span rounded up to
// 2 mod 16 float2 (4 mod 32 banks), so the teams a wavefront's 16-lane group
// spans (column teams are interleaved lane by lane) hit different banks.
template <int N>
constexpr int team_stride() { return lds_span<N>() + ((2 - lds_span<N>() % 16) + 16) % 16; }

// r butterfly stages (r <= 3) starting at stage t0, on a lane's 8 registers
// of each of K independent sequences (same indices and twiddles for all K).
// idx[a] tracks the reference-array index register a currently represents.
// FMA (the opt-in fast mode only): the rotation's products are fused,
// (dr*wr - di*wi) as fma(dr, wr, -(di*wi)), which is no longer the reference's
// rounding.
template <int T0, int RST, int K, bool FMA = false>
__host__ __device__ __forceinline__ void fft_pass(float (&re)[K][8], float (&im)[K][8],
                                                  uint32_t (&idx)[8], const float2* __restrict__ tw)
{
#pragma unroll
    for (int st = 0; st < RST; ++st) {
        const int t = T0 + st;
        const int m = 4 >> st;
        const uint32_t s = 1u << t;
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            if (a & m) continue;
            const int b = a | m;
            const uint32_t jl = idx[a];
            const float2 w = tw[jl & ~(s - 1u)];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const float ar = re[k][a], ai = im[k][a];
                const float br = re[k][b], bi = im[k][b];
                re[k][a] = ar + br;
                im[k][a] = ai + bi;
                const float dr = ar - br;
                const float di = ai - bi;
                if constexpr (FMA) {
                    re[k][b] = __builtin_fmaf(dr, w.x, -(di * w.y));
                    im[k][b] = __builtin_fmaf(dr, w.y, di * w.x);
                } else {
                    re[k][b] = dr * w.x - di * w.y;
                    im[k][b] = dr * w.y + di * w.x;
                }
            }
            const uint32_t base = ((jl >> t) << (t + 1)) | (jl & (s - 1u));
            idx[a] = base;
            idx[b] = base | s;
        }
    }
}

template <int N, int P>
struct PassPlan {
    static constexpr int T0 = 3 * P;
    static constexpr int RST = (FftShape<N>::LOGN - T0) < 3 ? (FftShape<N>::LOGN - T0) : 3;
};

// One LDS exchange between passes: every register goes to its current index
// and comes back as element J + a*T, for all K sequences between the same
// two barriers (sequence k uses the area xch + k * kstride).
template <int N, int K, typename Sync>
__host__ __device__ __forceinline__ void exchange(float (&re)[K][8], float (&im)[K][8],
                                                  uint32_t (&idx)[8], uint32_t J,
                                                  float2* __restrict__ xch, uint32_t kstride,
                                                  Sync sync)
{
    constexpr int T = FftShape<N>::T;
    sync();
#pragma unroll
    for (int k = 0; k < K; ++k) {
#pragma unroll
        for (int a = 0; a < 8; ++a) xch[k * kstride + lds_pad(idx[a])] = make_float2(re[k][a], im[k][a]);
    }
    sync();
#pragma unroll
    for (int k = 0; k < K; ++k) {
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const float2 v = xch[k * kstride + lds_pad(J + a * T)];
            re[k][a] = v.x;
            im[k][a] = v.y;
        }
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) idx[a] = J | (static_cast<uint32_t>(a) * T);
}

// Full inverse transform of K complex sequences held by a team.
//   re/im   : in: register a = element J + a*T; out: element idx[a]
//   xch     : this team's LDS exchange areas, K of them kstride float2 apart
//             (each lds_span<N>() float2; teams of a block are
//             team_stride<N>() * K apart)
//   tw      : LDS twiddles, N/2 entries (cos, +sin) of 2*pi*j/N
//   sync    : barrier callable (a block barrier in practice)
// Transforming K sequences together shares the index arithmetic and twiddle
// reads, and gives each barrier K times the independent work.
template <int N, int K, typename Sync, bool FMA = false>
__host__ __device__ __forceinline__ void fft_team_k(float (&re)[K][8], float (&im)[K][8],
                                                    uint32_t (&idx)[8], uint32_t J,
                                                    float2* __restrict__ xch, uint32_t kstride,
                                                    const float2* __restrict__ tw, Sync sync)
{
    constexpr int T = FftShape<N>::T;
    constexpr int PASSES = FftShape<N>::PASSES;
#pragma unroll
    for (int a = 0; a < 8; ++a) idx[a] = J | (static_cast<uint32_t>(a) * T);

    fft_pass<PassPlan<N, 0>::T0, PassPlan<N, 0>::RST, K, FMA>(re, im, idx, tw);
    if constexpr (PASSES > 1) {
        exchange<N, K>(re, im, idx, J, xch, kstride, sync);
        fft_pass<PassPlan<N, 1>::T0, PassPlan<N, 1>::RST, K, FMA>(re, im, idx, tw);
    }
    if constexpr (PASSES > 2) {
        exchange<N, K>(re, im, idx, J, xch, kstride, sync);
        fft_pass<PassPlan<N, 2>::T0, PassPlan<N, 2>::RST, K, FMA>(re, im, idx, tw);
    }
    if constexpr (PASSES > 3) {
        exchange<N, K>(re, im, idx, J, xch, kstride, sync);
        fft_pass<PassPlan<N, 3>::T0, PassPlan<N, 3>::RST, K, FMA>(re, im, idx, tw);
    }
    static_assert(PASSES <= 4, "N > 4096 is outside the reference's range");
}

// One sequence (K = 1).
template <int N, typename Sync, bool FMA = false>
__host__ __device__ __forceinline__ void fft_team(float (&re)[8], float (&im)[8], uint32_t (&idx)[8],
                                                  uint32_t J, float2* __restrict__ xch,
                                                  const float2* __restrict__ tw, Sync sync)
{
    float r1[1][8], i1[1][8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        r1[0][a] = re[a];
        i1[0][a] = im[a];
    }
    fft_team_k<N, 1, Sync, FMA>(r1, i1, idx, J, xch, 0u, tw, sync);
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        re[a] = r1[0][a];
        im[a] = i1[0][a];
    }
}

}  // namespace oceanamd

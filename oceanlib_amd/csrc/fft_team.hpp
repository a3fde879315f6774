// Radix-2 Stockham inverse FFT executed by a "team" of lanes, R complex
// points per lane (R = 4, 8 or 16), in registers, with LDS exchanges between
// passes.
//
// The butterfly graph is exactly the reference's (src/core/fft.cpp:36-82 and
// src/core/fft_kernel_scalar.cpp:5-35): at stage t (s = 2^t) element j < N/2
// pairs with j + N/2, the sum goes to ((j>>t)<<(t+1)) | (j & (s-1)) and the
// difference, rotated by w[j & ~(s-1)], to that index | s. Every butterfly
// is computed as (a+b, (a-b)*w) with the complex product formed as
// (dr*wr - di*wi, dr*wi + di*wr) and no FMA, so each output is bit-identical
// to the reference's, including the trivial twiddle w = (1, 0).
//
// What changes is where the data lives. A lane holds registers a = 0..R-1 at
// indices J | a*T (T = N/R lanes per sequence), i.e. the R points that differ
// only in the top log2 R index bits. Those are exactly the points the next
// log2 R stages combine with each other, so that many stages run in registers
// before the team swaps data through LDS (log2 N stages = ceil(log2 N /
// log2 R) passes). R is the radix grouping: R = 8 groups stages three at a
// time; R = 4 groups them in pairs over twice the lanes (more wavefronts, one
// or two more exchanges); R = 16 four at a time over half the lanes.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace oceanamd {

constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v >> 1); }

template <int N, int R_ = 8>
struct FftShape {
    static constexpr int R = R_;                // points per lane
    static constexpr int T = N / R;             // lanes per sequence
    static constexpr int LOGN = ilog2(N);
    static constexpr int LR = ilog2(R);         // stages per pass
    static constexpr int PASSES = (LOGN + LR - 1) / LR;
    static_assert(N >= 16 && (N & (N - 1)) == 0, "N must be a power of two >= 16");
    static_assert(R == 4 || R == 8 || R == 16, "4, 8 or 16 points per lane");
    static_assert(N >= 2 * R, "at least two lanes per sequence");
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

// r butterfly stages (r <= log2 R) starting at stage t0, on a lane's R
// registers of each of K independent sequences (same indices and twiddles for
// all K). At the pass's st-th stage, register a pairs with a + (R/2 >> st).
// idx[a] tracks the reference-array index register a currently represents.
// FMA (the opt-in fast mode only): the rotation's products are fused,
// (dr*wr - di*wi) as fma(dr, wr, -(di*wi)), which is no longer the reference's
// rounding.
template <int T0, int RST, int K, bool FMA = false, int R = 8>
__host__ __device__ __forceinline__ void fft_pass(float (&re)[K][R], float (&im)[K][R],
                                                  uint32_t (&idx)[R], const float2* __restrict__ tw)
{
#pragma unroll
    for (int st = 0; st < RST; ++st) {
        const int t = T0 + st;
        const int m = (R / 2) >> st;
        const uint32_t s = 1u << t;
#pragma unroll
        for (int a = 0; a < R; ++a) {
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

template <int N, int P, int R = 8>
struct PassPlan {
    static constexpr int LR = FftShape<N, R>::LR;
    static constexpr int T0 = LR * P;
    static constexpr int RST = (FftShape<N, R>::LOGN - T0) < LR ? (FftShape<N, R>::LOGN - T0) : LR;
};

// One LDS exchange between passes: every register goes to its current index
// and comes back as element J + a*T, for all K sequences between the same
// two barriers (sequence k uses the area xch + k * kstride).
template <int N, int K, typename Sync, int R = 8>
__host__ __device__ __forceinline__ void exchange(float (&re)[K][R], float (&im)[K][R],
                                                  uint32_t (&idx)[R], uint32_t J,
                                                  float2* __restrict__ xch, uint32_t kstride,
                                                  Sync sync)
{
    constexpr int T = FftShape<N, R>::T;
    sync();
#pragma unroll
    for (int k = 0; k < K; ++k) {
#pragma unroll
        for (int a = 0; a < R; ++a) xch[k * kstride + lds_pad(idx[a])] = make_float2(re[k][a], im[k][a]);
    }
    sync();
#pragma unroll
    for (int k = 0; k < K; ++k) {
#pragma unroll
        for (int a = 0; a < R; ++a) {
            const float2 v = xch[k * kstride + lds_pad(J + a * T)];
            re[k][a] = v.x;
            im[k][a] = v.y;
        }
    }
#pragma unroll
    for (int a = 0; a < R; ++a) idx[a] = J | (static_cast<uint32_t>(a) * T);
}

// Passes P .. PASSES-1, each after an exchange.
template <int N, int K, typename Sync, bool FMA, int R, int P>
__host__ __device__ __forceinline__ void later_passes(float (&re)[K][R], float (&im)[K][R],
                                                      uint32_t (&idx)[R], uint32_t J,
                                                      float2* __restrict__ xch, uint32_t kstride,
                                                      const float2* __restrict__ tw, Sync sync)
{
    if constexpr (P < FftShape<N, R>::PASSES) {
        exchange<N, K, Sync, R>(re, im, idx, J, xch, kstride, sync);
        fft_pass<PassPlan<N, P, R>::T0, PassPlan<N, P, R>::RST, K, FMA, R>(re, im, idx, tw);
        later_passes<N, K, Sync, FMA, R, P + 1>(re, im, idx, J, xch, kstride, tw, sync);
    }
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
template <int N, int K, typename Sync, bool FMA = false, int R = 8>
__host__ __device__ __forceinline__ void fft_team_k(float (&re)[K][R], float (&im)[K][R],
                                                    uint32_t (&idx)[R], uint32_t J,
                                                    float2* __restrict__ xch, uint32_t kstride,
                                                    const float2* __restrict__ tw, Sync sync)
{
    constexpr int T = FftShape<N, R>::T;
#pragma unroll
    for (int a = 0; a < R; ++a) idx[a] = J | (static_cast<uint32_t>(a) * T);
    constexpr int PASSES = FftShape<N, R>::PASSES;
    fft_pass<PassPlan<N, 0, R>::T0, PassPlan<N, 0, R>::RST, K, FMA, R>(re, im, idx, tw);
    if constexpr (PASSES > 1) {
        exchange<N, K, Sync, R>(re, im, idx, J, xch, kstride, sync);
        fft_pass<PassPlan<N, 1, R>::T0, PassPlan<N, 1, R>::RST, K, FMA, R>(re, im, idx, tw);
    }
    if constexpr (PASSES > 2) {
        exchange<N, K, Sync, R>(re, im, idx, J, xch, kstride, sync);
        fft_pass<PassPlan<N, 2, R>::T0, PassPlan<N, 2, R>::RST, K, FMA, R>(re, im, idx, tw);
    }
    if constexpr (PASSES > 3) {
        exchange<N, K, Sync, R>(re, im, idx, J, xch, kstride, sync);
        fft_pass<PassPlan<N, 3, R>::T0, PassPlan<N, 3, R>::RST, K, FMA, R>(re, im, idx, tw);
    }
    later_passes<N, K, Sync, FMA, R, 4>(re, im, idx, J, xch, kstride, tw, sync);
}

// One sequence (K = 1).
template <int N, typename Sync, bool FMA = false, int R = 8>
__host__ __device__ __forceinline__ void fft_team(float (&re)[R], float (&im)[R], uint32_t (&idx)[R],
                                                  uint32_t J, float2* __restrict__ xch,
                                                  const float2* __restrict__ tw, Sync sync)
{
    float r1[1][R], i1[1][R];
#pragma unroll
    for (int a = 0; a < R; ++a) {
        r1[0][a] = re[a];
        i1[0][a] = im[a];
    }
    fft_team_k<N, 1, Sync, FMA, R>(r1, i1, idx, J, xch, 0u, tw, sync);
#pragma unroll
    for (int a = 0; a < R; ++a) {
        re[a] = r1[0][a];
        im[a] = i1[0][a];
    }
}

}  // namespace oceanamd

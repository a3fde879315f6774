// The two passes of one frame, as block-level device functions shared by the
// single-ocean kernels and the batched persistent kernels (frame_kernels.hip).
//
//   row_block : evolve h(k,t) for TEAMS rows of cells straight into registers
//               (reference src/core/evolve.cpp:28-184) and run each row's
//               inverse FFT for all F complex fields (src/ocean.cpp:193-208);
//               writes the fields row-major to the level's `spec` scratch.
//   col_block : column inverse FFT of C columns for all F fields
//               (src/ocean.cpp:214-233) and the finalize pack
//               (src/core/evolve.cpp:186-327) from registers into the RGBA32F
//               displacement / normal (/ velocity) buffers.
//
// The block must have exactly TEAMS*N/R (resp. C*N/R) threads, R the points
// per lane of the FFT (fft_team.hpp; 8 unless a shape picks 4 or 16); all of
// them reach every barrier. Work on `tw` and `xch` (LDS) starts with a barrier, so
// a persistent block can call these back to back for different levels.
#pragma once

#include "device_math.hpp"
#include "fft_team.hpp"
#include "hip_launch.h"

#include <hip/hip_runtime.h>

#include <type_traits>

namespace oceanamd {

// Streaming store of the row pass's fields: whole rows, read next by the
// column pass mostly on other XCDs, so keeping the lines in this XCD's L2
// only displaces the constant tables the next frame reads again.
// A team's lanes write T consecutive float2 per register: below T = 16
// (N = 64 and smaller) a store covers part of a 128-byte line, and those are
// left to the L2 to merge.
typedef float nt_f32x2 __attribute__((ext_vector_type(2)));
typedef float nt_f32x4 __attribute__((ext_vector_type(4)));
template <int N>
__device__ __forceinline__ void store_stream(float2* p, float x, float y)
{
    if constexpr (FftShape<N>::T * 8 >= 128)
        __builtin_nontemporal_store(nt_f32x2{x, y}, reinterpret_cast<nt_f32x2*>(p));
    else
        *p = make_float2(x, y);
}

// The column pass's RGBA32F outputs: streamed in the batched passes, stored
// normally in the single ocean's. A batch's outputs are read later, if at
// all, and their lines in the writing XCD's L2 only displaced the row pass's
// tables (measured A/B per group: column pass -9% to -22% at N <= 256). A
// single ocean's are gathered right away by the foam step and the queries,
// which then found them further away: streamed, its column pass was 1.5 us
// faster and the whole frame with the foam step and queries in one launch
// 1.6% slower. Where a store covers only part of a 128-byte line (4 columns
// per item at N = 512) the streamed lines are written as partial lines
// (WRITE_SIZE 1.27x); there only the normal is streamed, which measured
// faster than both (-3%) or neither (-3.5%).
template <bool STREAM>
__device__ __forceinline__ void store_out(float4* p, float x, float y, float z, float w)
{
    if constexpr (STREAM)
        __builtin_nontemporal_store(nt_f32x4{x, y, z, w}, reinterpret_cast<nt_f32x4*>(p));
    else
        *p = make_float4(x, y, z, w);
}

struct BlockSync {
    __device__ void operator()() const { __syncthreads(); }
};

// A team whose lanes all lie in one wavefront needs no block barrier around
// its LDS exchanges: a wavefront's LDS operations execute in program order,
// so it is enough that the compiler keeps them in order. Row teams are T = N/8
// consecutive lanes, inside one wavefront for N <= 512; column teams are
// interleaved with the block's other columns (column-fastest lanes), inside
// one wavefront only when the whole C * T block share is.
struct WaveSync {
    __device__ void operator()() const
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
};
template <bool IN_WAVE>
using TeamSync = typename std::conditional<IN_WAVE, WaveSync, BlockSync>::type;

// Precision modes of the two passes (template argument MODE).
//   kModeExact    : the reference's arithmetic, bit for bit (the default).
//   kModeFast     : the opt-in fast mode: the hardware's sine and cosine of
//                   the phase in revolutions, fused butterfly rotations and
//                   a hardware reciprocal square root for the normal.
//   kModeFastHalf : kModeFast with the row pass's fields stored as fp16
//                   pairs, halving the bytes between the passes.
constexpr int kModeExact = 0;
constexpr int kModeFast = 1;
constexpr int kModeFastHalf = 2;

typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

// Element (z, x) of field f in the level's `spec` scratch, in the layout W:
// rows of N elements (W = 0), or tiles of W columns x N rows whose rows are W
// elements long (W > 0), so that a column team's loads of W adjacent columns
// over consecutive rows are one contiguous run instead of one piece per row.
template <int N, int W>
__device__ __forceinline__ size_t spec_at(uint32_t f, uint32_t z, uint32_t x)
{
    constexpr size_t CELLS = size_t(N) * N;
    if constexpr (W == 0)
        return size_t(f) * CELLS + size_t(z) * N + x;
    else
        return size_t(f) * CELLS + size_t(x / W) * (size_t(N) * W) + size_t(z) * W + (x % W);
}

// One complex value of field f at cell `cell` (element offset within the
// level's `spec`), in the mode's storage: float2, or fp16 pairs. Row-major
// rows are streamed out (whole lines); tiled rows are pieces of W elements,
// left to the L2 to merge.
template <int N, int MODE, int W = 0>
__device__ __forceinline__ void store_field(void* spec, size_t at, float x, float y)
{
    if constexpr (MODE == kModeFastHalf) {
        reinterpret_cast<f16x2*>(spec)[at] = f16x2{static_cast<_Float16>(x), static_cast<_Float16>(y)};
    } else if constexpr (W != 0) {
        reinterpret_cast<float2*>(spec)[at] = make_float2(x, y);
    } else {
        store_stream<N>(reinterpret_cast<float2*>(spec) + at, x, y);
    }
}

// A single ocean's column pass reads each field element once, written by the
// row pass on other XCDs: with NT a streaming load, so the fields do not
// displace the row pass's constant tables (h0, omega, 1/|k|: 4 MB at 512^2,
// an eighth of it per XCD) from the XCD's L2 between frames (in the frame
// loop the row pass then runs at its back-to-back time: 10.9 -> 8.8 us at
// 512^2). The batched passes load normally: there streaming loads cost the
// whole batch 23% (164k -> 126k patch-updates/s, A/B on one box).
template <int MODE, bool NT = false>
__device__ __forceinline__ float2 load_field(const void* spec, size_t at)
{
    if constexpr (MODE == kModeFastHalf) {
        const f16x2 h = reinterpret_cast<const f16x2*>(spec)[at];
        return make_float2(static_cast<float>(h.x), static_cast<float>(h.y));
    } else if constexpr (NT) {
        const nt_f32x2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_f32x2*>(spec) + at);
        return make_float2(v.x, v.y);
    } else {
        return reinterpret_cast<const float2*>(spec)[at];
    }
}

// The phase's sine and cosine in the mode's arithmetic.
template <int MODE>
__device__ __forceinline__ void phase_sincos(float w, double time_over_two_pi, float& s, float& c)
{
    if constexpr (MODE == kModeExact) {
        const float phase = reduced_phase(w, time_over_two_pi);
        sincos_poly(phase, s, c);
    } else {
        sincos_fast(w, time_over_two_pi, s, c);
    }
}

// Team g handles row first(q) or second(q) of mirror pair q, so a block's
// teams read each other's h0 rows as their mirrors.
template <int N>
__device__ __forceinline__ uint32_t row_of_team(uint32_t g)
{
    const uint32_t q = g >> 1;
    if ((g & 1u) == 0u) return q;
    return q == 0u ? uint32_t(N / 2) : uint32_t(N) - q;
}

template <int N, int F, int TEAMS, int MODE = kModeExact, int R = 8, int W = 0>
__device__ __forceinline__ void row_block(const ocean_hip_level& L, uint32_t blk,
                                          double time_over_two_pi, float2* __restrict__ tw,
                                          float2* __restrict__ xch, uint32_t tid = threadIdx.x)
{
    constexpr int T = FftShape<N, R>::T;
    constexpr int BLOCK = TEAMS * T;
    const float2* __restrict__ h0 = reinterpret_cast<const float2*>(L.h0);
    const float* __restrict__ omega = L.omega;
    const float* __restrict__ kinv = L.kinv;
    const float* __restrict__ kxs = L.kx;
    const float2* __restrict__ twg = reinterpret_cast<const float2*>(L.twiddles);
    void* __restrict__ spec = L.spec;
    const float patch_length = L.patch_length;

    __syncthreads();
    for (int i = tid; i < N / 2; i += BLOCK) tw[i] = twg[i];
    __syncthreads();

    const uint32_t team = tid / T;
    const uint32_t J = tid % T;
    const uint32_t z = row_of_team<N>(blk * TEAMS + team);
    const uint32_t mz = (uint32_t(N) - z) & uint32_t(N - 1);

    float fr[F][R], fi[F][R];
    const float kz = wave_component(z, N, patch_length);

#pragma unroll
    for (int a = 0; a < R; ++a) {
        const uint32_t x = J + uint32_t(a) * T;
        const uint32_t mx = (uint32_t(N) - x) & uint32_t(N - 1);
        const uint32_t i = z * N + x;
        const float w = omega[i];
        const float2 hv = h0[i];
        const float2 hm = h0[mz * N + mx];

        // reference src/core/evolve.cpp:45-121 (h0c = conj(h0(-k)), spectrum.cpp:329-338)
        float s, c;
        phase_sincos<MODE>(w, time_over_two_pi, s, c);
        const float a_re = hv.x, a_im = hv.y;
        const float b_re = hm.x, b_im = -hm.y;
        const float fwd_re = a_re * c - a_im * s;
        const float fwd_im = a_re * s + a_im * c;
        const float bwd_re = b_re * c + b_im * s;
        const float bwd_im = b_im * c - b_re * s;
        const float hr = fwd_re + bwd_re;
        const float hi = fwd_im + bwd_im;

        const float kx = kxs[x];
        const float ki = kinv[i];
        const float sx = kx * ki;
        const float sz = kz * ki;
        const float dx_re = sx * hi, dx_im = -sx * hr;
        const float dz_re = sz * hi, dz_im = -sz * hr;
        const float hx_re = -kx * hi, hx_im = kx * hr;
        const float hz_re = -kz * hi, hz_im = kz * hr;
        const float jxx = kx * sx, jzz = kz * sz, jxz = kx * sz;
        const float jxx_re = jxx * hr, jxx_im = jxx * hi;
        const float jzz_re = jzz * hr, jzz_im = jzz * hi;
        const float jxz_re = jxz * hr, jxz_im = jxz * hi;
        fr[0][a] = hr - dx_im;      fi[0][a] = hi + dx_re;
        fr[1][a] = dz_re - hx_im;   fi[1][a] = dz_im + hx_re;
        fr[2][a] = hz_re - jxx_im;  fi[2][a] = hz_im + jxx_re;
        fr[3][a] = jzz_re - jxz_im; fi[3][a] = jzz_im + jxz_re;
        if constexpr (F == 6) {
            // reference src/core/evolve.cpp:150-182
            const float dr = fwd_re - bwd_re;
            const float di = fwd_im - bwd_im;
            const float wsx = w * sx, wsz = w * sz;
            const float ux_re = -wsx * dr, ux_im = -wsx * di;
            const float uz_re = -wsz * dr, uz_im = -wsz * di;
            const float uy_re = -w * di, uy_im = w * dr;
            fr[4][a] = ux_re - uy_im; fi[4][a] = ux_im + uy_re;
            fr[5][a] = uz_re;         fi[5][a] = uz_im;
        }
    }

    float2* x_team = xch + team * team_stride<N>();
#pragma unroll
    for (int f = 0; f < F; ++f) {
        uint32_t idx[R];
        using Sync = TeamSync<(T <= 64)>;
        fft_team<N, Sync, MODE != kModeExact, R>(fr[f], fi[f], idx, J, x_team, tw, Sync{});
#pragma unroll
        for (int a = 0; a < R; ++a)
            store_field<N, MODE, W>(spec, spec_at<N, W>(uint32_t(f), z, idx[a]), fr[f][a], fi[f][a]);
    }
}

// The finalize of one cell from its eight real fields (h, Dx, Dz, dh/dx,
// dh/dz, dDx/dx, dDz/dz, dDx/dz): reference src/core/evolve.cpp:256-325.
template <bool STREAM_DISP, bool STREAM_NORMAL, int MODE = kModeExact>
__device__ __forceinline__ void pack_cell(float h, float d_x, float d_z, float hx, float hz,
                                          float jxx, float jzz, float jxz, float lambda,
                                          float foam_threshold, float inv_threshold,
                                          float4* __restrict__ disp, float4* __restrict__ normal,
                                          size_t o)
{
    const float disp_x = lambda * d_x;
    const float disp_z = lambda * d_z;
    const float axx = 1.0f + lambda * jxx;
    const float azz = 1.0f + lambda * jzz;
    const float axz = lambda * jxz;
    const float jac = axx * azz - axz * axz;
    float nx = -(hx * azz - axz * hz);
    float ny = jac;
    float nz = -(axx * hz - hx * axz);
    const float len2 = nx * nx + ny * ny + nz * nz;
    float inv_len;
    if constexpr (MODE == kModeExact)
        inv_len = (len2 > 0.0f) ? 1.0f / sqrtf(len2) : 0.0f;
    else
        inv_len = (len2 > 0.0f) ? __builtin_amdgcn_rsqf(len2) : 0.0f;
    nx *= inv_len;
    ny *= inv_len;
    nz *= inv_len;
    float foam = (foam_threshold - jac) * inv_threshold;
    if (foam < 0.0f) foam = 0.0f;
    if (foam > 1.0f) foam = 1.0f;
    store_out<STREAM_DISP>(disp + o, disp_x, h, disp_z, foam);
    store_out<STREAM_NORMAL>(normal + o, nx, ny, nz, jac);
}

template <int N, int F, int C, bool STREAM = false, int MODE = kModeExact, int R = 8, int W = 0>
__device__ __forceinline__ void col_block(const ocean_hip_level& L, uint32_t blk,
                                          float2* __restrict__ tw, float2* __restrict__ xch,
                                          uint32_t tid = threadIdx.x)
{
    constexpr int T = FftShape<N, R>::T;
    constexpr int BLOCK = C * T;
    const void* __restrict__ spec = L.spec;
    const float2* __restrict__ twg = reinterpret_cast<const float2*>(L.twiddles);
    float4* __restrict__ disp = reinterpret_cast<float4*>(L.disp);
    float4* __restrict__ normal = reinterpret_cast<float4*>(L.normal);
    float4* __restrict__ vel = reinterpret_cast<float4*>(L.vel);
    const float lambda = L.choppiness;
    const float foam_threshold = L.foam_threshold;
    const float inv_threshold = L.inv_foam_threshold;

    __syncthreads();
    for (int i = tid; i < N / 2; i += BLOCK) tw[i] = twg[i];
    __syncthreads();

    const uint32_t c = tid % C;
    const uint32_t J = tid / C;
    const uint32_t x = blk * C + c;

    float fr[F][R], fi[F][R];
#pragma unroll
    for (int f = 0; f < F; ++f) {
#pragma unroll
        for (int a = 0; a < R; ++a) {
            const float2 v = load_field<MODE, !STREAM>(spec, spec_at<N, W>(uint32_t(f), J + uint32_t(a * T), x));
            fr[f][a] = v.x;
            fi[f][a] = v.y;
        }
    }

    uint32_t idx[R];
#pragma unroll
    for (int f = 0; f < F; ++f) {
        using Sync = TeamSync<(C * T <= 64)>;
        fft_team<N, Sync, MODE != kModeExact, R>(fr[f], fi[f], idx, J, xch + c * team_stride<N>(), tw,
                                              Sync{});
    }

#pragma unroll
    for (int a = 0; a < R; ++a) {
        const size_t o = size_t(idx[a]) * N + x;
        pack_cell<(STREAM && C * 16 >= 128), STREAM, MODE>(fr[0][a], fi[0][a], fr[1][a], fi[1][a],
                                                     fr[2][a], fi[2][a], fr[3][a], fi[3][a], lambda,
                                                     foam_threshold, inv_threshold, disp, normal, o);
        if constexpr (F == 6) {
            // reference src/core/evolve.cpp:186-203
            store_out<STREAM>(vel + o, fr[4][a], fi[4][a], fr[5][a], 0.0f);
        }
    }
}

// ---- field-split blocks ------------------------------------------------------
// The same passes with the F fields spread over G = F / 2 teams per row (or
// column): team group g transforms fields 2g and 2g + 1, so a frame has G
// times the wavefronts, each with a third (F = 6) or half (F = 4) of the
// serial FFT work. Every field is computed by the same expressions as above.
// The column finalize needs fields 0-3 of a cell together: groups 0 and 1
// swap half of their fields through LDS (`xfer`) and each packs half of the
// cells; group 2 (velocity) writes its output itself. Requires T = N/8 >= 64, so a team is whole wavefronts and the
// group branches are uniform (T = N/R). KX = 2 transforms a team's two fields together
// (fft_team_k: shared index arithmetic and twiddle reads, one pair of
// barriers per exchange for both), KX = 1 one after the other.

template <int N, int F, int TEAMS, int KX, int MODE = kModeExact, int R = 8, int W = 0>
__device__ __forceinline__ void row_block_split(const ocean_hip_level& L, uint32_t blk,
                                                double time_over_two_pi, float2* __restrict__ tw,
                                                float2* __restrict__ xch)
{
    constexpr int T = FftShape<N, R>::T;
    constexpr int G = F / 2;
    constexpr int BLOCK = TEAMS * G * T;
    static_assert(T >= 64 && F % 2 == 0, "field split needs whole-wavefront teams");
    const float2* __restrict__ h0 = reinterpret_cast<const float2*>(L.h0);
    const float* __restrict__ omega = L.omega;
    const float* __restrict__ kinv = L.kinv;
    const float* __restrict__ kxs = L.kx;
    const float2* __restrict__ twg = reinterpret_cast<const float2*>(L.twiddles);
    void* __restrict__ spec = L.spec;
    const float patch_length = L.patch_length;

    __syncthreads();
    for (int i = threadIdx.x; i < N / 2; i += BLOCK) tw[i] = twg[i];
    __syncthreads();

    const uint32_t tt = threadIdx.x / T;
    const uint32_t g = tt % G;
    const uint32_t rteam = tt / G;
    const uint32_t J = threadIdx.x % T;
    const uint32_t z = row_of_team<N>(blk * TEAMS + rteam);
    const uint32_t mz = (uint32_t(N) - z) & uint32_t(N - 1);

    float fr[2][R], fi[2][R];
    const float kz = wave_component(z, N, patch_length);

#pragma unroll
    for (int a = 0; a < R; ++a) {
        const uint32_t x = J + uint32_t(a) * T;
        const uint32_t mx = (uint32_t(N) - x) & uint32_t(N - 1);
        const uint32_t i = z * N + x;
        const float w = omega[i];
        const float2 hv = h0[i];
        const float2 hm = h0[mz * N + mx];

        // reference src/core/evolve.cpp:45-121 (h0c = conj(h0(-k)), spectrum.cpp:329-338)
        float s, c;
        phase_sincos<MODE>(w, time_over_two_pi, s, c);
        const float a_re = hv.x, a_im = hv.y;
        const float b_re = hm.x, b_im = -hm.y;
        const float fwd_re = a_re * c - a_im * s;
        const float fwd_im = a_re * s + a_im * c;
        const float bwd_re = b_re * c + b_im * s;
        const float bwd_im = b_im * c - b_re * s;
        const float hr = fwd_re + bwd_re;
        const float hi = fwd_im + bwd_im;

        const float kx = kxs[x];
        const float ki = kinv[i];
        const float sx = kx * ki;
        const float sz = kz * ki;
        if (g == 0u) {
            const float dx_re = sx * hi, dx_im = -sx * hr;
            const float dz_re = sz * hi, dz_im = -sz * hr;
            const float hx_re = -kx * hi, hx_im = kx * hr;
            fr[0][a] = hr - dx_im;      fi[0][a] = hi + dx_re;
            fr[1][a] = dz_re - hx_im;   fi[1][a] = dz_im + hx_re;
        } else if (g == 1u) {
            const float hz_re = -kz * hi, hz_im = kz * hr;
            const float jxx = kx * sx, jzz = kz * sz, jxz = kx * sz;
            const float jxx_re = jxx * hr, jxx_im = jxx * hi;
            const float jzz_re = jzz * hr, jzz_im = jzz * hi;
            const float jxz_re = jxz * hr, jxz_im = jxz * hi;
            fr[0][a] = hz_re - jxx_im;  fi[0][a] = hz_im + jxx_re;
            fr[1][a] = jzz_re - jxz_im; fi[1][a] = jzz_im + jxz_re;
        } else {
            // reference src/core/evolve.cpp:150-182
            const float dr = fwd_re - bwd_re;
            const float di = fwd_im - bwd_im;
            const float wsx = w * sx, wsz = w * sz;
            const float ux_re = -wsx * dr, ux_im = -wsx * di;
            const float uz_re = -wsz * dr, uz_im = -wsz * di;
            const float uy_re = -w * di, uy_im = w * dr;
            fr[0][a] = ux_re - uy_im; fi[0][a] = ux_im + uy_re;
            fr[1][a] = uz_re;         fi[1][a] = uz_im;
        }
    }

    float2* x_team = xch + tt * KX * team_stride<N>();
    using RSync = TeamSync<(T <= 64)>;
    if constexpr (KX == 2) {
        uint32_t idx[R];
        fft_team_k<N, 2, RSync, MODE != kModeExact, R>(fr, fi, idx, J, x_team, team_stride<N>(), tw,
                                                    RSync{});
#pragma unroll
        for (int f = 0; f < 2; ++f) {
#pragma unroll
            for (int a = 0; a < R; ++a)
                store_field<N, MODE, W>(spec, spec_at<N, W>(2u * g + uint32_t(f), z, idx[a]), fr[f][a],
                                        fi[f][a]);
        }
    } else {
#pragma unroll
        for (int f = 0; f < 2; ++f) {
            uint32_t idx[R];
            fft_team<N, RSync, MODE != kModeExact, R>(fr[f], fi[f], idx, J, x_team, tw, RSync{});
#pragma unroll
            for (int a = 0; a < R; ++a)
                store_field<N, MODE, W>(spec, spec_at<N, W>(2u * g + uint32_t(f), z, idx[a]), fr[f][a],
                                        fi[f][a]);
        }
    }
}

template <int N, int F, int C, int KX, int MODE = kModeExact, int R = 8, int W = 0>
__device__ __forceinline__ void col_block_split(const ocean_hip_level& L, uint32_t blk,
                                                float2* __restrict__ tw, float2* __restrict__ xch,
                                                float2* __restrict__ xfer)
{
    constexpr int T = FftShape<N, R>::T;
    constexpr int G = F / 2;
    constexpr int BLOCK = C * T * G;
    static_assert(T >= 64 && F % 2 == 0, "field split needs whole-wavefront teams");
    const void* __restrict__ spec = L.spec;
    const float2* __restrict__ twg = reinterpret_cast<const float2*>(L.twiddles);
    float4* __restrict__ disp = reinterpret_cast<float4*>(L.disp);
    float4* __restrict__ normal = reinterpret_cast<float4*>(L.normal);
    float4* __restrict__ vel = reinterpret_cast<float4*>(L.vel);
    const float lambda = L.choppiness;
    const float foam_threshold = L.foam_threshold;
    const float inv_threshold = L.inv_foam_threshold;

    __syncthreads();
    for (int i = threadIdx.x; i < N / 2; i += BLOCK) tw[i] = twg[i];
    __syncthreads();

    const uint32_t g = threadIdx.x / uint32_t(C * T);
    const uint32_t l = threadIdx.x % uint32_t(C * T);
    const uint32_t c = l % C;
    const uint32_t J = l / C;
    const uint32_t x = blk * C + c;

    float fr[2][R], fi[2][R];
#pragma unroll
    for (int f = 0; f < 2; ++f) {
#pragma unroll
        for (int a = 0; a < R; ++a) {
            const float2 v =
                load_field<MODE, true>(spec, spec_at<N, W>(2u * g + uint32_t(f), J + uint32_t(a * T), x));
            fr[f][a] = v.x;
            fi[f][a] = v.y;
        }
    }

    uint32_t idx[R];
    using CSync = TeamSync<(C * T <= 64)>;
    if constexpr (KX == 2) {
        fft_team_k<N, 2, CSync, MODE != kModeExact, R>(fr, fi, idx, J,
                                                    xch + (g * C + c) * 2 * team_stride<N>(),
                                                    team_stride<N>(), tw, CSync{});
    } else {
#pragma unroll
        for (int f = 0; f < 2; ++f)
            fft_team<N, CSync, MODE != kModeExact, R>(fr[f], fi[f], idx, J,
                                                   xch + (g * C + c) * team_stride<N>(), tw, CSync{});
    }

    // Groups 0 and 1 share the finalize: group 0 packs the cells of registers
    // a < R/2, group 1 those of a >= R/2. Lanes with the same (c, J) hold the
    // same indices in both groups, and the two halves' cells are disjoint, so
    // one LDS area per column carries both hand-overs: group 1's fields 2, 3
    // of a < R/2 and group 0's fields 0, 1 of a >= R/2.
    float2* xf0 = xfer + (c * 2u + 0u) * N;
    float2* xf1 = xfer + (c * 2u + 1u) * N;
    if (g == 0u) {
#pragma unroll
        for (int a = R / 2; a < R; ++a) {
            xf0[idx[a]] = make_float2(fr[0][a], fi[0][a]);
            xf1[idx[a]] = make_float2(fr[1][a], fi[1][a]);
        }
    } else if (g == 1u) {
#pragma unroll
        for (int a = 0; a < R / 2; ++a) {
            xf0[idx[a]] = make_float2(fr[0][a], fi[0][a]);
            xf1[idx[a]] = make_float2(fr[1][a], fi[1][a]);
        }
    }
    if constexpr (F == 6) {
        if (g == 2u) {
            // reference src/core/evolve.cpp:186-203
#pragma unroll
            for (int a = 0; a < R; ++a)
                store_out<false>(vel + size_t(idx[a]) * N + x, fr[0][a], fi[0][a], fr[1][a], 0.0f);
        }
    }
    __syncthreads();
    if (g == 0u) {
#pragma unroll
        for (int a = 0; a < R / 2; ++a) {
            const float2 v2 = xf0[idx[a]];
            const float2 v3 = xf1[idx[a]];
            pack_cell<false, false, MODE>(fr[0][a], fi[0][a], fr[1][a], fi[1][a], v2.x, v2.y, v3.x,
                                    v3.y, lambda, foam_threshold, inv_threshold, disp, normal,
                                    size_t(idx[a]) * N + x);
        }
    } else if (g == 1u) {
#pragma unroll
        for (int a = R / 2; a < R; ++a) {
            const float2 v0 = xf0[idx[a]];
            const float2 v1 = xf1[idx[a]];
            pack_cell<false, false, MODE>(v0.x, v0.y, v1.x, v1.y, fr[0][a], fi[0][a], fr[1][a],
                                    fi[1][a], lambda, foam_threshold, inv_threshold, disp, normal,
                                    size_t(idx[a]) * N + x);
        }
    }
}

}  // namespace oceanamd

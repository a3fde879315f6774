// Per-frame wave synthesis kernels for gfx950 (MI355X).
//
// Two launches per frame and ocean level replace the reference's four
// barriered CPU stages (src/ocean.cpp:251-268):
//
//   row_pass : evolve h(k,t) for a row of cells straight into registers
//              (src/core/evolve.cpp:28-184), then the row inverse FFT of all
//              F complex fields (src/ocean.cpp:193-208). Writes the F fields
//              row-major to the `spec` scratch in HBM.
//   col_pass : column inverse FFT of the F fields (src/ocean.cpp:214-233),
//              then the finalize pack (src/core/evolve.cpp:186-327) straight
//              from registers into the RGBA32F displacement / normal
//              (/ velocity) buffers.
//
// Only h0 (8 B/cell) and omega (4 B/cell) are read per frame: kx, kz and
// 1/|k| are recomputed from the cell index with the reference's own float /
// double operations, and conj(h0(-k)) is read from the mirrored cell of h0
// instead of a second table.
//
// Plus: the batched height/sample query kernel (src/ocean.cpp:309-452) and
// the foam advection step (src/foam.cpp:22-139).
#include "device_math.hpp"
#include "fft_team.hpp"
#include "hip_launch.h"
#include "sample_device.hpp"

#include <hip/hip_runtime.h>

namespace oceanamd {
namespace {

constexpr int kBlock = 256;

template <int N>
struct RowShape {
    static constexpr int T = FftShape<N>::T;
    static constexpr int B0 = T > kBlock ? T : kBlock;
    static constexpr int BLOCK = B0 < N * T ? B0 : N * T;   // never more teams than rows
    static constexpr int TEAMS = BLOCK / T;
};

template <int N>
struct ColShape {
    static constexpr int T = FftShape<N>::T;
    static constexpr int C0 = (kBlock / T) > 8 ? (kBlock / T) : 8;
    static constexpr int C1 = C0 < N ? C0 : N;
    static constexpr int C2 = (C1 * T) > 512 ? (512 / T) : C1;   // <= 512 lanes: no spills at F = 6
    static constexpr int C = (C2 * N * 8) > 65536 ? (65536 / (N * 8)) : C2;
    static constexpr int BLOCK = C * T;
};

struct BlockSync {
    __device__ void operator()() const { __syncthreads(); }
};

// Team g of the row pass handles row first(q) or second(q) of mirror pair q,
// so a block's teams read each other's h0 rows as their mirrors.
template <int N>
__device__ __forceinline__ uint32_t row_of_team(uint32_t g)
{
    const uint32_t q = g >> 1;
    if ((g & 1u) == 0u) return q;
    return q == 0u ? uint32_t(N / 2) : uint32_t(N) - q;
}

template <int N, int F>
__global__ __launch_bounds__(RowShape<N>::BLOCK) void row_pass_kernel(
    const float2* __restrict__ h0, const float* __restrict__ omega,
    const float2* __restrict__ twg, float2* __restrict__ spec, double time_over_two_pi,
    float patch_length)
{
    constexpr int T = RowShape<N>::T;
    constexpr int TEAMS = RowShape<N>::TEAMS;
    constexpr uint32_t CELLS = uint32_t(N) * N;
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[TEAMS][lds_span<N>()];

    for (int i = threadIdx.x; i < N / 2; i += RowShape<N>::BLOCK) tw[i] = twg[i];
    __syncthreads();

    const uint32_t team = threadIdx.x / T;
    const uint32_t J = threadIdx.x % T;
    const uint32_t z = row_of_team<N>(blockIdx.x * TEAMS + team);
    const uint32_t mz = (uint32_t(N) - z) & uint32_t(N - 1);

    float fr[F][8], fi[F][8];
    const float kz = wave_component(z, N, patch_length);

#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const uint32_t x = J + uint32_t(a) * T;
        const uint32_t mx = (uint32_t(N) - x) & uint32_t(N - 1);
        const uint32_t i = z * N + x;
        const float w = omega[i];
        const float2 hv = h0[i];
        const float2 hm = h0[mz * N + mx];

        // reference src/core/evolve.cpp:45-121 (h0c = conj(h0(-k)), spectrum.cpp:329-338)
        const float phase = reduced_phase(w, time_over_two_pi);
        float s, c;
        sincos_poly(phase, s, c);
        const float a_re = hv.x, a_im = hv.y;
        const float b_re = hm.x, b_im = -hm.y;
        const float fwd_re = a_re * c - a_im * s;
        const float fwd_im = a_re * s + a_im * c;
        const float bwd_re = b_re * c + b_im * s;
        const float bwd_im = b_im * c - b_re * s;
        const float hr = fwd_re + bwd_re;
        const float hi = fwd_im + bwd_im;

        const float kx = wave_component(x, N, patch_length);
        const float ki = inv_wavenumber(kx, kz);
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

    float2* x_team = xch[team];
#pragma unroll
    for (int f = 0; f < F; ++f) {
        uint32_t idx[8];
        fft_team<N>(fr[f], fi[f], idx, J, x_team, tw, BlockSync{});
        float2* dst = spec + size_t(f) * CELLS + size_t(z) * N;
#pragma unroll
        for (int a = 0; a < 8; ++a) dst[idx[a]] = make_float2(fr[f][a], fi[f][a]);
    }
}

template <int N, int F>
__global__ __launch_bounds__(ColShape<N>::BLOCK) void col_pass_kernel(
    const float2* __restrict__ spec, const float2* __restrict__ twg,
    float4* __restrict__ disp, float4* __restrict__ normal, float4* __restrict__ vel,
    float lambda, float foam_threshold, float inv_threshold)
{
    constexpr int T = ColShape<N>::T;
    constexpr int C = ColShape<N>::C;
    constexpr uint32_t CELLS = uint32_t(N) * N;
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[C][lds_span<N>()];

    for (int i = threadIdx.x; i < N / 2; i += ColShape<N>::BLOCK) tw[i] = twg[i];
    __syncthreads();

    const uint32_t c = threadIdx.x % C;
    const uint32_t J = threadIdx.x / C;
    const uint32_t x = blockIdx.x * C + c;

    float fr[F][8], fi[F][8];
#pragma unroll
    for (int f = 0; f < F; ++f) {
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const float2 v = spec[size_t(f) * CELLS + size_t(J + a * T) * N + x];
            fr[f][a] = v.x;
            fi[f][a] = v.y;
        }
    }

    uint32_t idx[8];
#pragma unroll
    for (int f = 0; f < F; ++f) fft_team<N>(fr[f], fi[f], idx, J, xch[c], tw, BlockSync{});

    // reference src/core/evolve.cpp:256-325
#pragma unroll
    for (int a = 0; a < 8; ++a) {
        const float h = fr[0][a], d_x = fi[0][a];
        const float d_z = fr[1][a], hx = fi[1][a];
        const float hz = fr[2][a], jxx = fi[2][a];
        const float jzz = fr[3][a], jxz = fi[3][a];
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
        const float inv_len = (len2 > 0.0f) ? 1.0f / sqrtf(len2) : 0.0f;
        nx *= inv_len;
        ny *= inv_len;
        nz *= inv_len;
        float foam = (foam_threshold - jac) * inv_threshold;
        if (foam < 0.0f) foam = 0.0f;
        if (foam > 1.0f) foam = 1.0f;
        const size_t o = size_t(idx[a]) * N + x;
        disp[o] = make_float4(disp_x, h, disp_z, foam);
        normal[o] = make_float4(nx, ny, nz, jac);
        if constexpr (F == 6) {
            // reference src/core/evolve.cpp:186-203
            vel[o] = make_float4(fr[4][a], fi[4][a], fr[5][a], 0.0f);
        }
    }
}

__global__ __launch_bounds__(kBlock) void sample_kernel(
    const float4* __restrict__ disp, const float4* __restrict__ normal,
    const float4* __restrict__ vel, uint32_t n, float inv_cell,
    const float2* __restrict__ xz, uint32_t count, ocean_hip_surface* __restrict__ out,
    float* __restrict__ heights)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const float2 p = xz[q];
    const ocean_hip_surface s = sample_level(disp, normal, vel, n, inv_cell, p.x, p.y);
    if (heights != nullptr) heights[q] = s.height;
    if (out != nullptr) out[q] = s;
}

__global__ __launch_bounds__(kBlock) void cascade_sample_kernel(
    const ocean_hip_level* __restrict__ levels, uint32_t level_count,
    const float2* __restrict__ xz, uint32_t count, ocean_hip_surface* __restrict__ out,
    float* __restrict__ heights)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const float2 p = xz[q];
    const ocean_hip_surface o = sample_cascade(levels, level_count, p.x, p.y);
    if (heights != nullptr) heights[q] = o.height;
    if (out != nullptr) out[q] = o;
}

// ---------------------------------------------------------------------------
// Foam advection step: reference src/foam.cpp:22-41 (sample_wrap) and 89-139.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void foam_step_kernel(
    const float4* __restrict__ disp, const float4* __restrict__ vel,
    const float* __restrict__ cur, float* __restrict__ next, uint32_t n,
    ocean_hip_foam_params p)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * n) return;
    const uint32_t x = i % n, z = i / n;
    const float4 w = vel[i];
    const float vx = w.x * p.scale + p.drift_x;
    const float vz = w.z * p.scale + p.drift_z;
    const float su = static_cast<float>(x) - vx * p.dt * p.inv_cell;
    const float sv = static_cast<float>(z) - vz * p.dt * p.inv_cell;
    float f = sample_wrap(cur, su, sv, static_cast<int>(n));
    const float src = disp[i].w;
    f = f * p.decay + src * p.gain;
    if (f > 1.0f) f = 1.0f;
    if (f < 0.0f) f = 0.0f;
    next[i] = f;
}

// FoamField::foam_at: reference src/foam.cpp:291-305.
__global__ __launch_bounds__(kBlock) void foam_query_kernel(
    const float4* __restrict__ disp, const float4* __restrict__ normal, uint32_t n,
    float inv_cell, const float* __restrict__ foam, float cell, const float2* __restrict__ xz,
    uint32_t count, float* __restrict__ out)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const float2 p = xz[q];
    const ocean_hip_surface s = sample_level(disp, normal, nullptr, n, inv_cell, p.x, p.y);
    const float u = (p.x - s.offset_x) / cell;
    const float v = (p.y - s.offset_z) / cell;
    out[q] = sample_wrap(foam, u, v, static_cast<int>(n));
}

constexpr uint32_t kCoverageBlocks = 256;

__global__ __launch_bounds__(kBlock) void coverage_partial_kernel(const float* __restrict__ v,
                                                                  uint32_t n,
                                                                  double* __restrict__ partial)
{
    __shared__ double red[kBlock];
    double acc = 0.0;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        acc += static_cast<double>(v[i]);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < uint32_t(w)) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void coverage_final_kernel(const double* __restrict__ partial, uint32_t blocks,
                                      uint32_t n, double* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (uint32_t b = 0; b < blocks; ++b) s += partial[b];
    *out = s / static_cast<double>(n);
}

template <int N>
hipError_t launch_frame_n(const ocean_hip_level* L, double time_over_two_pi, hipStream_t s,
                          unsigned passes)
{
    const bool v = L->vel != nullptr;
    const dim3 rg(N / RowShape<N>::TEAMS), rb(RowShape<N>::BLOCK);
    const dim3 cg(N / ColShape<N>::C), cb(ColShape<N>::BLOCK);
    const float2* h0 = reinterpret_cast<const float2*>(L->h0);
    const float2* tw = reinterpret_cast<const float2*>(L->twiddles);
    float2* spec = reinterpret_cast<float2*>(L->spec);
    float4* disp = reinterpret_cast<float4*>(L->disp);
    float4* nrm = reinterpret_cast<float4*>(L->normal);
    float4* vel = reinterpret_cast<float4*>(L->vel);
    if (v) {
        if (passes & 1u)
            hipLaunchKernelGGL((row_pass_kernel<N, 6>), rg, rb, 0, s, h0, L->omega, tw, spec,
                               time_over_two_pi, L->patch_length);
        if (passes & 2u)
            hipLaunchKernelGGL((col_pass_kernel<N, 6>), cg, cb, 0, s, spec, tw, disp, nrm, vel,
                               L->choppiness, L->foam_threshold, L->inv_foam_threshold);
    } else {
        if (passes & 1u)
            hipLaunchKernelGGL((row_pass_kernel<N, 4>), rg, rb, 0, s, h0, L->omega, tw, spec,
                               time_over_two_pi, L->patch_length);
        if (passes & 2u)
            hipLaunchKernelGGL((col_pass_kernel<N, 4>), cg, cb, 0, s, spec, tw, disp, nrm, vel,
                               L->choppiness, L->foam_threshold, L->inv_foam_threshold);
    }
    return hipGetLastError();
}

}  // namespace
}  // namespace oceanamd

using namespace oceanamd;

extern "C" hipError_t ocean_hip_frame_passes(const ocean_hip_level* L, double time_over_two_pi,
                                             hipStream_t s, unsigned passes)
{
    switch (L->n) {
        case 16: return launch_frame_n<16>(L, time_over_two_pi, s, passes);
        case 32: return launch_frame_n<32>(L, time_over_two_pi, s, passes);
        case 64: return launch_frame_n<64>(L, time_over_two_pi, s, passes);
        case 128: return launch_frame_n<128>(L, time_over_two_pi, s, passes);
        case 256: return launch_frame_n<256>(L, time_over_two_pi, s, passes);
        case 512: return launch_frame_n<512>(L, time_over_two_pi, s, passes);
        case 1024: return launch_frame_n<1024>(L, time_over_two_pi, s, passes);
        case 2048: return launch_frame_n<2048>(L, time_over_two_pi, s, passes);
        case 4096: return launch_frame_n<4096>(L, time_over_two_pi, s, passes);
        default: return hipErrorInvalidValue;
    }
}

extern "C" hipError_t ocean_hip_frame(const ocean_hip_level* L, double time_over_two_pi,
                                      hipStream_t s)
{
    return ocean_hip_frame_passes(L, time_over_two_pi, s, 3u);
}

extern "C" hipError_t ocean_hip_sample(const ocean_hip_level* L, const float* xz, uint32_t count,
                                       ocean_hip_surface* out, float* heights, hipStream_t s)
{
    if (count == 0) return hipSuccess;
    const float inv_cell = static_cast<float>(L->n) / L->patch_length;
    const dim3 g((count + kBlock - 1) / kBlock), b(kBlock);
    hipLaunchKernelGGL(sample_kernel, g, b, 0, s, reinterpret_cast<const float4*>(L->disp),
                       reinterpret_cast<const float4*>(L->normal),
                       reinterpret_cast<const float4*>(L->vel), L->n, inv_cell,
                       reinterpret_cast<const float2*>(xz), count, out, heights);
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_cascade_sample(const ocean_hip_level* levels,
                                               uint32_t level_count, const float* xz,
                                               uint32_t count, ocean_hip_surface* out,
                                               float* heights, hipStream_t s)
{
    if (count == 0) return hipSuccess;
    const dim3 g((count + kBlock - 1) / kBlock), b(kBlock);
    hipLaunchKernelGGL(cascade_sample_kernel, g, b, 0, s, levels, level_count,
                       reinterpret_cast<const float2*>(xz), count, out, heights);
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_foam_query(const ocean_hip_level* L, const float* foam,
                                           float cell, const float* xz, uint32_t count,
                                           float* out, hipStream_t s)
{
    if (count == 0) return hipSuccess;
    const float inv_cell = static_cast<float>(L->n) / L->patch_length;
    const dim3 g((count + kBlock - 1) / kBlock), b(kBlock);
    hipLaunchKernelGGL(foam_query_kernel, g, b, 0, s, reinterpret_cast<const float4*>(L->disp),
                       reinterpret_cast<const float4*>(L->normal), L->n, inv_cell, foam, cell,
                       reinterpret_cast<const float2*>(xz), count, out);
    return hipGetLastError();
}

extern "C" uint32_t ocean_hip_coverage_blocks(uint32_t n)
{
    const uint32_t want = (n + kBlock - 1) / kBlock;
    return want < kCoverageBlocks ? (want == 0 ? 1 : want) : kCoverageBlocks;
}

extern "C" hipError_t ocean_hip_coverage(const float* v, uint32_t n, double* partial,
                                         double* out, hipStream_t s)
{
    const uint32_t blocks = ocean_hip_coverage_blocks(n);
    hipLaunchKernelGGL(coverage_partial_kernel, dim3(blocks), dim3(kBlock), 0, s, v, n, partial);
    hipLaunchKernelGGL(coverage_final_kernel, dim3(1), dim3(64), 0, s, partial, blocks, n, out);
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_foam_step(const ocean_hip_level* L, const float* cur, float* next,
                                          const ocean_hip_foam_params* p, hipStream_t s)
{
    const uint32_t cells = L->n * L->n;
    const dim3 g((cells + kBlock - 1) / kBlock), b(kBlock);
    hipLaunchKernelGGL(foam_step_kernel, g, b, 0, s, reinterpret_cast<const float4*>(L->disp),
                       reinterpret_cast<const float4*>(L->vel), cur, next, L->n, *p);
    return hipGetLastError();
}

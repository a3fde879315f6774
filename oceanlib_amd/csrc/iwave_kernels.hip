// iWave interaction field kernels for gfx950.
//
// Per substep (reference src/interaction.cpp:560-587): source stamps
// (inject, 322-391), obstruction (295-318), one leapfrog step
// (src/core/iwave_step.cpp:7-54); per update the RGBA finalize
// (src/core/iwave_step.cpp:93-124). Every per-cell operation is the
// reference's, in its order, with no FMA:
//   - the convolution accumulates its (2P+1)^2 taps j-outer, i-inner into one
//     float, exactly as the scalar reference does, from a zero-haloed LDS tile;
//   - a stamp's mean is the reference's sequential double sum over the
//     clipped box in z-major order, and the per-cell stamp values use
//     expf_libm, bit-identical to the host expf the reference calls;
//   - overlapping stamps are applied per cell in queue order.
#include "device_expf.hpp"
#include "exact_sum.hpp"
#include "hip_launch.h"
#include "sample_device.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>

namespace oceanamd {
namespace {

constexpr int kBlock = 256;
constexpr int kTile = 16;                 // step kernel: 16x16 outputs per block

// reference src/interaction.cpp:71-75
__device__ __forceinline__ float ricker(float r2, float inv_two_sigma2)
{
    const float u = r2 * inv_two_sigma2;
    return (1.0f - u) * expf_libm(-u);
}

__device__ __forceinline__ float stamp_value(const ocean_hip_stamp& st, int x, int z, float cell)
{
    const float px = (static_cast<float>(x) + 0.5f) - st.gx;
    const float pz = (static_cast<float>(z) + 0.5f) - st.gz;
    return ricker((px * px + pz * pz) * cell * cell, st.inv2s2);
}

// One block per stamp, over windows of up to kWindow 256-value chunks of the
// stamp's clipped box in z-major order (the reference's order). Per window
// the wavefronts compute the values and gather the exactness statistics
// (exact_sum.hpp) of each chunk and of each group of four; then one lane
// carries the reference's sequential double sum through the window: a group
// the statistics prove rounding-free is added as its exact total in one step,
// else each chunk that is, and any other chunk term by term, its values
// recomputed by that lane (the same function of the cell, so the same bits).
// The result is the sequential sum bit for bit either way. Only the
// statistics are held in LDS, so the blocks share their CUs with the other
// streams' passes (the cascade's, in a scene).
constexpr int kSumChunk = 256;
constexpr int kGroup = 4;     // chunks per group
constexpr int kWindow = 128;  // chunks per window
constexpr int kMeanBlock = 1024;

// A substep's stamps reach the kernels either in device memory or, when there
// are at most kArgStamps of them, inside the kernel arguments (no upload).
struct DevStamps {
    const ocean_hip_stamp* __restrict__ p;
    __device__ __forceinline__ ocean_hip_stamp operator[](uint32_t i) const { return p[i]; }
};
constexpr uint32_t kArgStamps = 32;
struct ArgStamps {
    ocean_hip_stamp s[kArgStamps];
    __device__ __forceinline__ ocean_hip_stamp operator[](uint32_t i) const { return s[i]; }
};

template <typename Stamps>
__global__ __launch_bounds__(kMeanBlock) void stamp_mean_kernel(const Stamps stamps, float cell,
                                                                float* __restrict__ means)
{
    __shared__ double c_total[kWindow], c_abs[kWindow];
    __shared__ int c_q[kWindow], c_finite[kWindow];
    __shared__ ChunkStats small[kWindow];
    __shared__ ChunkStats big[kWindow / kGroup];
    const ocean_hip_stamp st = stamps[blockIdx.x];
    const int w = st.x1 - st.x0 + 1;
    const int h = st.z1 - st.z0 + 1;
    const int cnt = w * h;
    const int chunks = (cnt + kSumChunk - 1) / kSumChunk;
    const int lane = threadIdx.x % warpSize;
    double s = 0.0;  // the running sum, carried by thread 0
    for (int c0 = 0; c0 < chunks; c0 += kWindow) {
        const int wc = chunks - c0 < kWindow ? chunks - c0 : kWindow;
        const int groups = (wc + kGroup - 1) / kGroup;
        for (int c = threadIdx.x / warpSize; c < wc; c += kMeanBlock / warpSize) {
            const int begin = (c0 + c) * kSumChunk;
            const int end = cnt - begin < kSumChunk ? cnt : begin + kSumChunk;
            ChunkAcc acc;
            for (int i = begin + lane; i < end; i += warpSize)
                acc.add(stamp_value(st, st.x0 + i % w, st.z0 + i / w, cell));
            for (int off = warpSize / 2; off >= 1; off /= 2)
                acc.merge(__shfl_xor(acc.total, off), __shfl_xor(acc.abs_sum, off),
                          __shfl_xor(acc.q, off), __shfl_xor(int(acc.finite), off) != 0);
            if (lane == 0) {
                c_total[c] = acc.total;
                c_abs[c] = acc.abs_sum;
                c_q[c] = acc.q;
                c_finite[c] = acc.finite ? 1 : 0;
                small[c] = acc.finish();
            }
        }
        __syncthreads();
        for (int g = threadIdx.x; g < groups; g += kMeanBlock) {
            ChunkAcc acc;
            const int end = wc < (g + 1) * kGroup ? wc : (g + 1) * kGroup;
            for (int c = g * kGroup; c < end; ++c) acc.merge(c_total[c], c_abs[c], c_q[c], c_finite[c] != 0);
            big[g] = acc.finish();
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int g = 0; g < groups; ++g) {
                if (chunk_is_exact(s, big[g])) {
                    s += big[g].total;
                    continue;
                }
                const int c_end = wc < (g + 1) * kGroup ? wc : (g + 1) * kGroup;
                for (int c = g * kGroup; c < c_end; ++c) {
                    if (chunk_is_exact(s, small[c])) {
                        s += small[c].total;
                        continue;
                    }
                    const int begin = (c0 + c) * kSumChunk;
                    const int end = cnt - begin < kSumChunk ? cnt : begin + kSumChunk;
                    for (int i = begin; i < end; ++i)
                        s += static_cast<double>(stamp_value(st, st.x0 + i % w, st.z0 + i / w, cell));
                }
            }
        }
        __syncthreads();  // the window's statistics are read before the next window writes them
    }
    if (threadIdx.x == 0)
        means[blockIdx.x] = static_cast<float>(s / (static_cast<double>(w) * static_cast<double>(h)));
}

template <typename Stamps>
__global__ __launch_bounds__(kBlock) void stamp_apply_kernel(ocean_hip_iwave f, const Stamps stamps,
                                                             uint32_t count,
                                                             const float* __restrict__ means)
{
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int z = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= int(f.n) || z >= int(f.n)) return;
    const size_t ci = size_t(z) * f.stride + x;
    float c = 0.0f, o = 0.0f;
    bool touched = false;
    for (uint32_t r = 0; r < count; ++r) {
        const ocean_hip_stamp st = stamps[r];
        if (x < st.x0 || x > st.x1 || z < st.z0 || z > st.z1) continue;
        if (!touched) {
            c = f.cur[ci];
            o = f.old[ci];
            touched = true;
        }
        const float sv = stamp_value(st, x, z, f.cell) - means[r];
        const float dd = st.disp * sv;
        const float dv = st.vel * sv;
        c += dd;
        o += dd - dv * f.dt;
    }
    if (touched) {
        f.cur[ci] = c;
        f.old[ci] = o;
    }
}

// The solid list and its count come from the device (the mask tables), so
// the grid strides over an upper bound.
__global__ __launch_bounds__(kBlock) void obstruct_kernel(ocean_hip_iwave f,
                                                          const uint32_t* __restrict__ solid,
                                                          const uint32_t* __restrict__ source,
                                                          const uint32_t* __restrict__ count_p,
                                                          int dirichlet)
{
    const uint32_t count = *count_p;
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < count; k += gridDim.x * kBlock) {
        const uint32_t i = solid[k];
        const size_t dst = size_t(i / f.n) * f.stride + (i % f.n);
        if (dirichlet) {
            f.cur[dst] = 0.0f;
        } else {
            const uint32_t src = source[k];
            if (src == 0xFFFFFFFFu) continue;
            f.cur[dst] = f.cur[size_t(src / f.n) * f.stride + (src % f.n)];
        }
    }
}

// reference src/core/iwave_step.cpp:7-54, on a (16+2P)^2 LDS tile.
template <int P>
__global__ __launch_bounds__(kBlock) void step_kernel(ocean_hip_iwave f)
{
    constexpr int W = 2 * P + 1;
    constexpr int TW = kTile + 2 * P;
    __shared__ float tile[TW][TW + 1];
    const int bx = blockIdx.x * kTile, bz = blockIdx.y * kTile;
    for (int i = threadIdx.x; i < TW * TW; i += kBlock) {
        const int tz = i / TW, tx = i % TW;
        const int gz = bz + tz - P, gx = bx + tx - P;   // within [-P, n+P): halo is zero
        tile[tz][tx] = f.cur[ptrdiff_t(gz) * f.stride + gx];
    }
    __syncthreads();
    const int lx = threadIdx.x & 15, lz = threadIdx.x >> 4;
    const int x = bx + lx, z = bz + lz;
    if (x >= int(f.n) || z >= int(f.n)) return;
    float acc = 0.0f;
#pragma unroll
    for (int j = 0; j < W; ++j) {
#pragma unroll
        for (int i = 0; i < W; ++i) acc += f.taps[j * W + i] * tile[lz + j][lx + i];
    }
    const size_t row = size_t(z) * f.stride;
    const size_t orow = size_t(z) * f.n;
    const float hc = tile[lz + P][lx + P];
    const float hp = f.old[row + x];
    const float v = f.c1[orow + x] * (2.0f * hc - f.c2[orow + x] * hp - f.gdt2 * acc);
    f.old[row + x] = (v > -1.0e-30f && v < 1.0e-30f) ? 0.0f : v;
}

// reference src/core/iwave_step.cpp:93-124
__global__ __launch_bounds__(kBlock) void finalize_kernel(ocean_hip_iwave f)
{
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int z = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= int(f.n) || z >= int(f.n)) return;
    const ptrdiff_t st = f.stride;
    const float* c = f.cur + z * st + x;
    const float gx = (c[1] - c[-1]) * f.inv_two_cell;
    const float gz = (c[st] - c[-st]) * f.inv_two_cell;
    const float vy = (*c - f.old[z * st + x]) * f.inv_dt;
    reinterpret_cast<float4*>(f.out)[size_t(z) * f.n + x] = make_float4(*c, gx, gz, vy);
}

// reference src/interaction.cpp:620-668 (shift_plane): integer shift, zero fill.
__global__ __launch_bounds__(kBlock) void shift_kernel(const float* __restrict__ src,
                                                       float* __restrict__ dst, uint32_t n,
                                                       uint32_t stride, int32_t dx, int32_t dz)
{
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int z = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= int(n) || z >= int(n)) return;
    const int sx = x + dx, sz = z + dz;
    const bool in = sx >= 0 && sx < int(n) && sz >= 0 && sz < int(n);
    dst[size_t(z) * stride + x] = in ? src[size_t(sz) * stride + sx] : 0.0f;
}

// InteractionField::sample_at, reference src/interaction.cpp:729-773.
__device__ __forceinline__ float4 iwave_sample_one(const ocean_hip_iwave& f, float wx, float wz)
{
    const float gx = (wx - static_cast<float>(f.origin_x) * f.cell) / f.cell - 0.5f;
    const float gz = (wz - static_cast<float>(f.origin_z) * f.cell) / f.cell - 0.5f;
    if (gx < 0.0f || gz < 0.0f || gx >= static_cast<float>(f.n - 1) ||
        gz >= static_cast<float>(f.n - 1))
        return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const int x0 = static_cast<int>(gx);
    const int z0 = static_cast<int>(gz);
    const float tx = gx - static_cast<float>(x0);
    const float tz = gz - static_cast<float>(z0);
    const float4* o = reinterpret_cast<const float4*>(f.out);
    const size_t i00 = size_t(z0) * f.n + x0;
    const float4 a = o[i00], b = o[i00 + 1], c = o[i00 + f.n], d = o[i00 + f.n + 1];
    const float w00 = (1.0f - tx) * (1.0f - tz);
    const float w10 = tx * (1.0f - tz);
    const float w01 = (1.0f - tx) * tz;
    const float w11 = tx * tz;
    float4 r;
    r.x = a.x * w00 + b.x * w10 + c.x * w01 + d.x * w11;
    r.y = a.y * w00 + b.y * w10 + c.y * w01 + d.y * w11;
    r.z = a.z * w00 + b.z * w10 + c.z * w01 + d.z * w11;
    r.w = a.w * w00 + b.w * w10 + c.w * w01 + d.w * w11;
    return r;
}

__global__ __launch_bounds__(kBlock) void iwave_sample_kernel(ocean_hip_iwave f,
                                                              const float2* __restrict__ xz,
                                                              uint32_t count,
                                                              float4* __restrict__ out)
{
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= count) return;
    out[q] = iwave_sample_one(f, xz[q].x, xz[q].y);
}

// WaterSurface::sample_at, reference src/interaction.cpp:863-901.
__global__ __launch_bounds__(kBlock) void combined_sample_kernel(
    const ocean_hip_level* __restrict__ levels, uint32_t level_count, int as_cascade,
    ocean_hip_iwave f, int has_field, const float2* __restrict__ xz, uint32_t count,
    ocean_hip_surface* __restrict__ out)
{
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= count) return;
    const float2 p = xz[q];
    ocean_hip_surface s;
    if (as_cascade) {
        s = sample_cascade(levels, level_count, p.x, p.y);
    } else {
        const ocean_hip_level L = levels[0];
        s = sample_level(reinterpret_cast<const float4*>(L.disp),
                         reinterpret_cast<const float4*>(L.normal),
                         reinterpret_cast<const float4*>(L.vel), L.n,
                         static_cast<float>(L.n) / L.patch_length, p.x, p.y);
    }
    if (has_field) {
        const float4 e = iwave_sample_one(f, p.x, p.y);
        s.height += e.x;
        s.velocity_y += e.w;
        if (s.normal_y > 1.0e-6f) {
            const float inv = 1.0f / s.normal_y;
            const float sx = -s.normal_x * inv + e.y;
            const float sz = -s.normal_z * inv + e.z;
            const float len = sqrtf(sx * sx + sz * sz + 1.0f);
            const float il = 1.0f / len;
            s.normal_x = -sx * il;
            s.normal_y = il;
            s.normal_z = -sz * il;
        }
    }
    out[q] = s;
}

// Per-cell terms of the reference's energy() (src/interaction.cpp:803-846):
// (hc - hp)/dt squared, and hc times the double tap sum over h^{n-1}.
__global__ __launch_bounds__(kBlock) void energy_terms_kernel(ocean_hip_iwave f, double inv_dt,
                                                              double* __restrict__ kin,
                                                              double* __restrict__ pot)
{
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int z = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= int(f.n) || z >= int(f.n)) return;
    const int p = int(f.p), w = 2 * p + 1;
    const ptrdiff_t st = f.stride;
    const double hc = f.cur[z * st + x];
    const double hp = f.old[z * st + x];
    const double v = (hc - hp) * inv_dt;
    double acc = 0.0;
    const float* tap = f.taps;
    for (int j = -p; j <= p; ++j) {
        const float* srow = f.old + (z + j) * st + x - p;
        for (int i = 0; i < w; ++i) acc += tap[i] * srow[i];
        tap += w;
    }
    const size_t k = size_t(z) * f.n + x;
    kin[k] = v * v;
    pot[k] = hc * acc;
}

// Sequential sums in the reference's cell order, then 0.5 (kin + g pot).
__global__ void energy_sum_kernel(const double* __restrict__ kin, const double* __restrict__ pot,
                                  uint32_t cells, double gravity, double* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double k = 0.0, p = 0.0;
    for (uint32_t i = 0; i < cells; ++i) {
        k += kin[i];
        p += pot[i];
    }
    *out = 0.5 * (k + gravity * p);
}

template <int P>
hipError_t launch_step(const ocean_hip_iwave* f, hipStream_t s)
{
    const dim3 g((f->n + kTile - 1) / kTile, (f->n + kTile - 1) / kTile);
    hipLaunchKernelGGL(step_kernel<P>, g, dim3(kBlock), 0, s, *f);
    return hipGetLastError();
}

dim3 cell_grid(uint32_t n) { return dim3((n + 15) / 16, (n + 15) / 16); }

}  // namespace
}  // namespace oceanamd

using namespace oceanamd;

extern "C" hipError_t ocean_hip_iwave_inject(const ocean_hip_iwave* f, const ocean_hip_stamp* stamps,
                                             uint32_t count, float* means, hipStream_t s)
{
    if (count == 0) return hipSuccess;
    const DevStamps d{stamps};
    hipLaunchKernelGGL(stamp_mean_kernel<DevStamps>, dim3(count), dim3(kMeanBlock), 0, s, d, f->cell,
                       means);
    hipLaunchKernelGGL(stamp_apply_kernel<DevStamps>, cell_grid(f->n), dim3(kBlock), 0, s, *f, d,
                       count, static_cast<const float*>(means));
    return hipGetLastError();
}

extern "C" uint32_t ocean_hip_iwave_arg_stamps() { return kArgStamps; }

extern "C" hipError_t ocean_hip_iwave_inject_args(const ocean_hip_iwave* f,
                                                  const ocean_hip_stamp* host_stamps, uint32_t count,
                                                  float* means, hipStream_t s)
{
    if (count == 0) return hipSuccess;
    if (count > kArgStamps) return hipErrorInvalidValue;
    ArgStamps a{};
    for (uint32_t i = 0; i < count; ++i) a.s[i] = host_stamps[i];
    hipLaunchKernelGGL(stamp_mean_kernel<ArgStamps>, dim3(count), dim3(kMeanBlock), 0, s, a, f->cell,
                       means);
    hipLaunchKernelGGL(stamp_apply_kernel<ArgStamps>, cell_grid(f->n), dim3(kBlock), 0, s, *f, a,
                       count, static_cast<const float*>(means));
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_iwave_obstruct(const ocean_hip_iwave* f, const uint32_t* solid,
                                               const uint32_t* source, const uint32_t* count,
                                               uint32_t max_count, int dirichlet, hipStream_t s)
{
    if (max_count == 0) return hipSuccess;
    const uint32_t blocks = std::min<uint32_t>((max_count + kBlock - 1) / kBlock, 512u);
    hipLaunchKernelGGL(obstruct_kernel, dim3(blocks), dim3(kBlock), 0, s, *f, solid, source, count,
                       dirichlet);
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_iwave_step(const ocean_hip_iwave* f, hipStream_t s)
{
    switch (f->p) {
        case 1: return launch_step<1>(f, s);
        case 2: return launch_step<2>(f, s);
        case 3: return launch_step<3>(f, s);
        case 4: return launch_step<4>(f, s);
        case 5: return launch_step<5>(f, s);
        case 6: return launch_step<6>(f, s);
        case 7: return launch_step<7>(f, s);
        case 8: return launch_step<8>(f, s);
        case 9: return launch_step<9>(f, s);
        case 10: return launch_step<10>(f, s);
        case 11: return launch_step<11>(f, s);
        case 12: return launch_step<12>(f, s);
        default: return hipErrorInvalidValue;
    }
}

extern "C" hipError_t ocean_hip_iwave_energy(const ocean_hip_iwave* f, double gravity,
                                             double* scratch, double* out, hipStream_t s)
{
    const uint32_t cells = f->n * f->n;
    hipLaunchKernelGGL(energy_terms_kernel, cell_grid(f->n), dim3(kBlock), 0, s, *f,
                       1.0 / static_cast<double>(f->dt), scratch, scratch + cells);
    hipLaunchKernelGGL(energy_sum_kernel, dim3(1), dim3(64), 0, s,
                       static_cast<const double*>(scratch),
                       static_cast<const double*>(scratch + cells), cells, gravity, out);
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_iwave_finalize(const ocean_hip_iwave* f, hipStream_t s)
{
    hipLaunchKernelGGL(finalize_kernel, cell_grid(f->n), dim3(kBlock), 0, s, *f);
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_iwave_shift(const float* src, float* dst, uint32_t n,
                                            uint32_t stride, int32_t dx, int32_t dz, hipStream_t s)
{
    hipLaunchKernelGGL(shift_kernel, cell_grid(n), dim3(kBlock), 0, s, src, dst, n, stride, dx, dz);
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_iwave_sample(const ocean_hip_iwave* f, const float* xz,
                                             uint32_t count, float* out, hipStream_t s)
{
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(iwave_sample_kernel, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       *f, reinterpret_cast<const float2*>(xz), count,
                       reinterpret_cast<float4*>(out));
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_combined_sample(const ocean_hip_level* levels, uint32_t level_count,
                                                int as_cascade, const ocean_hip_iwave* field,
                                                const float* xz, uint32_t count,
                                                ocean_hip_surface* out, hipStream_t s)
{
    if (count == 0) return hipSuccess;
    ocean_hip_iwave f{};
    if (field) f = *field;
    hipLaunchKernelGGL(combined_sample_kernel, dim3((count + kBlock - 1) / kBlock), dim3(kBlock),
                       0, s, levels, level_count, as_cascade, f, field != nullptr ? 1 : 0,
                       reinterpret_cast<const float2*>(xz), count, out);
    return hipGetLastError();
}

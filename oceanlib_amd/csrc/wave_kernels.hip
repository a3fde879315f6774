// Query and foam kernels for gfx950 (MI355X).
//
//   sample_kernel / cascade_sample_kernel : batched Ocean::sample_at and
//       CascadeStack::sample_at (reference src/ocean.cpp:309-452,
//       src/cascade.cpp:82-153), one lane per point;
//   foam_step_kernel / foam_query_kernel : FoamField's semi-Lagrangian step
//       and foam_at (reference src/foam.cpp:22-139,291-305);
//   coverage_kernel : FoamField::coverage, the reference's sequential double
//       sum bit for bit, by one workgroup (ordered_sum.hpp).
// The frame kernels themselves are in frame_kernels.hip.
#include "hip_launch.h"
#include "ordered_sum.hpp"
#include "sample_device.hpp"

#include <hip/hip_runtime.h>

namespace oceanamd {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ void sample_point(uint32_t q, const float4* __restrict__ disp,
                                             const float4* __restrict__ normal,
                                             const float4* __restrict__ vel, uint32_t n,
                                             float inv_cell, const float2* __restrict__ xz,
                                             uint32_t count, ocean_hip_surface* __restrict__ out,
                                             float* __restrict__ heights)
{
    if (q >= count) return;
    const float2 p = xz[q];
    if (out == nullptr) {
        // Heights only (Ocean::height_at): the same taps, without the normal
        // and velocity loads.
        if (heights != nullptr) heights[q] = height_level(disp, n, inv_cell, p.x, p.y);
        return;
    }
    const ocean_hip_surface s = sample_level(disp, normal, vel, n, inv_cell, p.x, p.y);
    if (heights != nullptr) heights[q] = s.height;
    out[q] = s;
}

__global__ __launch_bounds__(kBlock) void sample_kernel(
    const float4* __restrict__ disp, const float4* __restrict__ normal,
    const float4* __restrict__ vel, uint32_t n, float inv_cell,
    const float2* __restrict__ xz, uint32_t count, ocean_hip_surface* __restrict__ out,
    float* __restrict__ heights)
{
    sample_point(blockIdx.x * blockDim.x + threadIdx.x, disp, normal, vel, n, inv_cell, xz, count,
                 out, heights);
}

__global__ __launch_bounds__(kBlock) void cascade_sample_kernel(
    const ocean_hip_level* __restrict__ levels, uint32_t level_count,
    const float2* __restrict__ xz, uint32_t count, ocean_hip_surface* __restrict__ out,
    float* __restrict__ heights)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const float2 p = xz[q];
    if (out == nullptr) {
        // Heights only: each level's height taps alone, summed in level order
        // from 0 as the full query sums them.
        if (heights == nullptr) return;
        float height = 0.0f;
        for (uint32_t l = 0; l < level_count; ++l) {
            const ocean_hip_level L = levels[l];
            height += height_level(reinterpret_cast<const float4*>(L.disp), L.n,
                                   static_cast<float>(L.n) / L.patch_length, p.x, p.y);
        }
        heights[q] = height;
        return;
    }
    const ocean_hip_surface o = sample_cascade(levels, level_count, p.x, p.y);
    if (heights != nullptr) heights[q] = o.height;
    if (out != nullptr) out[q] = o;
}

// CascadeStack::height_at for large batches, one launch per level: every
// query taps one level's displacement (4 MB at 512^2, which an XCD's L2
// holds) at a time instead of all levels' at once. The running sum is kept in
// `heights` between the launches, with the reference's float additions in
// level order (the first level's launch adds to 0.0f, as the reference's sum
// starts, src/cascade.cpp:94,118-120).
// Two lanes per point (height_level_lane_pair): kBlock / 2 points per block;
// a pair past the end follows the last point and stores nothing.
__global__ __launch_bounds__(kBlock) void cascade_height_level_kernel(
    const ocean_hip_level* __restrict__ level, bool first, const float2* __restrict__ xz,
    uint32_t count, float* __restrict__ heights)
{
    const uint32_t q = (blockIdx.x * blockDim.x + threadIdx.x) >> 1;
    const bool odd = (threadIdx.x & 1u) != 0u;
    const ocean_hip_level L = *level;
    const float2 p = xz[q < count ? q : count - 1u];
    const float h = height_level_lane_pair(reinterpret_cast<const float4*>(L.disp), L.n,
                                           static_cast<float>(L.n) / L.patch_length, p.x, p.y, odd);
    if (q < count && !odd) heights[q] = (first ? 0.0f : heights[q]) + h;
}

// Heights alone of one ocean for large batches, two lanes per point as above.
__global__ __launch_bounds__(kBlock) void height_lane_pair_kernel(
    const float4* __restrict__ disp, uint32_t n, float inv_cell, const float2* __restrict__ xz,
    uint32_t count, float* __restrict__ heights)
{
    const uint32_t q = (blockIdx.x * blockDim.x + threadIdx.x) >> 1;
    const bool odd = (threadIdx.x & 1u) != 0u;
    const float2 p = xz[q < count ? q : count - 1u];
    const float h = height_level_lane_pair(disp, n, inv_cell, p.x, p.y, odd);
    if (q < count && !odd) heights[q] = h;
}

// reference src/foam.cpp:89-139
__device__ __forceinline__ void foam_step_cell(uint32_t i, const float4* __restrict__ disp,
                                               const float4* __restrict__ vel,
                                               const float* __restrict__ cur,
                                               float* __restrict__ next, uint32_t n,
                                               const ocean_hip_foam_params& p)
{
    if (i >= n * n) return;
    const uint32_t x = i & (n - 1u), z = i >> __builtin_ctz(n);  // n is a power of two
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

__global__ __launch_bounds__(kBlock) void foam_step_kernel(
    const float4* __restrict__ disp, const float4* __restrict__ vel,
    const float* __restrict__ cur, float* __restrict__ next, uint32_t n,
    ocean_hip_foam_params p)
{
    foam_step_cell(blockIdx.x * blockDim.x + threadIdx.x, disp, vel, cur, next, n, p);
}

// One foam step and a batch of queries of the same frame in one launch: the
// first `qblocks` workgroups answer the queries (their dependent tap chains
// start first), the rest step the foam. Neither reads what the other writes.
__global__ __launch_bounds__(kBlock) void foam_step_sample_kernel(
    const float4* __restrict__ disp, const float4* __restrict__ normal,
    const float4* __restrict__ vel, uint32_t n, float inv_cell, const float* __restrict__ cur,
    float* __restrict__ next, ocean_hip_foam_params p, const float2* __restrict__ xz,
    uint32_t count, ocean_hip_surface* __restrict__ out, float* __restrict__ heights,
    uint32_t qblocks)
{
    if (blockIdx.x < qblocks)
        sample_point(blockIdx.x * blockDim.x + threadIdx.x, disp, normal, vel, n, inv_cell, xz,
                     count, out, heights);
    else
        foam_step_cell((blockIdx.x - qblocks) * blockDim.x + threadIdx.x, disp, vel, cur, next, n,
                       p);
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

// FoamField::coverage (reference src/foam.cpp:307-313): s = 0; s += v[i] in
// index order, in double, then s / n. One workgroup walks the array in chunks
// of 4096 terms (4 consecutive per lane); within a chunk, runs of terms that
// keep s in one binade are added in one exact step from a scan of the terms
// rounded to that binade's spacing, and the term that ends a run is added the
// plain way (ordered_sum.hpp). Same bits as the sequential loop.
constexpr uint32_t kCovLanes = 1024;
constexpr uint32_t kCovPer = 4;
constexpr uint32_t kCovChunk = kCovLanes * kCovPer;

__device__ __forceinline__ void coverage_load(const float* __restrict__ v, uint32_t n,
                                              uint32_t first, float (&t)[kCovPer])
{
    if (first + kCovPer <= n) {
        const float4 q = *reinterpret_cast<const float4*>(v + first);
        t[0] = q.x; t[1] = q.y; t[2] = q.z; t[3] = q.w;
    } else {
#pragma unroll
        for (uint32_t j = 0; j < kCovPer; ++j) t[j] = first + j < n ? v[first + j] : 0.0f;
    }
}

__global__ __launch_bounds__(kCovLanes) void coverage_kernel(const float* __restrict__ v,
                                                             uint32_t n, double* __restrict__ out)
{
    __shared__ double wave_sum[kCovLanes / 64];
    __shared__ uint32_t wave_end[kCovLanes / 64];
    __shared__ double bc_prefix;
    __shared__ float bc_term;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const bool vec = (reinterpret_cast<uintptr_t>(v) & 15u) == 0u;
    double s = 0.0;
    float t[kCovPer], next[kCovPer];
    if (vec) {
        coverage_load(v, n, tid * kCovPer, next);
    } else {
#pragma unroll
        for (uint32_t j = 0; j < kCovPer; ++j) next[j] = tid * kCovPer + j < n ? v[tid * kCovPer + j] : 0.0f;
    }
    for (uint32_t base = 0; base < n; base += kCovChunk) {
#pragma unroll
        for (uint32_t j = 0; j < kCovPer; ++j) t[j] = next[j];
        const uint32_t nb = base + kCovChunk + tid * kCovPer;  // the next chunk, in flight
        if (vec) {
            coverage_load(v, n, nb, next);
        } else {
#pragma unroll
            for (uint32_t j = 0; j < kCovPer; ++j) next[j] = nb + j < n ? v[nb + j] : 0.0f;
        }
        // Terms past n read as +0.0: they neither end a run nor change s.
        uint32_t pos = 0;  // chunk-local index of the first term not yet in s
        while (pos < kCovChunk) {
            const Binade bin = binade_of(s);
            if (bin.mode == 2) {
                // s is negative or not finite: the rest term by term, one lane
                if (tid == 0) {
                    for (uint32_t i = base + pos; i < n; ++i) s += static_cast<double>(v[i]);
                    *out = s / static_cast<double>(n);
                }
                return;
            }
            double p[kCovPer];
            bool stop[kCovPer];
            double run = 0.0;
#pragma unroll
            for (uint32_t j = 0; j < kCovPer; ++j) {
                bool st = false;
                double r = 0.0;
                if (tid * kCovPer + j >= pos) r = rounded_term(t[j], bin, st);
                run += r;
                p[j] = run;
                stop[j] = st;
            }
            double x = run;  // inclusive scan of the lanes' sums over the wavefront
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const double y = __shfl_up(x, off, 64);
                if (lane >= uint32_t(off)) x += y;
            }
            double lane_before = __shfl_up(x, 1, 64);
            if (lane == 0) lane_before = 0.0;
            if (lane == 63) wave_sum[wave] = x;
            __syncthreads();
            double before = 0.0;
            for (uint32_t w = 0; w < wave; ++w) before += wave_sum[w];
            before += lane_before;  // this chunk's r before this lane's terms
            uint32_t end = kCovChunk;
#pragma unroll
            for (int j = kCovPer - 1; j >= 0; --j)
                if (tid * kCovPer + j >= pos && ends_run(stop[j], before + p[j], bin))
                    end = tid * kCovPer + j;
            uint32_t m = end;
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) m = min(m, uint32_t(__shfl_xor(int(m), off, 64)));
            if (lane == 0) wave_end[wave] = m;
            __syncthreads();
            uint32_t e = kCovChunk;
            for (uint32_t w = 0; w < kCovLanes / 64; ++w) e = min(e, wave_end[w]);
            if (e == kCovChunk) {
                double total = 0.0;
                for (uint32_t w = 0; w < kCovLanes / 64; ++w) total += wave_sum[w];
                s += total;  // exact: the whole chunk stays in the binade
                pos = kCovChunk;
            } else {
                if (tid == e / kCovPer) {
                    const uint32_t j = e % kCovPer;
                    double pre = before;
#pragma unroll
                    for (uint32_t k = 0; k < kCovPer; ++k)
                        if (k + 1 == j) pre = before + p[k];
                    bc_prefix = pre;
#pragma unroll
                    for (uint32_t k = 0; k < kCovPer; ++k)
                        if (k == j) bc_term = t[k];
                }
                __syncthreads();
                s += bc_prefix;                       // exact: the run before the end
                s += static_cast<double>(bc_term);    // the ending term, rounded as usual
                pos = e + 1;
            }
            __syncthreads();  // wave_sum / wave_end / bc_* are rewritten next
        }
    }
    if (tid == 0) *out = s / static_cast<double>(n);
}

}  // namespace
}  // namespace oceanamd

using namespace oceanamd;

namespace {

// Query launches: a query is a chain of dependent taps, so a small batch (a
// frame's few thousand points) is latency-bound; one wavefront per workgroup
// spreads it over four times as many CUs (and their L1s) as 256-lane blocks.
dim3 query_block(uint32_t count) { return dim3(count <= 65536u ? 64u : uint32_t(kBlock)); }

}  // namespace

extern "C" hipError_t ocean_hip_sample(const ocean_hip_level* L, const float* xz, uint32_t count,
                                       ocean_hip_surface* out, float* heights, hipStream_t s)
{
    if (count == 0) return hipSuccess;
    const float inv_cell = static_cast<float>(L->n) / L->patch_length;
    if (out == nullptr && heights != nullptr && count > 65536u) {
        // heights alone, many points: two lanes per point
        hipLaunchKernelGGL(height_lane_pair_kernel, dim3((count + kBlock / 2 - 1) / (kBlock / 2)), dim3(kBlock), 0,
                           s, reinterpret_cast<const float4*>(L->disp), L->n, inv_cell,
                           reinterpret_cast<const float2*>(xz), count, heights);
        return hipGetLastError();
    }
    const dim3 b = query_block(count), g((count + b.x - 1) / b.x);
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
    if (out == nullptr && heights != nullptr && count > 65536u) {
        // heights alone, many points: level by level, two lanes per point
        const dim3 g((count + kBlock / 2 - 1) / (kBlock / 2)), b(kBlock);
        for (uint32_t l = 0; l < level_count; ++l)
            hipLaunchKernelGGL(cascade_height_level_kernel, g, b, 0, s, levels + l, l == 0,
                               reinterpret_cast<const float2*>(xz), count, heights);
        return hipGetLastError();
    }
    const dim3 b = query_block(count), g((count + b.x - 1) / b.x);
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

extern "C" hipError_t ocean_hip_coverage(const float* v, uint32_t n, double* out, hipStream_t s)
{
    if (n == 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(coverage_kernel, dim3(1), dim3(kCovLanes), 0, s, v, n, out);
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_foam_step_sample(const ocean_hip_level* L, const float* cur,
                                                 float* next, const ocean_hip_foam_params* p,
                                                 const float* xz, uint32_t count,
                                                 ocean_hip_surface* out, float* heights,
                                                 hipStream_t s)
{
    const uint32_t cells = L->n * L->n;
    const uint32_t b = 64;  // one wavefront per workgroup: the queries' latency chains spread
    const uint32_t qblocks = (count + b - 1) / b;
    const uint32_t fblocks = (cells + b - 1) / b;
    const float inv_cell = static_cast<float>(L->n) / L->patch_length;
    hipLaunchKernelGGL(foam_step_sample_kernel, dim3(qblocks + fblocks), dim3(b), 0, s,
                       reinterpret_cast<const float4*>(L->disp),
                       reinterpret_cast<const float4*>(L->normal),
                       reinterpret_cast<const float4*>(L->vel), L->n, inv_cell, cur, next, *p,
                       reinterpret_cast<const float2*>(xz), count, out, heights, qblocks);
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

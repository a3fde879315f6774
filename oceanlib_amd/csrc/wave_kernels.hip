// Query and foam kernels for gfx950 (MI355X).
//
//   sample_kernel / cascade_sample_kernel : batched Ocean::sample_at and
//       CascadeStack::sample_at (reference src/ocean.cpp:309-452,
//       src/cascade.cpp:82-153), one lane per point;
//   foam_step_kernel / foam_query_kernel : FoamField's semi-Lagrangian step
//       and foam_at (reference src/foam.cpp:22-139,291-305);
//   coverage kernels : FoamField::coverage as a device reduction.
// The frame kernels themselves are in frame_kernels.hip.
#include "hip_launch.h"
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
    const ocean_hip_surface s = sample_level(disp, normal, vel, n, inv_cell, p.x, p.y);
    if (heights != nullptr) heights[q] = s.height;
    if (out != nullptr) out[q] = s;
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
    const ocean_hip_surface o = sample_cascade(levels, level_count, p.x, p.y);
    if (heights != nullptr) heights[q] = o.height;
    if (out != nullptr) out[q] = o;
}

// reference src/foam.cpp:89-139
__device__ __forceinline__ void foam_step_cell(uint32_t i, const float4* __restrict__ disp,
                                               const float4* __restrict__ vel,
                                               const float* __restrict__ cur,
                                               float* __restrict__ next, uint32_t n,
                                               const ocean_hip_foam_params& p)
{
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

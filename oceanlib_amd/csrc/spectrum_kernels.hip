// The time-independent tables on the device: h0 (re/im), omega and the FFT
// twiddles, one lane per cell; the per-cell arithmetic and its certification
// flag are spectrum_cell.hpp's (the device libm through NativeLibm), flagged
// cells are recomputed on the host by the reference's chain (ocean.cpp).
#include "hip_launch.h"
#include "spectrum_cell.hpp"

#include <hip/hip_runtime.h>

#include <stdint.h>

namespace oceanamd {
namespace {

using cell::NativeLibm;
using cell::SpecConst;

__global__ __launch_bounds__(256) void spectrum_kernel(SpecConst c, uint32_t n, float patch_length,
                                                       double dk, uint64_t seed,
                                                       float2* __restrict__ h0,
                                                       float* __restrict__ omega_out,
                                                       uint8_t* __restrict__ flags)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n * n) return;
    float hr, hi, om;
    const bool risky =
        cell::spectrum_cell<NativeLibm>(c, n, patch_length, dk, seed, i % n, i / n, hr, hi, om);
    omega_out[i] = om;
    h0[i] = make_float2(hr, hi);
    flags[i] = risky ? 1u : 0u;
}

__global__ __launch_bounds__(256) void twiddle_kernel(uint32_t n, float2* __restrict__ tw,
                                                      uint8_t* __restrict__ flags)
{
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j >= n / 2) return;
    float cv, sv;
    const bool risky = cell::twiddle<NativeLibm>(n, j, cv, sv);
    tw[j] = make_float2(cv, sv);
    flags[j] = risky ? 1u : 0u;
}

struct Fix {
    uint32_t index;
    float h0r, h0i, omega;
};

__global__ void fix_kernel(const Fix* __restrict__ fixes, uint32_t count, float2* __restrict__ h0,
                           float* __restrict__ omega)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const Fix f = fixes[k];
    h0[f.index] = make_float2(f.h0r, f.h0i);
    omega[f.index] = f.omega;
}

}  // namespace
}  // namespace oceanamd

using namespace oceanamd;

extern "C" hipError_t ocean_hip_spectrum(const ocean_hip_spectrum_params* p, float* h0, float* omega,
                                         uint8_t* flags, hipStream_t s)
{
    if (p->n == 0 || (p->n & (p->n - 1u)) != 0u) return hipErrorInvalidValue;
    SpecConst c{p->gravity, p->depth, p->omega_p, p->alpha, p->gamma, p->wind, p->swell,
                p->wind_direction, p->cutoff_l, p->mu, p->q_den};
    const uint32_t cells = p->n * p->n;
    hipLaunchKernelGGL(spectrum_kernel, dim3((cells + 255u) / 256u), dim3(256), 0, s, c, p->n,
                       p->patch_length, p->dk, p->seed, reinterpret_cast<float2*>(h0), omega, flags);
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_twiddles(uint32_t n, float* tw, uint8_t* flags, hipStream_t s)
{
    if (n < 2) return hipSuccess;
    hipLaunchKernelGGL(twiddle_kernel, dim3((n / 2 + 255u) / 256u), dim3(256), 0, s, n,
                       reinterpret_cast<float2*>(tw), flags);
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_spectrum_fix(const void* fixes, uint32_t count, float* h0, float* omega,
                                             hipStream_t s)
{
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(fix_kernel, dim3((count + 255u) / 256u), dim3(256), 0, s,
                       static_cast<const Fix*>(fixes), count, reinterpret_cast<float2*>(h0), omega);
    return hipGetLastError();
}

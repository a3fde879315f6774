// Renderer handoff: RGBA16F copies of the output buffers, converted on the
// device (reference examples/viewer/half.hpp:58-130, the viewer's upload in
// examples/viewer/ocean_view.cpp:818-862, ARCHITECTURE.md ADR-026).
//
// The conversion is the viewer's float_to_half written as integer operations
// on the float's bits (half.hpp:58-100): round to nearest even with half
// subnormals, values past half's range to infinity, every NaN to the quiet
// NaN sign | 0x7E00. Integer-only, so the result does not depend on the
// hardware's f16 conversion modes. One lane per texel: a float4 in (and the
// alpha replacement, when given), four halves out as one 8-byte store, so a
// wavefront writes 512 contiguous bytes.
#include "hip_launch.h"

#include <hip/hip_runtime.h>

#include <cstdint>

namespace oceanamd {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint32_t half_bits(float value)
{
    const uint32_t x = __float_as_uint(value);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t e32 = (x >> 23) & 0xFFu;
    const uint32_t m32 = x & 0x007FFFFFu;
    if (e32 == 0xFFu) return sign | 0x7C00u | (m32 != 0u ? 0x0200u : 0u);  // inf / NaN
    const int32_t e = int32_t(e32) - 127 + 15;
    if (e >= 0x1F) return sign | 0x7C00u;  // past half's range
    if (e <= 0) {                           // half subnormal, or zero
        if (e < -10) return sign;
        const uint32_t m = m32 | 0x00800000u;
        const uint32_t s = uint32_t(14 - e);  // 14..24
        uint32_t h = m >> s;
        const uint32_t rem = m & ((1u << s) - 1u);
        const uint32_t tie = 1u << (s - 1u);
        if (rem > tie || (rem == tie && (h & 1u) != 0u)) ++h;
        return sign | h;
    }
    // Exponent packed before rounding, so a mantissa carry moves it (up to
    // infinity at the top).
    uint32_t h = (uint32_t(e) << 10) | (m32 >> 13);
    const uint32_t rem = m32 & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u) != 0u)) ++h;
    return sign | h;
}

__device__ __forceinline__ uint2 pack4(float a, float b, float c, float d)
{
    return make_uint2(half_bits(a) | (half_bits(b) << 16), half_bits(c) | (half_bits(d) << 16));
}

// RGB from src, A from alpha[cell] (or src's own w when alpha is null).
__global__ __launch_bounds__(kBlock) void rgba16_kernel(const float4* __restrict__ src,
                                                        const float* __restrict__ alpha,
                                                        uint2* __restrict__ dst, uint32_t cells)
{
    const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= cells) return;
    const float4 v = src[c];
    const float w = alpha != nullptr ? alpha[c] : v.w;
    dst[c] = pack4(v.x, v.y, v.z, w);
}

// Plain conversion of `count` floats: four per lane, the tail one by one.
__global__ __launch_bounds__(kBlock) void half_kernel(const float* __restrict__ src,
                                                      uint16_t* __restrict__ dst, uint64_t count)
{
    const uint64_t q = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    const uint64_t i = q * 4u;
    if (i + 4u <= count) {
        const float4 v = reinterpret_cast<const float4*>(src)[q];
        reinterpret_cast<uint2*>(dst)[q] = pack4(v.x, v.y, v.z, v.w);
    } else {
        for (uint64_t k = i; k < count; ++k) dst[k] = uint16_t(half_bits(src[k]));
    }
}

}  // namespace
}  // namespace oceanamd

using namespace oceanamd;

extern "C" hipError_t ocean_hip_pack_rgba16(const float* src, const float* alpha, uint16_t* dst,
                                            uint32_t cells, hipStream_t s)
{
    if (cells == 0) return hipSuccess;
    hipLaunchKernelGGL(rgba16_kernel, dim3((cells + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       reinterpret_cast<const float4*>(src), alpha, reinterpret_cast<uint2*>(dst),
                       cells);
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_pack_half(const float* src, uint16_t* dst, uint64_t count,
                                          hipStream_t s)
{
    if (count == 0) return hipSuccess;
    const uint64_t lanes = (count + 3u) / 4u;
    const uint64_t blocks = (lanes + kBlock - 1) / kBlock;
    if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(half_kernel, dim3(uint32_t(blocks)), dim3(kBlock), 0, s, src, dst, count);
    return hipGetLastError();
}

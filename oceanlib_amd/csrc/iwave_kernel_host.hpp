// iWave stencil construction (see iwave_kernel_host.cpp).
#pragma once

#include "hip_launch.h"

#include <cstdint>
#include <vector>

namespace oceanamd {

struct IWaveStencil {
    std::uint32_t radius = 0;
    std::uint32_t width = 0;
    float cell = 0.0f;             // the fit's cell size, stored as float
    std::vector<float> taps;       // width*width, row-major (j + P) * width + (i + P)
    double max_symbol = 0.0;       // largest realised symbol over the grid's k-square
    double min_symbol = 0.0;
    double band_low = 0.0;
    double peak_relative_error = 0.0;
};

enum class StencilStatus { Ok, Singular, DeviceError };

// Least-squares fit of radius `radius` (1..12) for `cell` metres and `depth`
// (<= 0 deep), with its normal equations and symbol scan on the current
// device, enqueued on `stream` (waited for before returning). Singular: the
// normal equations are not positive definite.
StencilStatus build_iwave_stencil(std::uint32_t radius, double cell, double depth,
                                  hipStream_t stream, IWaveStencil& out);

// The reference's stability test (src/core/iwave_kernel.hpp:166-169).
inline bool stencil_is_stable(const IWaveStencil& k)
{
    return k.min_symbol >= -1e-3 * k.max_symbol && k.max_symbol > 0.0;
}

}  // namespace oceanamd

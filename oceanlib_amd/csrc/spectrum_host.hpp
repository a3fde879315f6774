// Host construction of the per-ocean constant tables (see spectrum_host.cpp).
#pragma once

#include "ocean/ocean.hpp"

#include <cstdint>

namespace oceanamd {

// Rows [row_begin, row_end) of h0 (2 floats per cell, re/im interleaved) and
// omega (1 float per cell). Pure function of (desc, row): any row split gives
// identical bits.
void build_spectrum_tables(const ocean::OceanDesc& desc, float* h0, float* omega,
                           std::uint32_t row_begin, std::uint32_t row_end) noexcept;

// n/2 twiddles (cos, +sin)(2*pi*j/n), interleaved.
void build_twiddles(std::uint32_t n, float* tw) noexcept;

// Per-descriptor constants of the spectrum, computed by the host libm as the
// per-cell code computes them, for the device build.
struct SpectrumConstants {
    double gravity, depth, omega_p, alpha, gamma, wind, swell, wind_direction, cutoff_l, mu, q_den,
        dk;
};
SpectrumConstants spectrum_constants(const ocean::OceanDesc& desc) noexcept;
// One cell of the tables by the host chain (h0: 2 floats), and one twiddle.
void spectrum_cell(const ocean::OceanDesc& desc, std::uint32_t x, std::uint32_t y, float* h0,
                   float* omega) noexcept;
void twiddle_at(std::uint32_t n, std::uint32_t j, float* tw) noexcept;

}  // namespace oceanamd

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

}  // namespace oceanamd

// Host-side construction of the time-independent tables the per-frame
// kernels read: h0 (interleaved re/im), omega, and the FFT twiddles.
//
// Runs once per ocean at creation, in double with the host libm, in the
// reference's operation order, so the tables are bit-identical to the
// reference's (src/core/spectrum.cpp:264-339, src/core/fft.cpp:24-33). Only
// h0 and omega are kept: kx, kz and 1/|k| are recomputed on the device from
// the cell index, and conj(h0(-k)) is read from the mirrored h0 cell.
#include "spectrum_host.hpp"

#include <cmath>
#include <cstdint>

namespace oceanamd {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 6.28318530717958647692;

// ---- PCG32 / SplitMix64 / Box-Muller: reference src/core/rng.hpp:27-97,
// src/core/rng.cpp:7-17 ---------------------------------------------------

struct Pcg32 {
    std::uint64_t state = 0;
    std::uint64_t inc = 1;
};

inline std::uint32_t next_u32(Pcg32& r) noexcept
{
    const std::uint64_t old = r.state;
    r.state = old * 6364136223846793005ULL + r.inc;
    const auto xs = static_cast<std::uint32_t>(((old >> 18u) ^ old) >> 27u);
    const auto rot = static_cast<std::uint32_t>(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

inline Pcg32 seed_pcg32(std::uint64_t seed, std::uint64_t stream) noexcept
{
    Pcg32 r;
    r.state = 0u;
    r.inc = (stream << 1u) | 1u;
    next_u32(r);
    r.state += seed;
    next_u32(r);
    return r;
}

inline double next_open01(Pcg32& r) noexcept
{
    return (static_cast<double>(next_u32(r)) + 0.5) * (1.0 / 4294967296.0);
}

inline std::uint64_t mix64(std::uint64_t x) noexcept
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

inline Pcg32 row_rng(std::uint64_t seed, std::uint32_t row) noexcept
{
    return seed_pcg32(mix64(seed ^ (0x9E3779B97F4A7C15ULL * (row + 1u))), 0);
}

inline void gaussian_pair(Pcg32& r, float& a, float& b) noexcept
{
    const double u1 = next_open01(r);
    const double u2 = next_open01(r);
    const double radius = std::sqrt(-2.0 * std::log(u1));
    const double theta = 6.283185307179586476925286766559 * u2;
    a = static_cast<float>(radius * std::cos(theta));
    b = static_cast<float>(radius * std::sin(theta));
}

// ---- spectrum: reference src/core/spectrum.cpp:14-244 -------------------

double wrap_angle(double a) noexcept
{
    a = std::fmod(a + kPi, kTwoPi);
    if (a < 0.0) a += kTwoPi;
    return a - kPi;
}

double dispersion_omega(double gravity, double k, double depth) noexcept
{
    if (!(k > 0.0)) return 0.0;
    if (depth <= 0.0) return std::sqrt(gravity * k);
    return std::sqrt(gravity * k * std::tanh(k * depth));
}

double dispersion_domega_dk(double gravity, double k, double omega, double depth) noexcept
{
    if (!(omega > 0.0)) return 0.0;
    if (depth <= 0.0) return gravity / (2.0 * omega);
    const double kh = k * depth;
    const double th = std::tanh(kh);
    const double sech2 = 1.0 - th * th;
    return gravity * (th + kh * sech2) / (2.0 * omega);
}

double peak_omega(const ocean::SpectrumDesc& d) noexcept
{
    const double g = d.gravity;
    const double u = static_cast<double>(d.wind_speed);
    const double f = static_cast<double>(d.fetch);
    if (u <= 0.0 || f <= 0.0) return 0.0;
    return 22.0 * std::cbrt(g * g / (u * f));
}

double alpha_of(const ocean::SpectrumDesc& d) noexcept
{
    const double g = d.gravity;
    const double u = static_cast<double>(d.wind_speed);
    const double f = static_cast<double>(d.fetch);
    if (u <= 0.0 || f <= 0.0) return 0.0;
    const double x_tilde = g * f / (u * u);
    return 0.076 * std::pow(x_tilde, -0.22);
}

double jonswap(const ocean::SpectrumDesc& d, double omega) noexcept
{
    if (!(omega > 0.0)) return 0.0;
    const double g = d.gravity;
    const double omega_p = peak_omega(d);
    if (!(omega_p > 0.0)) return 0.0;
    const double alpha = alpha_of(d);
    const double gamma = static_cast<double>(d.peak_enhancement);
    const double wp_over_w = omega_p / omega;
    const double w2 = omega * omega;
    const double w4 = w2 * w2;
    const double pm = (alpha * g * g / (w4 * omega)) * std::exp(-1.25 * std::pow(wp_over_w, 4.0));
    const double sigma = (omega <= omega_p) ? 0.07 : 0.09;
    const double dw = omega - omega_p;
    const double r = std::exp(-(dw * dw) / (2.0 * sigma * sigma * omega_p * omega_p));
    return pm * std::pow(gamma, r);
}

double directional_spread(const ocean::SpectrumDesc& d, double omega, double theta) noexcept
{
    const double omega_p = peak_omega(d);
    if (!(omega > 0.0) || !(omega_p > 0.0)) return 0.0;
    const double g = d.gravity;
    const double u = static_cast<double>(d.wind_speed);
    const double ratio = omega / omega_p;
    double s;
    if (ratio < 1.0) {
        s = 6.97 * std::pow(ratio, 4.06);
    } else {
        const double mu = -2.33 - 1.45 * ((u * omega_p / g) - 1.17);
        s = 9.77 * std::pow(ratio, mu);
    }
    const double swell = static_cast<double>(d.swell);
    if (swell > 0.0) s += 16.0 * std::tanh(omega_p / omega) * swell * swell;
    if (s < 0.0) s = 0.0;
    const double dtheta = wrap_angle(theta - static_cast<double>(d.wind_direction));
    const double c = std::cos(0.5 * dtheta);
    if (c <= 0.0) return 0.0;
    const double q = std::exp(std::lgamma(s + 1.0) - std::lgamma(s + 0.5)) / (2.0 * std::sqrt(kPi));
    return q * std::pow(c, 2.0 * s);
}

double wave_density(const ocean::SpectrumDesc& d, double kx, double kz) noexcept
{
    const double k2 = kx * kx + kz * kz;
    if (!(k2 > 0.0)) return 0.0;
    const double k = std::sqrt(k2);
    const double g = d.gravity;
    const double depth = static_cast<double>(d.depth);
    const double omega = dispersion_omega(g, k, depth);
    const double s_omega = jonswap(d, omega);
    if (!(s_omega > 0.0)) return 0.0;
    const double domega_dk = dispersion_domega_dk(g, k, omega, depth);
    const double theta = std::atan2(kz, kx);
    const double spread = directional_spread(d, omega, theta);
    double psi = s_omega * domega_dk * spread / k;
    const double l = static_cast<double>(d.small_wave_cutoff) / kTwoPi;
    psi *= std::exp(-k2 * l * l);
    return psi;
}

inline float wave_component(std::uint32_t i, std::uint32_t n, float patch_length) noexcept
{
    constexpr float kTwoPiF = 6.28318530717958647692f;
    const int si = (i < n / 2) ? static_cast<int>(i) : static_cast<int>(i) - static_cast<int>(n);
    return kTwoPiF * static_cast<float>(si) / patch_length;
}

}  // namespace

SpectrumConstants spectrum_constants(const ocean::OceanDesc& desc) noexcept
{
    const ocean::SpectrumDesc& d = desc.spectrum;
    SpectrumConstants c{};
    c.gravity = d.gravity;
    c.depth = static_cast<double>(d.depth);
    c.omega_p = peak_omega(d);
    c.alpha = alpha_of(d);
    c.gamma = static_cast<double>(d.peak_enhancement);
    c.wind = static_cast<double>(d.wind_speed);
    c.swell = static_cast<double>(d.swell);
    c.wind_direction = static_cast<double>(d.wind_direction);
    c.cutoff_l = static_cast<double>(d.small_wave_cutoff) / kTwoPi;
    // the per-cell expression of directional_spread, hoisted (same operations)
    c.mu = -2.33 - 1.45 * ((c.wind * c.omega_p / c.gravity) - 1.17);
    c.q_den = 2.0 * std::sqrt(kPi);
    c.dk = kTwoPi / static_cast<double>(desc.patch_length);
    return c;
}

void spectrum_cell(const ocean::OceanDesc& desc, std::uint32_t x, std::uint32_t y, float* h0,
                   float* omega) noexcept
{
    const std::uint32_t n = desc.size;
    const float L = desc.patch_length;
    const double dk = kTwoPi / static_cast<double>(L);
    Pcg32 rng = row_rng(desc.seed, y);
    for (std::uint32_t k = 0; k < 2u * x; ++k) next_u32(rng);
    const float kz_row = wave_component(y, n, L);
    const float kx_val = wave_component(x, n, L);
    const double k2 = static_cast<double>(kx_val) * kx_val + static_cast<double>(kz_row) * kz_row;
    const double k = std::sqrt(k2);
    *omega = static_cast<float>(
        dispersion_omega(desc.spectrum.gravity, k, static_cast<double>(desc.spectrum.depth)));
    const double psi = wave_density(desc.spectrum, kx_val, kz_row);
    const double amp = 0.5 * std::sqrt(psi) * dk;
    float ga, gb;
    gaussian_pair(rng, ga, gb);
    h0[0] = static_cast<float>(ga * amp);
    h0[1] = static_cast<float>(gb * amp);
}

void twiddle_at(std::uint32_t n, std::uint32_t j, float* tw) noexcept
{
    constexpr double k2pi = 6.283185307179586476925286766559;
    const double angle = k2pi * static_cast<double>(j) / static_cast<double>(n);
    tw[0] = static_cast<float>(std::cos(angle));
    tw[1] = static_cast<float>(std::sin(angle));
}

void build_spectrum_tables(const ocean::OceanDesc& desc, float* h0, float* omega,
                           std::uint32_t row_begin, std::uint32_t row_end) noexcept
{
    const std::uint32_t n = desc.size;
    const float L = desc.patch_length;
    const double dk = kTwoPi / static_cast<double>(L);
    for (std::uint32_t y = row_begin; y < row_end; ++y) {
        Pcg32 rng = row_rng(desc.seed, y);
        const float kz_row = wave_component(y, n, L);
        for (std::uint32_t x = 0; x < n; ++x) {
            const std::size_t i = static_cast<std::size_t>(y) * n + x;
            const float kx_val = wave_component(x, n, L);
            const double k2 = static_cast<double>(kx_val) * kx_val +
                              static_cast<double>(kz_row) * kz_row;
            const double k = std::sqrt(k2);
            omega[i] = static_cast<float>(dispersion_omega(
                desc.spectrum.gravity, k, static_cast<double>(desc.spectrum.depth)));
            const double psi = wave_density(desc.spectrum, kx_val, kz_row);
            const double amp = 0.5 * std::sqrt(psi) * dk;
            float ga, gb;
            gaussian_pair(rng, ga, gb);
            h0[2 * i + 0] = static_cast<float>(ga * amp);
            h0[2 * i + 1] = static_cast<float>(gb * amp);
        }
    }
}

void build_twiddles(std::uint32_t n, float* tw) noexcept
{
    constexpr double k2pi = 6.283185307179586476925286766559;
    for (std::uint32_t j = 0; j < n / 2; ++j) {
        const double angle = k2pi * static_cast<double>(j) / static_cast<double>(n);
        tw[2 * j + 0] = static_cast<float>(std::cos(angle));
        tw[2 * j + 1] = static_cast<float>(std::sin(angle));
    }
}

}  // namespace oceanamd

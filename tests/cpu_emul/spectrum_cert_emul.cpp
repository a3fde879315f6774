// CPU emulation of the device tables' certification: the device kernels'
// own per-cell code (oceanlib_amd/csrc/spectrum_cell.hpp) compiled for the
// host, with a libm whose every result (log, exp, pow, tanh, atan2, cos, sin,
// lgamma) is moved by a random number of double ulps in [-U, U] from the host
// libm's: a stand-in for a device libm that differs from the host's by up to
// U ulps per call. Beside it, the host chain the library repairs flagged
// cells with (spectrum_host.cpp). Used by
// tests/test_table_certification_emulation.py.
#include "../../oceanlib_amd/csrc/spectrum_cell.hpp"
#include "../../oceanlib_amd/csrc/spectrum_host.hpp"

#include <cmath>
#include <cstdint>
#include <cstring>

using namespace oceanamd;

namespace {

// SplitMix64 stream of the perturbations.
uint64_t g_state = 0;
int g_ulps = 0;

uint64_t next_draw()
{
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// r moved by k ulps (k in [-U, U], U itself half of the time so the bound is
// met at its edge); zeros, infinities and NaNs stay (every libm returns the
// exact results it can represent exactly at the points that give them).
double perturb(double r)
{
    if (g_ulps == 0 || r == 0.0 || !std::isfinite(r)) return r;
    const uint64_t d = next_draw();
    int64_t k;
    if (d & 1u)
        k = (d & 2u) ? g_ulps : -g_ulps;
    else
        k = int64_t((d >> 8) % uint64_t(2 * g_ulps + 1)) - g_ulps;
    int64_t bits;
    std::memcpy(&bits, &r, sizeof bits);
    bits += k;  // the magnitude's ulps, either sign
    double out;
    std::memcpy(&out, &bits, sizeof out);
    return std::isfinite(out) ? out : r;
}

struct PerturbedLibm {
    static double log(double x) { return perturb(std::log(x)); }
    static double exp(double x) { return perturb(std::exp(x)); }
    static double pow(double x, double y) { return perturb(std::pow(x, y)); }
    static double tanh(double x) { return perturb(std::tanh(x)); }
    static double atan2(double y, double x) { return perturb(std::atan2(y, x)); }
    static double cos(double x) { return perturb(std::cos(x)); }
    static double sin(double x) { return perturb(std::sin(x)); }
    static double lgamma(double x) { return perturb(std::lgamma(x)); }
};

ocean::OceanDesc make_desc(uint32_t n, float patch_length, uint64_t seed, const float* spec)
{
    ocean::OceanDesc d;
    d.size = n;
    d.patch_length = patch_length;
    d.seed = seed;
    d.spectrum.wind_speed = spec[0];
    d.spectrum.fetch = spec[1];
    d.spectrum.wind_direction = spec[2];
    d.spectrum.peak_enhancement = spec[3];
    d.spectrum.swell = spec[4];
    d.spectrum.small_wave_cutoff = spec[5];
    d.spectrum.gravity = spec[6];
    d.spectrum.depth = spec[7];
    return d;
}

}  // namespace

// spec: wind_speed, fetch, wind_direction, peak_enhancement, swell,
// small_wave_cutoff, gravity, depth. Outputs per cell: the emulated device
// cell (h0 re/im interleaved, omega, flag) under perturbations of up to
// `ulps` ulps from the stream `pseed`, and the host chain's cell.
extern "C" int emul_tables(uint32_t n, float patch_length, uint64_t seed, const float* spec, int ulps,
                           uint64_t pseed, float* h0, float* omega, uint8_t* flags, float* h0_host,
                           float* omega_host)
{
    if (n < 2 || (n & (n - 1u)) != 0u || ulps < 0) return -1;
    const ocean::OceanDesc d = make_desc(n, patch_length, seed, spec);
    const SpectrumConstants k = spectrum_constants(d);
    const cell::SpecConst c{k.gravity, k.depth, k.omega_p, k.alpha, k.gamma, k.wind, k.swell,
                            k.wind_direction, k.cutoff_l, k.mu, k.q_den};
    g_ulps = ulps;
    for (uint32_t i = 0; i < n * n; ++i) {
        g_state = pseed ^ (0xD1B54A32D192ED03ULL * (uint64_t(i) + 1u));  // a stream per cell
        const bool risky = cell::spectrum_cell<PerturbedLibm>(c, n, d.patch_length, k.dk, d.seed, i % n,
                                                              i / n, h0[2 * i], h0[2 * i + 1], omega[i]);
        flags[i] = risky ? 1u : 0u;
    }
    g_ulps = 0;
    build_spectrum_tables(d, h0_host, omega_host, 0, n);
    return 0;
}

// The n/2 twiddles (cos, sin interleaved) the same way.
extern "C" int emul_twiddles(uint32_t n, int ulps, uint64_t pseed, float* tw, uint8_t* flags,
                             float* tw_host)
{
    if (n < 2 || (n & (n - 1u)) != 0u || ulps < 0) return -1;
    g_state = pseed;
    g_ulps = ulps;
    for (uint32_t j = 0; j < n / 2; ++j)
        flags[j] = cell::twiddle<PerturbedLibm>(n, j, tw[2 * j], tw[2 * j + 1]) ? 1u : 0u;
    g_ulps = 0;
    build_twiddles(n, tw_host);
    return 0;
}

// Construction of the iWave convolution stencil, once per interaction field,
// on the host: the reference's least-squares symbol fit (the solver's method,
// src/core/iwave_kernel.cpp:177-437, LeastSquares branch) in the same double
// operation order, so the float taps the step kernel reads are bit-identical
// to the reference's, and the realised-symbol scan that yields the stability
// limit and the dispersion error.
#include "iwave_kernel_host.hpp"

#include <cmath>
#include <vector>

namespace oceanamd {
namespace {

constexpr double kPi = 3.14159265358979323846;

// Number of cells in the D4 orbit of (a, b), 0 <= b <= a (iwave_kernel.cpp:72-77).
int orbit_size(int a, int b) noexcept
{
    if (a == 0) return 1;
    if (b == 0) return 4;
    if (a == b) return 4;
    return 8;
}

// The orbit's cells in the reference's enumeration order (iwave_kernel.cpp:79-108).
int orbit_of(int a, int b, int (*out)[2]) noexcept
{
    int n = 0;
    if (a == 0) {
        out[n][0] = 0;
        out[n][1] = 0;
        return 1;
    }
    if (b == 0) {
        const int v[4][2] = {{a, 0}, {-a, 0}, {0, a}, {0, -a}};
        for (auto& c : v) {
            out[n][0] = c[0];
            out[n][1] = c[1];
            ++n;
        }
        return n;
    }
    const int sx[2] = {a, -a};
    const int sz[2] = {b, -b};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            out[n][0] = sx[i];
            out[n][1] = sz[j];
            ++n;
        }
    if (a != b)
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) {
                out[n][0] = sz[j];
                out[n][1] = sx[i];
                ++n;
            }
    return n;
}

// In-place Cholesky solve of the normal equations (iwave_kernel.cpp:110-150).
bool cholesky_solve(std::vector<double>& a, std::vector<double>& b, int n)
{
    const std::size_t sn = static_cast<std::size_t>(n);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j <= i; ++j) {
            double s = a[i * sn + j];
            for (int k = 0; k < j; ++k) s -= a[i * sn + k] * a[j * sn + k];
            if (i == j) {
                if (!(s > 0.0)) return false;
                a[i * sn + i] = std::sqrt(s);
            } else {
                a[i * sn + j] = s / a[j * sn + j];
            }
        }
    }
    for (int i = 0; i < n; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= a[i * sn + k] * b[k];
        b[i] = s / a[i * sn + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int k = i + 1; k < n; ++k) s -= a[k * sn + i] * b[k];
        b[i] = s / a[i * sn + i];
    }
    return true;
}

double symbol(double k, double depth) noexcept
{
    if (k <= 0.0) return 0.0;
    if (depth <= 0.0) return k;
    return k * std::tanh(k * depth);
}

// sum_ij G_ij cos(k . x_ij), with the float-stored cell (iwave_kernel.cpp:161-173)
double realised_symbol(const IWaveStencil& k, double kx, double kz) noexcept
{
    const int p = static_cast<int>(k.radius);
    const double dx = k.cell;
    double sum = 0.0;
    for (int j = -p; j <= p; ++j)
        for (int i = -p; i <= p; ++i)
            sum += static_cast<double>(k.taps[static_cast<std::size_t>(j + p) * k.width + (i + p)]) *
                   std::cos(kx * i * dx + kz * j * dx);
    return sum;
}

}  // namespace

bool build_iwave_stencil(std::uint32_t radius, double cell, double depth, IWaveStencil& out)
{
    constexpr int kFitSamples = 96;
    constexpr double kBandLowFrac = 0.125;
    constexpr double kOutOfBandWeight = 0.02;
    const int p = static_cast<int>(radius);
    const int width = 2 * p + 1;
    const double k_max = kPi / cell;
    out.radius = radius;
    out.width = static_cast<std::uint32_t>(width);
    out.cell = static_cast<float>(cell);
    out.taps.assign(static_cast<std::size_t>(width) * width, 0.0f);

    std::vector<int> ai, bi;
    for (int a = 0; a <= p; ++a)
        for (int b = 0; b <= a; ++b) {
            ai.push_back(a);
            bi.push_back(b);
        }
    const int m = static_cast<int>(ai.size());
    std::vector<double> g(static_cast<std::size_t>(m), 0.0);

    const int mm = m - 1;  // the centre tap is eliminated by the zero-sum constraint
    const std::size_t smm = static_cast<std::size_t>(mm);
    std::vector<double> ata(smm * smm, 0.0), atb(smm, 0.0), psi(smm, 0.0);
    const double k_lo = kBandLowFrac * k_max;
    int orb[8][2];
    for (int zi = 0; zi <= kFitSamples; ++zi) {
        for (int xi = 0; xi <= kFitSamples; ++xi) {
            const double kx = k_max * xi / kFitSamples;
            const double kz = k_max * zi / kFitSamples;
            const double kk = std::sqrt(kx * kx + kz * kz);
            const double s = symbol(kk, depth);
            const bool in_band = (kk >= k_lo && kk <= k_max);
            const double ref = in_band ? s : symbol((kk < k_lo) ? k_lo : k_max, depth);
            double w = 1.0 / (ref * ref);
            if (!in_band) w *= kOutOfBandWeight;
            for (int t = 1; t < m; ++t) {
                const int n = orbit_of(ai[t], bi[t], orb);
                double phi = 0.0;
                for (int e = 0; e < n; ++e)
                    phi += std::cos(kx * orb[e][0] * cell + kz * orb[e][1] * cell);
                psi[t - 1] = phi - n;
            }
            for (int r = 0; r < mm; ++r) {
                const double wr = w * psi[r];
                atb[r] += wr * s;
                for (int c = 0; c <= r; ++c) ata[r * smm + c] += wr * psi[c];
            }
        }
    }
    double trace = 0.0;
    for (int r = 0; r < mm; ++r) trace += ata[r * smm + r];
    const double ridge = 1e-12 * trace / mm;
    for (int r = 0; r < mm; ++r) {
        ata[r * smm + r] += ridge;
        for (int c = r + 1; c < mm; ++c) ata[r * smm + c] = ata[c * smm + r];
    }
    if (!cholesky_solve(ata, atb, mm)) return false;
    double sum_rest = 0.0;
    for (int t = 1; t < m; ++t) {
        g[t] = atb[t - 1];
        sum_rest += orbit_size(ai[t], bi[t]) * g[t];
    }
    g[0] = -sum_rest;

    for (int t = 0; t < m; ++t) {
        const int n = orbit_of(ai[t], bi[t], orb);
        const float v = static_cast<float>(g[t]);
        for (int e = 0; e < n; ++e)
            out.taps[static_cast<std::size_t>(orb[e][1] + p) * width + (orb[e][0] + p)] = v;
    }

    constexpr int kScan = 192;
    double worst = 0.0, peak = 0.0, trough = 0.0;
    out.band_low = kBandLowFrac * k_max;
    for (int zi = 0; zi <= kScan; ++zi) {
        for (int xi = 0; xi <= kScan; ++xi) {
            const double kx = k_max * xi / kScan;
            const double kz = k_max * zi / kScan;
            const double kk = std::sqrt(kx * kx + kz * kz);
            const double rs = realised_symbol(out, kx, kz);
            if (rs > peak) peak = rs;
            if (rs < trough) trough = rs;
            if (kk >= out.band_low && kk <= k_max) {
                const double tr = symbol(kk, depth);
                if (tr > 0.0) {
                    const double e = std::fabs(rs / tr - 1.0);
                    if (e > worst) worst = e;
                }
            }
        }
    }
    out.max_symbol = peak;
    out.min_symbol = trough;
    out.peak_relative_error = worst;
    return true;
}

}  // namespace oceanamd

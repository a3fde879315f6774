// Construction of the iWave convolution stencil, once per interaction field:
// the reference's least-squares symbol fit (src/core/iwave_kernel.cpp:177-437,
// LeastSquares branch) with the float taps bit-identical to the reference's,
// and the realised-symbol scan that yields the stability limit and the
// dispersion error.
//
// Where each part runs:
//   - the basis table psi (sample-major, one row of m - 1 values per fit
//     sample) and the per-sample weight and symbol: host threads, because
//     every value carries glibc's double cos / tanh bits into the sums;
//   - the normal equations A^T W A (lower triangle) and A^T W s: the device
//     (fit_normal_kernel), one lane per entry, summing the samples in the
//     reference's order with its roundings;
//   - ridge, Cholesky solve, the eliminated centre tap and the D4 expansion:
//     the host (a few thousand flops, in the reference's order);
//   - the realised symbol on the 193 x 193 grid: the device
//     (symbol_scan_kernel, device cos). Its max, min and peak relative error
//     are then certified: both evaluations sum the same 169..625 products in
//     the same order and differ only through the cosines, so they are within
//     B = 1e-12 * sum |taps| of each other (the test
//     test_symbol_scan_within_bound measures the actual gap, below 1e-15 *
//     sum |taps|); every point whose device value could be the extreme within
//     2B is recomputed on the host with glibc, and the extremes are taken
//     over those, which gives the reference's exact values.
#include "iwave_kernel_host.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

namespace oceanamd {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kFitSamples = 96;      // iwave_kernel.hpp fit_samples default
constexpr double kBandLowFrac = 0.125;
constexpr double kOutOfBandWeight = 0.02;
constexpr int kScan = 192;

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

// f(first, last) over [0, count) split across host threads.
template <typename F>
void parallel_ranges(std::size_t count, F&& f)
{
    unsigned hw = std::thread::hardware_concurrency();
    std::size_t threads = std::min<std::size_t>(hw == 0 ? 1u : hw, 16u);
    threads = std::max<std::size_t>(1u, std::min(threads, count / 64u));
    if (threads <= 1) {
        f(std::size_t(0), count);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(threads - 1);
    const std::size_t per = (count + threads - 1) / threads;
    for (std::size_t t = 1; t < threads; ++t) {
        const std::size_t a = t * per, b = std::min(count, a + per);
        if (a < b) pool.emplace_back([&f, a, b] { f(a, b); });
    }
    f(std::size_t(0), std::min(count, per));
    for (auto& th : pool) th.join();
}

struct DeviceScratch {
    void* p = nullptr;
    ~DeviceScratch()
    {
        if (p != nullptr) (void)hipFree(p);
    }
};

}  // namespace

StencilStatus build_iwave_stencil(std::uint32_t radius, double cell, double depth,
                                  hipStream_t stream, IWaveStencil& out)
{
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
    const int mm = m - 1;  // the centre tap is eliminated by the zero-sum constraint
    if (mm < 1 || mm > OCEAN_HIP_IWAVE_MAX_BASIS) return StencilStatus::DeviceError;
    const std::size_t smm = static_cast<std::size_t>(mm);
    const std::size_t samples = std::size_t(kFitSamples + 1) * (kFitSamples + 1);
    const std::size_t tri = smm * (smm + 1) / 2;
    const std::size_t scan_pts = std::size_t(kScan + 1) * (kScan + 1);

    // The basis table, weights and symbols (iwave_kernel.cpp:286-326), per sample.
    std::vector<double> psi(samples * smm), wv(samples), sv(samples);
    const double k_lo = kBandLowFrac * k_max;
    parallel_ranges(samples, [&](std::size_t first, std::size_t last) {
        int orb[8][2];
        for (std::size_t sidx = first; sidx < last; ++sidx) {
            const int zi = static_cast<int>(sidx / (kFitSamples + 1));
            const int xi = static_cast<int>(sidx % (kFitSamples + 1));
            const double kx = k_max * xi / kFitSamples;
            const double kz = k_max * zi / kFitSamples;
            const double kk = std::sqrt(kx * kx + kz * kz);
            const double s = symbol(kk, depth);
            const bool in_band = (kk >= k_lo && kk <= k_max);
            const double ref = in_band ? s : symbol((kk < k_lo) ? k_lo : k_max, depth);
            double w = 1.0 / (ref * ref);
            if (!in_band) w *= kOutOfBandWeight;
            double* row = psi.data() + sidx * smm;
            for (int t = 1; t < m; ++t) {
                const int n = orbit_of(ai[t], bi[t], orb);
                double phi = 0.0;
                for (int e = 0; e < n; ++e)
                    phi += std::cos(kx * orb[e][0] * cell + kz * orb[e][1] * cell);
                row[t - 1] = phi - n;
            }
            wv[sidx] = w;
            sv[sidx] = s;
        }
    });

    // Device: the normal equations.
    auto up = [](std::size_t v) { return (v + 255) / 256 * 256; };
    const std::size_t o_psi = 0;
    const std::size_t o_w = up(o_psi + samples * smm * 8);
    const std::size_t o_s = up(o_w + samples * 8);
    const std::size_t o_ata = up(o_s + samples * 8);
    const std::size_t o_atb = up(o_ata + tri * 8);
    const std::size_t o_taps = up(o_atb + smm * 8);
    const std::size_t o_scan = up(o_taps + out.taps.size() * 4);
    const std::size_t bytes = up(o_scan + scan_pts * 16);
    DeviceScratch scratch;
    if (hipMalloc(&scratch.p, bytes) != hipSuccess) {
        scratch.p = nullptr;
        return StencilStatus::DeviceError;
    }
    char* base = static_cast<char*>(scratch.p);
    auto dptr = [&](std::size_t o) { return reinterpret_cast<double*>(base + o); };
    std::vector<double> ata_lower(tri), atb(smm);
    bool ok = hipMemcpyAsync(base + o_psi, psi.data(), samples * smm * 8, hipMemcpyHostToDevice,
                             stream) == hipSuccess &&
              hipMemcpyAsync(base + o_w, wv.data(), samples * 8, hipMemcpyHostToDevice, stream) ==
                  hipSuccess &&
              hipMemcpyAsync(base + o_s, sv.data(), samples * 8, hipMemcpyHostToDevice, stream) ==
                  hipSuccess &&
              ocean_hip_iwave_fit_normal(dptr(o_psi), dptr(o_w), dptr(o_s),
                                         static_cast<uint32_t>(samples), static_cast<uint32_t>(mm),
                                         dptr(o_ata), dptr(o_atb), stream) == hipSuccess &&
              hipMemcpyAsync(ata_lower.data(), base + o_ata, tri * 8, hipMemcpyDeviceToHost,
                             stream) == hipSuccess &&
              hipMemcpyAsync(atb.data(), base + o_atb, smm * 8, hipMemcpyDeviceToHost, stream) ==
                  hipSuccess &&
              hipStreamSynchronize(stream) == hipSuccess;
    if (!ok) return StencilStatus::DeviceError;

    // Host: ridge, mirror, solve (iwave_kernel.cpp:337-362).
    std::vector<double> ata(smm * smm, 0.0);
    for (std::size_t r = 0; r < smm; ++r)
        for (std::size_t c = 0; c <= r; ++c) ata[r * smm + c] = ata_lower[r * (r + 1) / 2 + c];
    double trace = 0.0;
    for (int r = 0; r < mm; ++r) trace += ata[r * smm + r];
    const double ridge = 1e-12 * trace / mm;
    for (int r = 0; r < mm; ++r) {
        ata[r * smm + r] += ridge;
        for (int c = r + 1; c < mm; ++c) ata[r * smm + c] = ata[c * smm + r];
    }
    if (!cholesky_solve(ata, atb, mm)) return StencilStatus::Singular;
    std::vector<double> g(static_cast<std::size_t>(m), 0.0);
    double sum_rest = 0.0;
    for (int t = 1; t < m; ++t) {
        g[t] = atb[t - 1];
        sum_rest += orbit_size(ai[t], bi[t]) * g[t];
    }
    g[0] = -sum_rest;
    int orb[8][2];
    for (int t = 0; t < m; ++t) {
        const int n = orbit_of(ai[t], bi[t], orb);
        const float v = static_cast<float>(g[t]);
        for (int e = 0; e < n; ++e)
            out.taps[static_cast<std::size_t>(orb[e][1] + p) * width + (orb[e][0] + p)] = v;
    }

    // Device: the realised symbol on the scan grid (iwave_kernel.cpp:408-432).
    std::vector<double> scan(scan_pts * 2);
    ok = hipMemcpyAsync(base + o_taps, out.taps.data(), out.taps.size() * 4, hipMemcpyHostToDevice,
                        stream) == hipSuccess &&
         ocean_hip_iwave_symbol_scan(reinterpret_cast<const float*>(base + o_taps), radius,
                                     static_cast<double>(out.cell), k_max, depth, kScan,
                                     dptr(o_scan), stream) == hipSuccess &&
         hipMemcpyAsync(scan.data(), base + o_scan, scan_pts * 16, hipMemcpyDeviceToHost, stream) ==
             hipSuccess &&
         hipStreamSynchronize(stream) == hipSuccess;
    if (!ok) return StencilStatus::DeviceError;

    // Certified extremes: see the file comment for the bound B.
    double abs_taps = 0.0;
    for (float t : out.taps) abs_taps += std::fabs(static_cast<double>(t));
    const double B = 1e-12 * abs_taps;
    out.band_low = kBandLowFrac * k_max;
    auto kvec = [&](std::size_t pt, double& kx, double& kz) {
        const int zi = static_cast<int>(pt / (kScan + 1)), xi = static_cast<int>(pt % (kScan + 1));
        kx = k_max * xi / kScan;
        kz = k_max * zi / kScan;
    };
    double dmax = -INFINITY, dmin = INFINITY, elo = -INFINITY;
    std::vector<std::uint8_t> in_band(scan_pts, 0);
    std::vector<double> ebound(scan_pts, 0.0), edev(scan_pts, 0.0);
    for (std::size_t pt = 0; pt < scan_pts; ++pt) {
        const double rs = scan[2 * pt], tr = scan[2 * pt + 1];
        if (!std::isfinite(rs)) return StencilStatus::DeviceError;
        dmax = std::max(dmax, rs);
        dmin = std::min(dmin, rs);
        double kx, kz;
        kvec(pt, kx, kz);
        const double kk = std::sqrt(kx * kx + kz * kz);
        if (!(kk >= out.band_low && kk <= k_max)) continue;
        in_band[pt] = 1;
        if (!(tr > 1e-200) || !std::isfinite(tr)) {
            ebound[pt] = INFINITY;  // always recomputed
            edev[pt] = 0.0;
            continue;
        }
        const double q = rs / tr;
        edev[pt] = std::fabs(q - 1.0);
        ebound[pt] = 1.01 * B / tr + 3e-13 * (std::fabs(q) + 1.0);
        elo = std::max(elo, edev[pt] - ebound[pt]);
    }
    double peak = 0.0, trough = 0.0, worst = 0.0;
    for (std::size_t pt = 0; pt < scan_pts; ++pt) {
        const double rsd = scan[2 * pt];
        const bool c_peak = rsd >= dmax - 2.0 * B;
        const bool c_trough = rsd <= dmin + 2.0 * B;
        const bool c_worst = in_band[pt] && edev[pt] + ebound[pt] >= elo;
        if (!c_peak && !c_trough && !c_worst) continue;
        double kx, kz;
        kvec(pt, kx, kz);
        const double rs = realised_symbol(out, kx, kz);
        if (c_peak && rs > peak) peak = rs;
        if (c_trough && rs < trough) trough = rs;
        if (c_worst) {
            const double kk = std::sqrt(kx * kx + kz * kz);
            const double tr = symbol(kk, depth);
            if (tr > 0.0) {
                const double e = std::fabs(rs / tr - 1.0);
                if (e > worst) worst = e;
            }
        }
    }
    out.max_symbol = peak;
    out.min_symbol = trough;
    out.peak_relative_error = worst;
    return StencilStatus::Ok;
}

}  // namespace oceanamd

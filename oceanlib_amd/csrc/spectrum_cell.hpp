// One cell of the time-independent tables (h0 re/im and omega) and one FFT
// twiddle, with the certification flag of each: the device build's per-lane
// code (spectrum_kernels.hip), written once for the device and for the CPU
// emulation that checks the certification (tests/cpu_emul).
//
// The arithmetic restates the host build (spectrum_host.cpp, reference
// src/core/spectrum.cpp:14-339, src/core/rng.hpp:27-97, src/core/rng.cpp:7-17,
// src/core/fft.cpp:24-33) operation for operation. The libm calls whose
// results may differ between two libms (log, exp, pow, tanh, atan2, cos, sin,
// lgamma) go through the policy M; sqrt, fmod, floor and the arithmetic are
// correctly rounded or exact everywhere and are called directly.
//
// The reference's tables come from the host libm (glibc); the device libm
// can differ from it in the last bits of a double. Those differences reach a
// stored float only where a double lies next to a float rounding midpoint, or
// where a libm result decides a branch. Each cell therefore checks every
// double -> float rounding it performs (the two Gaussian draws, omega, h0 re
// and im) and every libm-dependent branch (omega vs omega_p, omega / omega_p
// vs 1, the spreading cosine vs 0), and flags the cell when the value lies
// closer to the decision than 4x a bound on how far the two libms' doubles
// can be apart there. The bound charges every libm call U = 8 double ulps
// (device and host error together) and carries it through the cell's own
// amplification factors (exp arguments, pow exponents, the cancellation in
// omega - omega_p, the spreading cosine's tangent, lgamma's magnitude).
// Flagged cells are recomputed on the host by the reference's own chain; the
// rest are certified equal to it.
//
// The PCG32 stream of a row is entered at the cell's draw with the LCG
// jump-ahead (O'Neill, "PCG: A Family of Simple Fast Space-Efficient
// Statistically Good Algorithms for Random Number Generation", 2014, the
// square-and-multiply advance of a 64-bit LCG).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <stdint.h>

namespace oceanamd {
namespace cell {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 6.28318530717958647692;
constexpr uint64_t kPcgMult = 6364136223846793005ULL;

// The libm of the build: the device's on the device, the host's on the host.
struct NativeLibm {
    __host__ __device__ static double log(double x) { return ::log(x); }
    __host__ __device__ static double exp(double x) { return ::exp(x); }
    __host__ __device__ static double pow(double x, double y) { return ::pow(x, y); }
    __host__ __device__ static double tanh(double x) { return ::tanh(x); }
    __host__ __device__ static double atan2(double y, double x) { return ::atan2(y, x); }
    __host__ __device__ static double cos(double x) { return ::cos(x); }
    __host__ __device__ static double sin(double x) { return ::sin(x); }
    __host__ __device__ static double lgamma(double x) { return ::lgamma(x); }
};

struct Pcg {
    uint64_t state, inc;
};

__host__ __device__ inline uint32_t next_u32(Pcg& r)
{
    const uint64_t old = r.state;
    r.state = old * kPcgMult + r.inc;
    const uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = uint32_t(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

__host__ __device__ inline Pcg seed_pcg(uint64_t seed, uint64_t stream)
{
    Pcg r{0u, (stream << 1u) | 1u};
    next_u32(r);
    r.state += seed;
    next_u32(r);
    return r;
}

// state after `delta` more steps
__host__ __device__ inline void advance(Pcg& r, uint64_t delta)
{
    uint64_t acc_mult = 1u, acc_plus = 0u, cur_mult = kPcgMult, cur_plus = r.inc;
    while (delta > 0u) {
        if (delta & 1u) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1u) * cur_plus;
        cur_mult *= cur_mult;
        delta >>= 1u;
    }
    r.state = acc_mult * r.state + acc_plus;
}

__host__ __device__ inline uint64_t mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

__host__ __device__ inline double open01(Pcg& r)
{
    return (double(next_u32(r)) + 0.5) * (1.0 / 4294967296.0);
}

// True when rounding v to float cannot be trusted: v lies within `tol` of a
// rounding midpoint (or is not finite).
__host__ __device__ inline bool near_midpoint(double v, double tol)
{
    if (!__builtin_isfinite(v)) return true;
    const float f = float(v);
    if (!__builtin_isfinite(f)) return true;
    const double lo = (double(f) + double(nextafterf(f, -3.4e38f))) * 0.5;
    const double hi = (double(f) + double(nextafterf(f, 3.4e38f))) * 0.5;
    return (v - lo) <= tol || (hi - v) <= tol;
}

struct SpecConst {  // per-descriptor constants, host-computed (host libm)
    double gravity, depth, omega_p, alpha, gamma, wind, swell, wind_dir, cutoff_l, mu, q_den;
};

__host__ __device__ inline double wrap_angle(double a)
{
    a = fmod(a + kPi, kTwoPi);
    if (a < 0.0) a += kTwoPi;
    return a - kPi;
}

template <class M>
__host__ __device__ inline double dispersion_omega(double g, double k, double depth)
{
    if (!(k > 0.0)) return 0.0;
    if (depth <= 0.0) return sqrt(g * k);
    return sqrt(g * k * M::tanh(k * depth));
}

template <class M>
__host__ __device__ inline double dispersion_domega_dk(double g, double k, double omega, double depth)
{
    if (!(omega > 0.0)) return 0.0;
    if (depth <= 0.0) return g / (2.0 * omega);
    const double kh = k * depth;
    const double th = M::tanh(kh);
    const double sech2 = 1.0 - th * th;
    return g * (th + kh * sech2) / (2.0 * omega);
}

constexpr double kUlp = 2.220446049250313e-16;  // 2^-52
constexpr double kU = 8.0;                        // ulps charged per libm call
constexpr double kSafety = 4.0;

// psi(k) with the branch checks and its error bound: `risky` is set when a
// libm-dependent branch is too close to call; `err` receives psi's relative
// error bound in ulps.
template <class M>
__host__ __device__ inline double wave_density(const SpecConst& c, double kx, double kz, double e_omega,
                                               bool& risky, double& err)
{
    err = 0.0;
    const double k2 = kx * kx + kz * kz;
    if (!(k2 > 0.0)) return 0.0;
    const double k = sqrt(k2);
    const double g = c.gravity;
    const double omega = dispersion_omega<M>(g, k, c.depth);
    // jonswap (reference src/core/spectrum.cpp:98-123)
    if (!(omega > 0.0) || !(c.omega_p > 0.0)) return 0.0;
    if (fabs(omega - c.omega_p) <= kSafety * e_omega * kUlp * omega) risky = true;
    const double wp_over_w = c.omega_p / omega;
    const double w2 = omega * omega;
    const double w4 = w2 * w2;
    const double x1 = -1.25 * M::pow(wp_over_w, 4.0);
    const double pm = (c.alpha * g * g / (w4 * omega)) * M::exp(x1);
    const double sigma = (omega <= c.omega_p) ? 0.07 : 0.09;
    const double dw = omega - c.omega_p;
    const double xr = -(dw * dw) / (2.0 * sigma * sigma * c.omega_p * c.omega_p);
    const double r = M::exp(xr);
    const double s_omega = pm * M::pow(c.gamma, r);
    if (!(s_omega > 0.0)) return 0.0;
    const double e_x1 = kU + 4.0 * e_omega + 3.0;
    const double e_xr = (dw != 0.0) ? 2.0 * e_omega * omega / fabs(dw) + 4.0 : 4.0;
    const double e_r = kU + fabs(xr) * e_xr;
    const double e_jonswap = 5.0 * e_omega + 5.0 + kU + fabs(x1) * e_x1 +
                             2.0 * kU + fabs(r * M::log(c.gamma > 0.0 ? c.gamma : 1.0)) * e_r;
    const double domega_dk = dispersion_domega_dk<M>(g, k, omega, c.depth);
    // Finite depth: f = th + kh (1 - th^2), th = tanh(kh). A change d in th
    // moves f by (1 - 2 kh th) d, and th / f <= 1, so tanh's kU ulps reach f
    // amplified by at most 1 + 2 kh (about 2 kh where th has saturated to 1
    // and 1 - th^2 cancels); the rounding of th^2 then differs by half an
    // ulp of 1, scaled by kh / f <= 1 + kh (th >= kh / (1 + kh)).
    const double kh = k * c.depth;
    const double e_dd = (c.depth > 0.0 ? kU * (2.0 + 2.0 * kh) + (1.0 + kh) : 0.0) + e_omega + 6.0;
    const double theta = M::atan2(kz, kx);
    // directional_spread (reference src/core/spectrum.cpp:126-192)
    const double ratio = omega / c.omega_p;
    const double e_ratio = e_omega + 1.0;
    if (fabs(ratio - 1.0) <= kSafety * e_ratio * kUlp) risky = true;
    double s, e_s;
    if (ratio < 1.0) {
        s = 6.97 * M::pow(ratio, 4.06);
        e_s = kU + 4.06 * e_ratio + 2.0;
    } else {
        s = 9.77 * M::pow(ratio, c.mu);
        e_s = kU + fabs(c.mu) * e_ratio + 2.0;
    }
    if (c.swell > 0.0) {
        s += 16.0 * M::tanh(c.omega_p / omega) * c.swell * c.swell;
        e_s = fmax(e_s, kU + e_ratio + 4.0) + 1.0;
    }
    if (s < 0.0) s = 0.0;
    const double dtheta = wrap_angle(theta - c.wind_dir);
    const double cs = M::cos(0.5 * dtheta);
    // absolute error of 0.5 * dtheta, in ulps of 1
    const double e_half = 0.5 * (kU * fabs(theta) + fabs(c.wind_dir) + 4.0 * kPi);
    const double sn = M::sin(0.5 * dtheta);
    if (fabs(cs) <= kSafety * (fabs(sn) * e_half + kU) * kUlp) risky = true;
    if (cs <= 0.0) return 0.0;
    const double e_cs = kU + fabs(sn / cs) * e_half;  // relative, ulps
    const double lg1 = M::lgamma(s + 1.0), lg2 = M::lgamma(s + 0.5);
    const double q = M::exp(lg1 - lg2) / c.q_den;
    const double e_q = kU + kU * (fabs(lg1) + fabs(lg2) + 2.0) + (2.0 * M::log(s + 2.0) + 4.0) * s * e_s + 2.0;
    const double lc = M::log(cs);
    const double spread = q * M::pow(cs, 2.0 * s);
    const double e_pc = kU + fabs(2.0 * s * lc) * e_s + 2.0 * s * e_cs;
    double psi = s_omega * domega_dk * spread / k;
    const double xc = -k2 * c.cutoff_l * c.cutoff_l;
    psi *= M::exp(xc);
    err = e_jonswap + e_dd + e_q + e_pc + kU + 3.0 * fabs(xc) + 8.0;
    return psi;
}

__host__ __device__ inline float wave_component(uint32_t i, uint32_t n, float patch_length)
{
    constexpr float kTwoPiF = 6.28318530717958647692f;
    const int si = (i < n / 2) ? int(i) : int(i) - int(n);
    return kTwoPiF * float(si) / patch_length;
}

// Cell (x, y) of the tables; returns true when the cell is flagged (its
// floats are not certified equal to the host chain's).
template <class M>
__host__ __device__ inline bool spectrum_cell(const SpecConst& c, uint32_t n, float patch_length, double dk,
                                              uint64_t seed, uint32_t x, uint32_t y, float& h0_re,
                                              float& h0_im, float& omega_out)
{
    bool risky = false;
    const float kz_row = wave_component(y, n, patch_length);
    const float kx_val = wave_component(x, n, patch_length);
    const double k2 = double(kx_val) * kx_val + double(kz_row) * kz_row;
    const double k = sqrt(k2);
    const double om = dispersion_omega<M>(c.gravity, k, c.depth);
    // omega's relative error bound (ulps): exact in deep water (sqrt), else tanh
    const double e_omega = c.depth > 0.0 ? 0.5 * kU + 3.0 : 1.0;
    if (c.depth > 0.0 && near_midpoint(om, kSafety * e_omega * kUlp * om)) risky = true;
    omega_out = float(om);
    double e_psi;
    const double psi = wave_density<M>(c, kx_val, kz_row, e_omega, risky, e_psi);
    const double amp = 0.5 * sqrt(psi) * dk;
    const double e_amp = 0.5 * e_psi + 3.0;
    // the row's stream at this cell's two draws (reference src/core/rng.hpp:94-97)
    Pcg r = seed_pcg(mix64(seed ^ (0x9E3779B97F4A7C15ULL * uint32_t(y + 1u))), 0);
    advance(r, 2ull * x);
    const double u1 = open01(r);
    const double u2 = open01(r);
    const double radius = sqrt(-2.0 * M::log(u1));
    const double theta = 6.283185307179586476925286766559 * u2;
    const double cth = M::cos(theta), sth = M::sin(theta);
    const double ad = radius * cth, bd = radius * sth;
    // absolute error bounds: radius (log, sqrt) relative, theta's ulp through cos / sin
    const double e_rad = 0.5 * kU + 2.0;
    const double ea = radius * (fabs(cth) * (e_rad + kU + 1.0) + fabs(sth) * 2.0 * theta) * kUlp;
    const double eb = radius * (fabs(sth) * (e_rad + kU + 1.0) + fabs(cth) * 2.0 * theta) * kUlp;
    if (near_midpoint(ad, kSafety * ea) || near_midpoint(bd, kSafety * eb)) risky = true;
    const float ga = float(ad), gb = float(bd);
    const double hr = double(ga) * amp, hi = double(gb) * amp;
    const double e_h = kSafety * (e_amp + 1.0) * kUlp;
    if (near_midpoint(hr, e_h * fabs(hr)) || near_midpoint(hi, e_h * fabs(hi))) risky = true;
    h0_re = float(hr);
    h0_im = float(hi);
    return risky;
}

// Twiddle j of an n-point transform, (cos, sin)(2 pi j / n); returns true
// when flagged.
template <class M>
__host__ __device__ inline bool twiddle(uint32_t n, uint32_t j, float& cv_out, float& sv_out)
{
    const double angle = 6.283185307179586476925286766559 * double(j) / double(n);
    const double cv = M::cos(angle), sv = M::sin(angle);
    // cos / sin of an angle with a few ulps of rounding: absolute error bounds
    const double e = (kU + 2.0 * angle + 2.0) * kUlp;
    cv_out = float(cv);
    sv_out = float(sv);
    return near_midpoint(cv, kSafety * e) || near_midpoint(sv, kSafety * e);
}

}  // namespace cell
}  // namespace oceanamd

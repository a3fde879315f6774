// The time-independent tables on the device: h0 (re/im), omega and the FFT
// twiddles, one lane per cell, restating the host build (spectrum_host.cpp,
// reference src/core/spectrum.cpp:14-339, src/core/rng.hpp:27-97,
// src/core/rng.cpp:7-17, src/core/fft.cpp:24-33) operation for operation
// with the device libm.
//
// The reference's tables come from the host libm (glibc); the device libm
// can differ from it in the last bits of a double. Those differences reach a
// stored float only where a double lies next to a float rounding midpoint, or
// where a libm result decides a branch. Each lane therefore checks every
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
#include "hip_launch.h"

#include <hip/hip_runtime.h>

#include <stdint.h>

namespace oceanamd {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kTwoPi = 6.28318530717958647692;
constexpr uint64_t kPcgMult = 6364136223846793005ULL;

struct Pcg {
    uint64_t state, inc;
};

__device__ __forceinline__ uint32_t next_u32(Pcg& r)
{
    const uint64_t old = r.state;
    r.state = old * kPcgMult + r.inc;
    const uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = uint32_t(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

__device__ __forceinline__ Pcg seed_pcg(uint64_t seed, uint64_t stream)
{
    Pcg r{0u, (stream << 1u) | 1u};
    next_u32(r);
    r.state += seed;
    next_u32(r);
    return r;
}

// state after `delta` more steps
__device__ __forceinline__ void advance(Pcg& r, uint64_t delta)
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

__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

__device__ __forceinline__ double open01(Pcg& r)
{
    return (double(next_u32(r)) + 0.5) * (1.0 / 4294967296.0);
}

// True when rounding v to float cannot be trusted: v lies within `tol` of a
// rounding midpoint (or is not finite).
__device__ __forceinline__ bool near_midpoint(double v, double tol)
{
    if (!isfinite(v)) return true;
    const float f = float(v);
    if (!isfinite(f)) return true;
    const double lo = (double(f) + double(nextafterf(f, -3.4e38f))) * 0.5;
    const double hi = (double(f) + double(nextafterf(f, 3.4e38f))) * 0.5;
    return (v - lo) <= tol || (hi - v) <= tol;
}

struct SpecConst {  // per-descriptor constants, host-computed (host libm)
    double gravity, depth, omega_p, alpha, gamma, wind, swell, wind_dir, cutoff_l, mu, q_den;
};

__device__ __forceinline__ double wrap_angle(double a)
{
    a = fmod(a + kPi, kTwoPi);
    if (a < 0.0) a += kTwoPi;
    return a - kPi;
}

__device__ __forceinline__ double dispersion_omega(double g, double k, double depth)
{
    if (!(k > 0.0)) return 0.0;
    if (depth <= 0.0) return sqrt(g * k);
    return sqrt(g * k * tanh(k * depth));
}

__device__ __forceinline__ double dispersion_domega_dk(double g, double k, double omega, double depth)
{
    if (!(omega > 0.0)) return 0.0;
    if (depth <= 0.0) return g / (2.0 * omega);
    const double kh = k * depth;
    const double th = tanh(kh);
    const double sech2 = 1.0 - th * th;
    return g * (th + kh * sech2) / (2.0 * omega);
}

constexpr double kUlp = 2.220446049250313e-16;  // 2^-52
constexpr double kU = 8.0;                        // ulps charged per libm call
constexpr double kSafety = 4.0;

// psi(k) with the branch checks and its error bound: `risky` is set when a
// libm-dependent branch is too close to call; `err` receives psi's relative
// error bound in ulps.
__device__ double wave_density(const SpecConst& c, double kx, double kz, double e_omega, bool& risky,
                               double& err)
{
    err = 0.0;
    const double k2 = kx * kx + kz * kz;
    if (!(k2 > 0.0)) return 0.0;
    const double k = sqrt(k2);
    const double g = c.gravity;
    const double omega = dispersion_omega(g, k, c.depth);
    // jonswap (reference src/core/spectrum.cpp:98-123)
    if (!(omega > 0.0) || !(c.omega_p > 0.0)) return 0.0;
    if (fabs(omega - c.omega_p) <= kSafety * e_omega * kUlp * omega) risky = true;
    const double wp_over_w = c.omega_p / omega;
    const double w2 = omega * omega;
    const double w4 = w2 * w2;
    const double x1 = -1.25 * pow(wp_over_w, 4.0);
    const double pm = (c.alpha * g * g / (w4 * omega)) * exp(x1);
    const double sigma = (omega <= c.omega_p) ? 0.07 : 0.09;
    const double dw = omega - c.omega_p;
    const double xr = -(dw * dw) / (2.0 * sigma * sigma * c.omega_p * c.omega_p);
    const double r = exp(xr);
    const double s_omega = pm * pow(c.gamma, r);
    if (!(s_omega > 0.0)) return 0.0;
    const double e_x1 = kU + 4.0 * e_omega + 3.0;
    const double e_xr = (dw != 0.0) ? 2.0 * e_omega * omega / fabs(dw) + 4.0 : 4.0;
    const double e_r = kU + fabs(xr) * e_xr;
    const double e_jonswap = 5.0 * e_omega + 5.0 + kU + fabs(x1) * e_x1 +
                             2.0 * kU + fabs(r * log(c.gamma > 0.0 ? c.gamma : 1.0)) * e_r;
    const double domega_dk = dispersion_domega_dk(g, k, omega, c.depth);
    const double e_dd = (c.depth > 0.0 ? 2.0 * kU : 0.0) + e_omega + 4.0;
    const double theta = atan2(kz, kx);
    // directional_spread (reference src/core/spectrum.cpp:126-192)
    const double ratio = omega / c.omega_p;
    const double e_ratio = e_omega + 1.0;
    if (fabs(ratio - 1.0) <= kSafety * e_ratio * kUlp) risky = true;
    double s, e_s;
    if (ratio < 1.0) {
        s = 6.97 * pow(ratio, 4.06);
        e_s = kU + 4.06 * e_ratio + 2.0;
    } else {
        s = 9.77 * pow(ratio, c.mu);
        e_s = kU + fabs(c.mu) * e_ratio + 2.0;
    }
    if (c.swell > 0.0) {
        s += 16.0 * tanh(c.omega_p / omega) * c.swell * c.swell;
        e_s = fmax(e_s, kU + e_ratio + 4.0) + 1.0;
    }
    if (s < 0.0) s = 0.0;
    const double dtheta = wrap_angle(theta - c.wind_dir);
    const double cs = cos(0.5 * dtheta);
    // absolute error of 0.5 * dtheta, in ulps of 1
    const double e_half = 0.5 * (kU * fabs(theta) + fabs(c.wind_dir) + 4.0 * kPi);
    const double sn = sin(0.5 * dtheta);
    if (fabs(cs) <= kSafety * (fabs(sn) * e_half + kU) * kUlp) risky = true;
    if (cs <= 0.0) return 0.0;
    const double e_cs = kU + fabs(sn / cs) * e_half;  // relative, ulps
    const double lg1 = lgamma(s + 1.0), lg2 = lgamma(s + 0.5);
    const double q = exp(lg1 - lg2) / c.q_den;
    const double e_q = kU + kU * (fabs(lg1) + fabs(lg2) + 2.0) + (2.0 * log(s + 2.0) + 4.0) * s * e_s + 2.0;
    const double lc = log(cs);
    const double spread = q * pow(cs, 2.0 * s);
    const double e_pc = kU + fabs(2.0 * s * lc) * e_s + 2.0 * s * e_cs;
    double psi = s_omega * domega_dk * spread / k;
    const double xc = -k2 * c.cutoff_l * c.cutoff_l;
    psi *= exp(xc);
    err = e_jonswap + e_dd + e_q + e_pc + kU + 3.0 * fabs(xc) + 8.0;
    return psi;
}

__device__ __forceinline__ float wave_component(uint32_t i, uint32_t n, float patch_length)
{
    constexpr float kTwoPiF = 6.28318530717958647692f;
    const int si = (i < n / 2) ? int(i) : int(i) - int(n);
    return kTwoPiF * float(si) / patch_length;
}

__global__ __launch_bounds__(256) void spectrum_kernel(SpecConst c, uint32_t n, float patch_length,
                                                       double dk, uint64_t seed,
                                                       float2* __restrict__ h0,
                                                       float* __restrict__ omega_out,
                                                       uint8_t* __restrict__ flags)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n * n) return;
    const uint32_t y = i / n, x = i % n;
    bool risky = false;
    const float kz_row = wave_component(y, n, patch_length);
    const float kx_val = wave_component(x, n, patch_length);
    const double k2 = double(kx_val) * kx_val + double(kz_row) * kz_row;
    const double k = sqrt(k2);
    const double om = dispersion_omega(c.gravity, k, c.depth);
    // omega's relative error bound (ulps): exact in deep water (sqrt), else tanh
    const double e_omega = c.depth > 0.0 ? 0.5 * kU + 3.0 : 1.0;
    if (c.depth > 0.0 && near_midpoint(om, kSafety * e_omega * kUlp * om)) risky = true;
    omega_out[i] = float(om);
    double e_psi;
    const double psi = wave_density(c, kx_val, kz_row, e_omega, risky, e_psi);
    const double amp = 0.5 * sqrt(psi) * dk;
    const double e_amp = 0.5 * e_psi + 3.0;
    // the row's stream at this cell's two draws (reference src/core/rng.hpp:94-97)
    Pcg r = seed_pcg(mix64(seed ^ (0x9E3779B97F4A7C15ULL * uint32_t(y + 1u))), 0);
    advance(r, 2ull * x);
    const double u1 = open01(r);
    const double u2 = open01(r);
    const double radius = sqrt(-2.0 * log(u1));
    const double theta = 6.283185307179586476925286766559 * u2;
    const double cth = cos(theta), sth = sin(theta);
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
    h0[i] = make_float2(float(hr), float(hi));
    flags[i] = risky ? 1u : 0u;
}

__global__ __launch_bounds__(256) void twiddle_kernel(uint32_t n, float2* __restrict__ tw,
                                                      uint8_t* __restrict__ flags)
{
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j >= n / 2) return;
    const double angle = 6.283185307179586476925286766559 * double(j) / double(n);
    const double cv = cos(angle), sv = sin(angle);
    // cos / sin of an angle with a few ulps of rounding: absolute error bounds
    const double e = (kU + 2.0 * angle + 2.0) * kUlp;
    const bool risky = near_midpoint(cv, kSafety * e) || near_midpoint(sv, kSafety * e);
    tw[j] = make_float2(float(cv), float(sv));
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

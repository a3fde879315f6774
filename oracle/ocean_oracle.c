/* CPU oracle: plain-C restatement of the reference's per-frame wave synthesis.
 * TEST INFRASTRUCTURE ONLY - see ocean_oracle.h. Every function cites the
 * reference file:line whose operation order it follows. */
#include "ocean_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

/* ======================================================================
 * PCG32 + SplitMix64 + Box-Muller (reference src/core/rng.hpp:27-97,
 * src/core/rng.cpp:7-17)
 * ====================================================================== */

typedef struct { uint64_t state, inc; } pcg32;

static const uint64_t kPcgMult = 6364136223846793005ULL;

/* rng.hpp:35-47 */
static uint32_t pcg_next(pcg32* r)
{
    const uint64_t old = r->state;
    r->state = old * kPcgMult + r->inc;
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

/* rng.hpp:53-62 */
static pcg32 pcg_seed(uint64_t seed, uint64_t stream)
{
    pcg32 r;
    r.state = 0u;
    r.inc = (stream << 1u) | 1u;
    pcg_next(&r);
    r.state += seed;
    pcg_next(&r);
    return r;
}

/* rng.hpp:68-71 */
static double open01(pcg32* r)
{
    return ((double)pcg_next(r) + 0.5) * (1.0 / 4294967296.0);
}

/* rng.hpp:83-89 */
static uint64_t mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

/* rng.hpp:94-97 */
static pcg32 row_rng(uint64_t seed, uint32_t row)
{
    return pcg_seed(mix64(seed ^ (0x9E3779B97F4A7C15ULL * (uint32_t)(row + 1u))), 0);
}

/* rng.cpp:7-17 */
static void gaussian_pair(pcg32* r, float* a, float* b)
{
    const double u1 = open01(r);
    const double u2 = open01(r);
    const double radius = sqrt(-2.0 * log(u1));
    const double theta = 6.283185307179586476925286766559 * u2;
    *a = (float)(radius * cos(theta));
    *b = (float)(radius * sin(theta));
}

EXPORT void orc_pcg32_draws(uint64_t seed, uint64_t stream, uint32_t count, uint32_t* out)
{
    pcg32 r = pcg_seed(seed, stream);
    for (uint32_t i = 0; i < count; ++i) out[i] = pcg_next(&r);
}

EXPORT void orc_row_gaussians(uint64_t seed, uint32_t row, uint32_t pairs, float* out)
{
    pcg32 r = row_rng(seed, row);
    for (uint32_t i = 0; i < pairs; ++i) gaussian_pair(&r, &out[2 * i], &out[2 * i + 1]);
}

/* ======================================================================
 * Own trig (reference src/core/trig.hpp:33-129)
 * ====================================================================== */

static const float kTwoOverPi = 0.636619772367581343076f;
static const float kPiOver2A = 1.5707963705062866f;
static const float kPiOver2B = -4.3711388286737929e-8f;
static const float kSin0 = -1.9515295891e-4f;
static const float kSin1 = 8.3321608736e-3f;
static const float kSin2 = -1.6666654611e-1f;
static const float kCos0 = 2.443315711809948e-5f;
static const float kCos1 = -1.388731625493765e-3f;
static const float kCos2 = 4.166664568298827e-2f;
static const double kTwoPiD = 6.28318530717958647692;
static const double kInvTwoPiD = 0.159154943091895335769;

/* trig.hpp:76-80 */
static float reduced_phase(float omega, double time_over_two_pi)
{
    const double d = (double)omega * time_over_two_pi;
    return (float)((d - floor(d)) * kTwoPiD);
}

/* trig.hpp:87-129 */
static void sincos_f32(float x, float* out_s, float* out_c)
{
    const float qf = floorf(x * kTwoOverPi + 0.5f);
    const int32_t q = (int32_t)qf;
    float y = x - qf * kPiOver2A;
    y = y - qf * kPiOver2B;
    const float z = y * y;
    const float sin_y = y + ((kSin0 * z + kSin1) * z + kSin2) * z * y;
    const float cos_y = 1.0f - 0.5f * z + ((kCos0 * z + kCos1) * z + kCos2) * z * z;
    const int32_t quadrant = q & 3;
    const int swap = (quadrant & 1) != 0;
    const float bs = swap ? cos_y : sin_y;
    const float bc = swap ? sin_y : cos_y;
    const int neg_s = (quadrant & 2) != 0;
    const int neg_c = ((quadrant + 1) & 2) != 0;
    *out_s = neg_s ? -bs : bs;
    *out_c = neg_c ? -bc : bc;
}

EXPORT void orc_sincos(const float* x, uint32_t count, float* s, float* c)
{
    for (uint32_t i = 0; i < count; ++i) sincos_f32(x[i], &s[i], &c[i]);
}

EXPORT void orc_reduced_phase(const float* omega, uint32_t count, double time, float* out)
{
    const double t2 = time * kInvTwoPiD;
    for (uint32_t i = 0; i < count; ++i) out[i] = reduced_phase(omega[i], t2);
}

/* ======================================================================
 * JONSWAP directional spectrum (reference src/core/spectrum.cpp:14-339,
 * src/core/spectrum.hpp:85-92)
 * ====================================================================== */

static const double kPi = 3.14159265358979323846;
static const double kTwoPi = 6.28318530717958647692;

/* spectrum.cpp:14-19 */
static double wrap_angle(double a)
{
    a = fmod(a + kPi, kTwoPi);
    if (a < 0.0) a += kTwoPi;
    return a - kPi;
}

/* spectrum.cpp:27-32 */
static double dispersion_omega(double gravity, double k, double depth)
{
    if (!(k > 0.0)) return 0.0;
    if (depth <= 0.0) return sqrt(gravity * k);
    return sqrt(gravity * k * tanh(k * depth));
}

/* spectrum.cpp:34-44 */
static double dispersion_domega_dk(double gravity, double k, double omega, double depth)
{
    if (!(omega > 0.0)) return 0.0;
    if (depth <= 0.0) return gravity / (2.0 * omega);
    const double kh = k * depth;
    const double th = tanh(kh);
    const double sech2 = 1.0 - th * th;
    return gravity * (th + kh * sech2) / (2.0 * omega);
}

/* spectrum.cpp:61-74 */
static double jonswap_peak_omega(const ocean_desc* d)
{
    const double g = d->spectrum.gravity;
    const double u = (double)d->spectrum.wind_speed;
    const double f = (double)d->spectrum.fetch;
    if (u <= 0.0 || f <= 0.0) return 0.0;
    return 22.0 * cbrt(g * g / (u * f));
}

/* spectrum.cpp:76-85 */
static double jonswap_alpha(const ocean_desc* d)
{
    const double g = d->spectrum.gravity;
    const double u = (double)d->spectrum.wind_speed;
    const double f = (double)d->spectrum.fetch;
    if (u <= 0.0 || f <= 0.0) return 0.0;
    const double x_tilde = g * f / (u * u);
    return 0.076 * pow(x_tilde, -0.22);
}

/* spectrum.cpp:87-123 */
static double jonswap(const ocean_desc* d, double omega)
{
    if (!(omega > 0.0)) return 0.0;
    const double g = d->spectrum.gravity;
    const double omega_p = jonswap_peak_omega(d);
    if (!(omega_p > 0.0)) return 0.0;
    const double alpha = jonswap_alpha(d);
    const double gamma = (double)d->spectrum.peak_enhancement;
    const double wp_over_w = omega_p / omega;
    const double w2 = omega * omega;
    const double w4 = w2 * w2;
    const double pm = (alpha * g * g / (w4 * omega)) * exp(-1.25 * pow(wp_over_w, 4.0));
    const double sigma = (omega <= omega_p) ? 0.07 : 0.09;
    const double dw = omega - omega_p;
    const double r = exp(-(dw * dw) / (2.0 * sigma * sigma * omega_p * omega_p));
    return pm * pow(gamma, r);
}

/* spectrum.cpp:129-192 */
static double directional_spread(const ocean_desc* d, double omega, double theta)
{
    const double omega_p = jonswap_peak_omega(d);
    if (!(omega > 0.0) || !(omega_p > 0.0)) return 0.0;
    const double g = d->spectrum.gravity;
    const double u = (double)d->spectrum.wind_speed;
    const double ratio = omega / omega_p;
    double s;
    if (ratio < 1.0) {
        s = 6.97 * pow(ratio, 4.06);
    } else {
        const double mu = -2.33 - 1.45 * ((u * omega_p / g) - 1.17);
        s = 9.77 * pow(ratio, mu);
    }
    const double swell = (double)d->spectrum.swell;
    if (swell > 0.0) s += 16.0 * tanh(omega_p / omega) * swell * swell;
    if (s < 0.0) s = 0.0;
    const double dtheta = wrap_angle(theta - (double)d->spectrum.wind_direction);
    const double c = cos(0.5 * dtheta);
    if (c <= 0.0) return 0.0;
    const double q = exp(lgamma(s + 1.0) - lgamma(s + 0.5)) / (2.0 * sqrt(kPi));
    return q * pow(c, 2.0 * s);
}

/* spectrum.cpp:198-244 */
static double wave_density(const ocean_desc* d, double kx, double kz)
{
    const double k2 = kx * kx + kz * kz;
    if (!(k2 > 0.0)) return 0.0;
    const double k = sqrt(k2);
    const double g = d->spectrum.gravity;
    const double depth = (double)d->water_depth;
    const double omega = dispersion_omega(g, k, depth);
    const double s_omega = jonswap(d, omega);
    if (!(s_omega > 0.0)) return 0.0;
    const double domega_dk = dispersion_domega_dk(g, k, omega, depth);
    const double theta = atan2(kz, kx);
    const double spread = directional_spread(d, omega, theta);
    double psi = s_omega * domega_dk * spread / k;
    const double l = (double)d->spectrum.small_wave_cutoff / kTwoPi;
    psi *= exp(-k2 * l * l);
    return psi;
}

EXPORT double orc_wave_density(const ocean_desc* d, double kx, double kz)
{
    return wave_density(d, kx, kz);
}

/* spectrum.hpp:85-92 */
static float wave_component(uint32_t i, uint32_t n, float patch_length)
{
    const float kTwoPiF = 6.28318530717958647692f;
    const int si = (i < n / 2) ? (int)i : (int)i - (int)n;
    return kTwoPiF * (float)si / patch_length;
}

typedef struct {
    uint32_t n;
    float *h0_re, *h0_im, *h0c_re, *h0c_im, *omega, *kx, *kz, *k_inv;
} tables_t;

/* spectrum.cpp:264-339 */
static void build_spectrum(const ocean_desc* desc, tables_t* t)
{
    const uint32_t n = desc->size;
    const float L = desc->patch_length;
    const double dk = kTwoPi / (double)L;
    for (uint32_t y = 0; y < n; ++y) {
        pcg32 rng = row_rng(desc->seed, y);
        const float kz_row = wave_component(y, n, L);
        for (uint32_t x = 0; x < n; ++x) {
            const size_t i = (size_t)y * n + x;
            const float kx_val = wave_component(x, n, L);
            t->kx[i] = kx_val;
            t->kz[i] = kz_row;
            const double k2 = (double)kx_val * kx_val + (double)kz_row * kz_row;
            const double k = sqrt(k2);
            t->omega[i] = (float)dispersion_omega(desc->spectrum.gravity, k,
                                                  (double)desc->water_depth);
            t->k_inv[i] = (k > 0.0) ? (float)(1.0 / k) : 0.0f;
            const double psi = wave_density(desc, kx_val, kz_row);
            const double amp = 0.5 * sqrt(psi) * dk;
            float ga, gb;
            gaussian_pair(&rng, &ga, &gb);
            t->h0_re[i] = (float)(ga * amp);
            t->h0_im[i] = (float)(gb * amp);
        }
    }
    for (uint32_t y = 0; y < n; ++y) {
        const uint32_t my = (n - y) % n;
        for (uint32_t x = 0; x < n; ++x) {
            const uint32_t mx = (n - x) % n;
            const size_t i = (size_t)y * n + x;
            const size_t m = (size_t)my * n + mx;
            t->h0c_re[i] = t->h0_re[m];
            t->h0c_im[i] = -t->h0_im[m];
        }
    }
}

EXPORT void orc_spectrum(const ocean_desc* d, float* h0re, float* h0im, float* h0cre,
                         float* h0cim, float* omega, float* kx, float* kz, float* kinv)
{
    tables_t t = {d->size, h0re, h0im, h0cre, h0cim, omega, kx, kz, kinv};
    build_spectrum(d, &t);
}

/* ======================================================================
 * Stockham radix-2 FFT (reference src/core/fft.cpp:11-179,
 * src/core/fft_kernel_scalar.cpp:5-35)
 * ====================================================================== */

/* fft.cpp:24-33: twiddles in double, narrowed; inverse uses +sin. */
static void twiddles(uint32_t n, float* tw_re, float* tw_im_inv)
{
    const double k2pi = 6.283185307179586476925286766559;
    for (uint32_t j = 0; j < n / 2; ++j) {
        const double angle = k2pi * (double)j / (double)n;
        tw_re[j] = (float)cos(angle);
        tw_im_inv[j] = (float)sin(angle);
    }
}

EXPORT void orc_twiddles(uint32_t n, float* tw_re, float* tw_im_inv)
{
    twiddles(n, tw_re, tw_im_inv);
}

/* fft_kernel_scalar.cpp:5-35 */
static void stage(const float* xr, const float* xi, float* yr, float* yi, const float* twr,
                  const float* twi, uint32_t len, uint32_t s, uint32_t tw_step)
{
    const uint32_t m = len >> 1;
    for (uint32_t p = 0; p < m; ++p) {
        const float wr = twr[(size_t)p * tw_step];
        const float wi = twi[(size_t)p * tw_step];
        const uint32_t src0 = s * p, src1 = s * (p + m);
        const uint32_t dst0 = s * (2u * p), dst1 = s * (2u * p + 1u);
        for (uint32_t q = 0; q < s; ++q) {
            const float ar = xr[src0 + q], ai = xi[src0 + q];
            const float br = xr[src1 + q], bi = xi[src1 + q];
            yr[dst0 + q] = ar + br;
            yi[dst0 + q] = ai + bi;
            const float dr = ar - br;
            const float di = ai - bi;
            yr[dst1 + q] = dr * wr - di * wi;
            yi[dst1 + q] = dr * wi + di * wr;
        }
    }
}

/* fft.cpp:36-82: gather, log2 N stages ping-pong, scatter. */
static void transform(uint32_t n, const float* twr, const float* twi, float* re, float* im,
                      ptrdiff_t stride, float* scratch)
{
    float *ar = scratch, *ai = scratch + n, *br = scratch + 2 * (size_t)n,
          *bi = scratch + 3 * (size_t)n;
    for (uint32_t j = 0; j < n; ++j) {
        ar[j] = re[(ptrdiff_t)j * stride];
        ai[j] = im[(ptrdiff_t)j * stride];
    }
    float *xr = ar, *xi = ai, *yr = br, *yi = bi;
    for (uint32_t len = n, s = 1; len > 1; len >>= 1, s <<= 1) {
        stage(xr, xi, yr, yi, twr, twi, len, s, s);
        float* t;
        t = xr; xr = yr; yr = t;
        t = xi; xi = yi; yi = t;
    }
    for (uint32_t j = 0; j < n; ++j) {
        re[(ptrdiff_t)j * stride] = xr[j];
        im[(ptrdiff_t)j * stride] = xi[j];
    }
}

/* fft.cpp:170-179 with ocean.cpp:193-233: all rows, then all columns. A
 * column transformed alone runs the same butterfly graph as the reference's
 * 8-wide batched column pass (fft.cpp:95-150: the batch only scales memory
 * addresses, tw_step stays s). */
static void transform_2d(uint32_t n, const float* twr, const float* twi, float* re, float* im,
                         float* scratch)
{
    for (uint32_t r = 0; r < n; ++r)
        transform(n, twr, twi, re + (size_t)r * n, im + (size_t)r * n, 1, scratch);
    for (uint32_t c = 0; c < n; ++c)
        transform(n, twr, twi, re + c, im + c, (ptrdiff_t)n, scratch);
}

EXPORT void orc_fft_inverse(uint32_t n, float* re, float* im)
{
    float* twr = malloc(sizeof(float) * (n / 2 + 1));
    float* twi = malloc(sizeof(float) * (n / 2 + 1));
    float* scratch = malloc(sizeof(float) * 4 * (size_t)n);
    twiddles(n, twr, twi);
    transform(n, twr, twi, re, im, 1, scratch);
    free(twr); free(twi); free(scratch);
}

EXPORT void orc_fft2d_inverse(uint32_t n, float* re, float* im)
{
    float* twr = malloc(sizeof(float) * (n / 2 + 1));
    float* twi = malloc(sizeof(float) * (n / 2 + 1));
    float* scratch = malloc(sizeof(float) * 4 * (size_t)n);
    twiddles(n, twr, twi);
    transform_2d(n, twr, twi, re, im, scratch);
    free(twr); free(twi); free(scratch);
}

/* ======================================================================
 * Evolve + finalize (reference src/core/evolve.cpp:28-327)
 * ====================================================================== */

/* evolve.cpp:28-121; planes = 2*f (re), 2*f+1 (im), each n*n. */
static void evolve(const tables_t* t, double time, float* planes, size_t cells)
{
    const uint32_t n = t->n;
    const double t2 = time * kInvTwoPiD;
    float *p0r = planes, *p0i = planes + cells, *p1r = planes + 2 * cells,
          *p1i = planes + 3 * cells, *p2r = planes + 4 * cells, *p2i = planes + 5 * cells,
          *p3r = planes + 6 * cells, *p3i = planes + 7 * cells;
    for (uint32_t y = 0; y < n; ++y) {
        for (uint32_t x = 0; x < n; ++x) {
            const size_t i = (size_t)y * n + x;
            const float phase = reduced_phase(t->omega[i], t2);
            float s, c;
            sincos_f32(phase, &s, &c);
            const float a_re = t->h0_re[i], a_im = t->h0_im[i];
            const float b_re = t->h0c_re[i], b_im = t->h0c_im[i];
            const float fwd_re = a_re * c - a_im * s;
            const float fwd_im = a_re * s + a_im * c;
            const float bwd_re = b_re * c + b_im * s;
            const float bwd_im = b_im * c - b_re * s;
            const float hr = fwd_re + bwd_re;
            const float hi = fwd_im + bwd_im;
            const float kx = t->kx[i], kz = t->kz[i], ki = t->k_inv[i];
            const float sx = kx * ki, sz = kz * ki;
            const float h_re = hr, h_im = hi;
            const float dx_re = sx * hi, dx_im = -sx * hr;
            const float dz_re = sz * hi, dz_im = -sz * hr;
            const float hx_re = -kx * hi, hx_im = kx * hr;
            const float hz_re = -kz * hi, hz_im = kz * hr;
            const float jxx = kx * sx, jzz = kz * sz, jxz = kx * sz;
            const float jxx_re = jxx * hr, jxx_im = jxx * hi;
            const float jzz_re = jzz * hr, jzz_im = jzz * hi;
            const float jxz_re = jxz * hr, jxz_im = jxz * hi;
            p0r[i] = h_re - dx_im;   p0i[i] = h_im + dx_re;
            p1r[i] = dz_re - hx_im;  p1i[i] = dz_im + hx_re;
            p2r[i] = hz_re - jxx_im; p2i[i] = hz_im + jxx_re;
            p3r[i] = jzz_re - jxz_im; p3i[i] = jzz_im + jxz_re;
        }
    }
}

/* evolve.cpp:123-184 */
static void evolve_velocity(const tables_t* t, double time, float* planes, size_t cells)
{
    const uint32_t n = t->n;
    const double t2 = time * kInvTwoPiD;
    float *p4r = planes + 8 * cells, *p4i = planes + 9 * cells, *p5r = planes + 10 * cells,
          *p5i = planes + 11 * cells;
    for (uint32_t y = 0; y < n; ++y) {
        for (uint32_t x = 0; x < n; ++x) {
            const size_t i = (size_t)y * n + x;
            const float phase = reduced_phase(t->omega[i], t2);
            float s, c;
            sincos_f32(phase, &s, &c);
            const float a_re = t->h0_re[i], a_im = t->h0_im[i];
            const float b_re = t->h0c_re[i], b_im = t->h0c_im[i];
            const float fwd_re = a_re * c - a_im * s;
            const float fwd_im = a_re * s + a_im * c;
            const float bwd_re = b_re * c + b_im * s;
            const float bwd_im = b_im * c - b_re * s;
            const float dr = fwd_re - bwd_re;
            const float di = fwd_im - bwd_im;
            const float w = t->omega[i], kx = t->kx[i], kz = t->kz[i], ki = t->k_inv[i];
            const float sx = kx * ki, sz = kz * ki;
            const float wsx = w * sx, wsz = w * sz;
            const float ux_re = -wsx * dr, ux_im = -wsx * di;
            const float uz_re = -wsz * dr, uz_im = -wsz * di;
            const float uy_re = -w * di, uy_im = w * dr;
            p4r[i] = ux_re - uy_im; p4i[i] = ux_im + uy_re;
            p5r[i] = uz_re;         p5i[i] = uz_im;
        }
    }
}

/* evolve.cpp:238-327 */
static void finalize(const float* planes, size_t cells, uint32_t n, float lambda,
                     float foam_threshold, float* disp, float* normal)
{
    const float *f0r = planes, *f0i = planes + cells, *f1r = planes + 2 * cells,
                *f1i = planes + 3 * cells, *f2r = planes + 4 * cells, *f2i = planes + 5 * cells,
                *f3r = planes + 6 * cells, *f3i = planes + 7 * cells;
    const float inv_thr = (foam_threshold > 1e-6f) ? (1.0f / foam_threshold) : 0.0f;
    for (size_t i = 0; i < (size_t)n * n; ++i) {
        const float h = f0r[i], d_x = f0i[i], d_z = f1r[i], hx = f1i[i];
        const float hz = f2r[i], jxx = f2i[i], jzz = f3r[i], jxz = f3i[i];
        const float disp_x = lambda * d_x;
        const float disp_z = lambda * d_z;
        const float axx = 1.0f + lambda * jxx;
        const float azz = 1.0f + lambda * jzz;
        const float axz = lambda * jxz;
        const float jac = axx * azz - axz * axz;
        float nx = -(hx * azz - axz * hz);
        float ny = jac;
        float nz = -(axx * hz - hx * axz);
        const float len2 = nx * nx + ny * ny + nz * nz;
        const float inv_len = (len2 > 0.0f) ? 1.0f / sqrtf(len2) : 0.0f;
        nx *= inv_len; ny *= inv_len; nz *= inv_len;
        float foam = (foam_threshold - jac) * inv_thr;
        if (foam < 0.0f) foam = 0.0f;
        if (foam > 1.0f) foam = 1.0f;
        disp[4 * i + 0] = disp_x; disp[4 * i + 1] = h; disp[4 * i + 2] = disp_z;
        disp[4 * i + 3] = foam;
        normal[4 * i + 0] = nx; normal[4 * i + 1] = ny; normal[4 * i + 2] = nz;
        normal[4 * i + 3] = jac;
    }
}

/* evolve.cpp:186-203 */
static void finalize_velocity(const float* planes, size_t cells, float* vel)
{
    const float *f4r = planes + 8 * cells, *f4i = planes + 9 * cells, *f5r = planes + 10 * cells;
    for (size_t i = 0; i < cells; ++i) {
        vel[4 * i + 0] = f4r[i];
        vel[4 * i + 1] = f4i[i];
        vel[4 * i + 2] = f5r[i];
        vel[4 * i + 3] = 0.0f;
    }
}

/* ======================================================================
 * The oracle ocean: reference src/ocean.cpp:65-289 (allocation at create,
 * run(0) before returning, run = evolve -> rows -> columns -> finalize).
 * ====================================================================== */

struct orc_ocean {
    ocean_desc desc;
    double time;
    size_t cells;
    tables_t t;
    float *planes, *twr, *twi, *scratch;
    float *disp, *normal, *vel;
};

static void ocean_run(orc_ocean* o, double time)
{
    const uint32_t n = o->desc.size;
    const int nf = o->desc.compute_velocity ? 6 : 4;
    evolve(&o->t, time, o->planes, o->cells);
    if (o->desc.compute_velocity) evolve_velocity(&o->t, time, o->planes, o->cells);
    for (int f = 0; f < nf; ++f)
        transform_2d(n, o->twr, o->twi, o->planes + 2 * f * o->cells,
                     o->planes + (2 * f + 1) * o->cells, o->scratch);
    finalize(o->planes, o->cells, n, o->desc.choppiness, o->desc.foam_threshold, o->disp,
             o->normal);
    if (o->desc.compute_velocity) finalize_velocity(o->planes, o->cells, o->vel);
    o->time = time;
}

EXPORT orc_ocean* orc_ocean_create(const ocean_desc* d)
{
    orc_ocean* o = calloc(1, sizeof(*o));
    o->desc = *d;
    const uint32_t n = d->size;
    o->cells = (size_t)n * n;
    float* tab = malloc(sizeof(float) * 8 * o->cells);
    o->t.n = n;
    o->t.h0_re = tab; o->t.h0_im = tab + o->cells; o->t.h0c_re = tab + 2 * o->cells;
    o->t.h0c_im = tab + 3 * o->cells; o->t.omega = tab + 4 * o->cells;
    o->t.kx = tab + 5 * o->cells; o->t.kz = tab + 6 * o->cells; o->t.k_inv = tab + 7 * o->cells;
    build_spectrum(d, &o->t);
    o->planes = calloc(12 * o->cells, sizeof(float));
    o->twr = malloc(sizeof(float) * (n / 2 + 1));
    o->twi = malloc(sizeof(float) * (n / 2 + 1));
    twiddles(n, o->twr, o->twi);
    o->scratch = malloc(sizeof(float) * 4 * (size_t)n);
    o->disp = calloc(4 * o->cells, sizeof(float));
    o->normal = calloc(4 * o->cells, sizeof(float));
    o->vel = d->compute_velocity ? calloc(4 * o->cells, sizeof(float)) : NULL;
    ocean_run(o, 0.0);
    return o;
}

EXPORT void orc_ocean_destroy(orc_ocean* o)
{
    if (!o) return;
    free(o->t.h0_re); free(o->planes); free(o->twr); free(o->twi); free(o->scratch);
    free(o->disp); free(o->normal); free(o->vel);
    free(o);
}

EXPORT void orc_ocean_update(orc_ocean* o, double time) { ocean_run(o, time); }
EXPORT const float* orc_ocean_displacement(const orc_ocean* o) { return o->disp; }
EXPORT const float* orc_ocean_normal(const orc_ocean* o) { return o->normal; }
EXPORT const float* orc_ocean_velocity(const orc_ocean* o) { return o->vel; }
EXPORT const float* orc_ocean_planes(const orc_ocean* o) { return o->planes; }

/* Evolve only (no transform), into caller planes: the reference's FieldSet
 * after task_evolve (ocean.cpp:175-188). */
EXPORT void orc_ocean_evolve(orc_ocean* o, double time, float* planes)
{
    evolve(&o->t, time, planes, o->cells);
    if (o->desc.compute_velocity) evolve_velocity(&o->t, time, planes, o->cells);
}

/* ======================================================================
 * Queries (reference src/ocean.cpp:309-452)
 * ====================================================================== */

typedef struct { size_t i00, i10, i01, i11; float w00, w10, w01, w11; } tap_t;

/* ocean.cpp:322-327 */
static uint32_t wrap_index(int i, int n)
{
    i %= n;
    if (i < 0) i += n;
    return (uint32_t)i;
}

/* ocean.cpp:330-352 */
static tap_t make_tap(float u, float v, uint32_t n, float inv_cell)
{
    const float gu = u * inv_cell;
    const float gv = v * inv_cell;
    const float fu = floorf(gu);
    const float fv = floorf(gv);
    const int ni = (int)n;
    const uint32_t x0 = wrap_index((int)fu, ni);
    const uint32_t z0 = wrap_index((int)fv, ni);
    const uint32_t x1 = (x0 + 1u) % n;
    const uint32_t z1 = (z0 + 1u) % n;
    const float tu = gu - fu;
    const float tv = gv - fv;
    tap_t t;
    t.i00 = ((size_t)z0 * n + x0) * 4;
    t.i10 = ((size_t)z0 * n + x1) * 4;
    t.i01 = ((size_t)z1 * n + x0) * 4;
    t.i11 = ((size_t)z1 * n + x1) * 4;
    t.w00 = (1.0f - tu) * (1.0f - tv);
    t.w10 = tu * (1.0f - tv);
    t.w01 = (1.0f - tu) * tv;
    t.w11 = tu * tv;
    return t;
}

/* ocean.cpp:354-358 */
static float fetch(const float* buf, const tap_t* t, int c)
{
    return buf[t->i00 + c] * t->w00 + buf[t->i10 + c] * t->w10 + buf[t->i01 + c] * t->w01 +
           buf[t->i11 + c] * t->w11;
}

/* ocean.cpp:375-452; kInversionSteps = 4 (ocean.cpp:366). */
static void sample_at(const ocean_desc* d, const float* disp, const float* nrm, const float* vel,
                      float wx, float wz, ocean_surface_ex* s)
{
    const uint32_t n = d->size;
    const float inv_cell = (float)n / d->patch_length;
    float u = wx, v = wz;
    for (int it = 0; it < 4; ++it) {
        const tap_t t = make_tap(u, v, n, inv_cell);
        u = wx - fetch(disp, &t, 0);
        v = wz - fetch(disp, &t, 2);
    }
    const tap_t t = make_tap(u, v, n, inv_cell);
    memset(s, 0, sizeof(*s));
    s->offset_x = fetch(disp, &t, 0);
    s->height = fetch(disp, &t, 1);
    s->offset_z = fetch(disp, &t, 2);
    s->foam = fetch(disp, &t, 3);
    float nx = fetch(nrm, &t, 0), ny = fetch(nrm, &t, 1), nz = fetch(nrm, &t, 2);
    const float len2 = nx * nx + ny * ny + nz * nz;
    if (len2 > 0.0f) {
        const float inv = 1.0f / sqrtf(len2);
        nx *= inv; ny *= inv; nz *= inv;
    } else {
        nx = 0.0f; ny = 1.0f; nz = 0.0f;
    }
    s->normal_x = nx; s->normal_y = ny; s->normal_z = nz;
    if (vel) {
        s->velocity_x = fetch(vel, &t, 0);
        s->velocity_y = fetch(vel, &t, 1);
        s->velocity_z = fetch(vel, &t, 2);
    }
}

EXPORT void orc_ocean_sample(const orc_ocean* o, float wx, float wz, ocean_surface_ex* out)
{
    sample_at(&o->desc, o->disp, o->normal, o->vel, wx, wz, out);
}

/* Batched form over caller-held buffers (any ocean state, e.g. a GPU frame). */
EXPORT void orc_sample_buffers(const ocean_desc* d, const float* disp, const float* nrm,
                               const float* vel, const float* xz, uint32_t count,
                               ocean_surface_ex* out)
{
    for (uint32_t q = 0; q < count; ++q)
        sample_at(d, disp, nrm, vel, xz[2 * q], xz[2 * q + 1], &out[q]);
}

/* Cascade query: reference src/cascade.cpp:82-153, levels in order. */
EXPORT void orc_cascade_sample(orc_ocean* const* levels, uint32_t count, float wx, float wz,
                               ocean_surface_ex* out)
{
    float height = 0.0f, ox = 0.0f, oz = 0.0f, nx = 0.0f, ny = 0.0f, nz = 0.0f;
    float transp = 1.0f, vx = 0.0f, vy = 0.0f, vz = 0.0f;
    for (uint32_t l = 0; l < count; ++l) {
        ocean_surface_ex s;
        orc_ocean_sample(levels[l], wx, wz, &s);
        height += s.height; ox += s.offset_x; oz += s.offset_z;
        nx += s.normal_x; ny += s.normal_y; nz += s.normal_z;
        vx += s.velocity_x; vy += s.velocity_y; vz += s.velocity_z;
        transp *= (1.0f - s.foam);
    }
    memset(out, 0, sizeof(*out));
    out->height = height; out->offset_x = ox; out->offset_z = oz;
    out->foam = 1.0f - transp;
    out->velocity_x = vx; out->velocity_y = vy; out->velocity_z = vz;
    const float len2 = nx * nx + ny * ny + nz * nz;
    if (len2 > 0.0f) {
        const float inv = 1.0f / sqrtf(len2);
        out->normal_x = nx * inv; out->normal_y = ny * inv; out->normal_z = nz * inv;
    } else {
        out->normal_x = 0.0f; out->normal_y = 1.0f; out->normal_z = 0.0f;
    }
}

/* ======================================================================
 * Persistent advected foam (reference src/foam.cpp:22-252,291-305)
 * ====================================================================== */

struct orc_foam {
    const orc_ocean* ocean;
    orc_foam_params p;
    uint32_t n;
    size_t cells;
    float cell;
    float *a, *b, *cur, *next;
    double accum;
    uint32_t last;
};

/* foam.cpp:22-41 */
static float sample_wrap(const float* f, float u, float v, int n)
{
    const float fu = floorf(u);
    const float fv = floorf(v);
    const float tu = u - fu;
    const float tv = v - fv;
    int x0 = (int)fu % n; if (x0 < 0) x0 += n;
    int z0 = (int)fv % n; if (z0 < 0) z0 += n;
    const int x1 = (x0 + 1) % n;
    const int z1 = (z0 + 1) % n;
    const float a = f[(size_t)z0 * n + x0];
    const float b = f[(size_t)z0 * n + x1];
    const float c = f[(size_t)z1 * n + x0];
    const float dd = f[(size_t)z1 * n + x1];
    return a * (1.0f - tu) * (1.0f - tv) + b * tu * (1.0f - tv) + c * (1.0f - tu) * tv +
           dd * tu * tv;
}

/* foam.cpp:89-139 */
static void foam_step(orc_foam* m, const float* disp, const float* vel, float dt)
{
    const int ni = (int)m->n;
    const float inv_cell = 1.0f / m->cell;
    const float decay = expf(-m->p.decay * dt);
    const float gain = m->p.source_gain * dt;
    const float scale = m->p.advect_scale;
    for (uint32_t z = 0; z < m->n; ++z) {
        for (uint32_t x = 0; x < m->n; ++x) {
            const size_t i = (size_t)z * m->n + x;
            const float vx = vel[4 * i + 0] * scale + m->p.wind_drift_x;
            const float vz = vel[4 * i + 2] * scale + m->p.wind_drift_z;
            const float su = (float)x - vx * dt * inv_cell;
            const float sv = (float)z - vz * dt * inv_cell;
            float f = sample_wrap(m->cur, su, sv, ni);
            const float src = disp[4 * i + 3];
            f = f * decay + src * gain;
            if (f > 1.0f) f = 1.0f;
            if (f < 0.0f) f = 0.0f;
            m->next[i] = f;
        }
    }
}

EXPORT orc_foam* orc_foam_create(const orc_ocean* o, const orc_foam_params* p)
{
    if (!o->vel) return NULL;
    orc_foam* m = calloc(1, sizeof(*m));
    m->ocean = o;
    m->p = *p;
    m->n = o->desc.size;
    m->cells = o->cells;
    m->cell = o->desc.patch_length / (float)o->desc.size;
    m->a = calloc(m->cells, sizeof(float));
    m->b = calloc(m->cells, sizeof(float));
    m->cur = m->a;
    m->next = m->b;
    return m;
}

EXPORT void orc_foam_destroy(orc_foam* m)
{
    if (!m) return;
    free(m->a); free(m->b); free(m);
}

/* foam.cpp:218-252 */
EXPORT void orc_foam_update(orc_foam* m, float dt)
{
    if (!(dt > 0.0f)) dt = 0.0f;
    m->accum += (double)dt;
    const double step = (double)m->p.fixed_dt;
    const double eps = step * 1e-6;
    uint32_t steps = 0;
    if (m->accum + eps >= step) steps = (uint32_t)((m->accum + eps) / step);
    if (steps > m->p.max_substeps) {
        steps = m->p.max_substeps;
        m->accum = 0.0;
    } else {
        m->accum -= (double)steps * step;
        if (m->accum < 0.0) m->accum = 0.0;
    }
    m->last = steps;
    for (uint32_t s = 0; s < steps; ++s) {
        foam_step(m, m->ocean->disp, m->ocean->vel, m->p.fixed_dt);
        float* t = m->cur; m->cur = m->next; m->next = t;
    }
}

EXPORT const float* orc_foam_data(const orc_foam* m) { return m->cur; }
EXPORT uint32_t orc_foam_last_substeps(const orc_foam* m) { return m->last; }

/* foam.cpp:291-305 */
EXPORT float orc_foam_at(const orc_foam* m, float wx, float wz)
{
    ocean_surface_ex s;
    orc_ocean_sample(m->ocean, wx, wz, &s);
    const float u = (wx - s.offset_x) / m->cell;
    const float v = (wz - s.offset_z) / m->cell;
    return sample_wrap(m->cur, u, v, (int)m->n);
}

/* foam.cpp:307-313: the sequential double sum of the field over the cells. */
EXPORT double orc_foam_coverage(const orc_foam* m)
{
    double sum = 0.0;
    for (size_t i = 0; i < m->cells; ++i) sum += m->cur[i];
    return sum / (double)m->cells;
}

// The reference's behavioural tests of the interaction field (iWave), restated
// against this library's InteractionField / WaterSurface and the
// ocean_interaction_* C functions on the GPU. Each case names the reference
// test it follows (tests/test_interaction.cpp, tests/test_c_interaction.cpp)
// and keeps its tolerance. The kernel's realised dispersion comes from the
// library's own tap fit (ocean_amd_build_iwave_taps), as the reference's
// test takes it from its own. Not restated: the SIMD-vs-scalar step test
// (the device step is compared with the oracle's scalar step bit for bit in
// tests/test_gpu_interaction.py).
#include "harness.hpp"

#include "ocean/interaction.hpp"
#include "ocean/ocean.h"
#include "ocean/ocean.hpp"
#include "ocean/ocean_amd.h"

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

using namespace ocean;
using namespace cases;

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr float kFrame = 1.0f / 60.0f;

InteractionDesc base_desc(std::uint32_t n = 256, float extent = 64.0f)  // test_interaction.cpp:28-38
{
    InteractionDesc d;
    d.size = n;
    d.extent = extent;
    d.kernel_radius = 6;
    d.damping = 0.0f;
    d.absorb_cells = 16;
    d.max_sources = 32;
    return d;
}

Disturbance rock(float x, float z, float radius = 0.5f, float strength = 0.2f)
{
    Disturbance d;
    d.world_x = x;
    d.world_z = z;
    d.radius = radius;
    d.strength = strength;
    d.kind = SourceKind::Impulse;
    return d;
}

void run(InteractionField& f, int frames, float dt = kFrame)
{
    for (int i = 0; i < frames; ++i) f.update(dt);
}

std::size_t words(const InteractionBuffers& b) { return std::size_t(b.size) * b.size * 4; }

float height(const InteractionBuffers& b, std::uint32_t x, std::uint32_t z)
{
    return b.field[4 * (std::size_t(z) * b.size + x)];
}

double net_volume(const InteractionBuffers& b)
{
    double v = 0.0;
    for (std::size_t i = 0; i < std::size_t(b.size) * b.size; ++i) v += b.field[4 * i];
    return v;
}

double sum_squares(const InteractionBuffers& b)
{
    double e = 0.0;
    for (std::size_t i = 0; i < std::size_t(b.size) * b.size; ++i) e += double(b.field[4 * i]) * b.field[4 * i];
    return e;
}

// Energy-weighted deep-water group velocity of a Ricker impulse of scale
// sigma: radial energy density k^5 exp(-sigma^2 k^2), c_g = sqrt(g/k) / 2,
// integrated to the grid's k_max (test_interaction.cpp:66-87).
double ideal_group_velocity(double sigma, double gravity, double k_max)
{
    double num = 0.0, den = 0.0;
    for (int i = 1; i <= 20000; ++i) {
        const double k = k_max * i / 20000.0;
        const double w = std::pow(k, 5.0) * std::exp(-sigma * sigma * k * k);
        num += w * 0.5 * std::sqrt(gravity / k);
        den += w;
    }
    return num / den;
}

// The same weighting over the group velocity the solver should deliver: the
// derivative of omega(k) from the fitted stencil's realised symbol
// S(k) = sum tap(i, j) cos(k i dx) along x, through the leapfrog's discrete
// dispersion sin(omega dt / 2) = (dt / 2) sqrt(g S) (test_interaction.cpp:89-131).
double realised_group_velocity(double sigma, double gravity, double cell, std::uint32_t radius,
                               double dt)
{
    const int p = int(radius), w = 2 * p + 1;
    std::vector<float> taps(std::size_t(w) * w);
    double max_symbol = 0.0;
    if (ocean_amd_build_iwave_taps(radius, cell, 0.0, taps.data(), &max_symbol) != w * w)
        fail("ocean_amd_build_iwave_taps(%u) failed", radius);
    auto symbol = [&](double k) {
        double s = 0.0;
        for (int j = -p; j <= p; ++j)
            for (int i = -p; i <= p; ++i) s += double(taps[std::size_t(j + p) * w + (i + p)]) * std::cos(k * i * cell);
        return s;
    };
    auto omega = [&](double k) {
        const double s = symbol(k);
        if (s <= 0.0) return 0.0;
        const double a = 0.5 * dt * std::sqrt(gravity * s);
        return a >= 1.0 ? 2.0 / dt : (2.0 / dt) * std::asin(a);
    };
    const double k_max = kPi / cell;
    const int steps = 4000;
    const double h = k_max / steps;
    double num = 0.0, den = 0.0;
    for (int i = 2; i < steps; ++i) {
        const double k = k_max * i / steps;
        const double wgt = std::pow(k, 5.0) * std::exp(-sigma * sigma * k * k);
        num += wgt * (omega(k + h) - omega(k - h)) / (2.0 * h);
        den += wgt;
    }
    return num / den;
}

// Radius of the field's energy centroid about the grid centre, in metres.
double centroid_radius(const InteractionBuffers& b)
{
    const double cell = double(b.extent) / b.size, c = 0.5 * b.size;
    double num = 0.0, den = 0.0;
    for (std::uint32_t z = 0; z < b.size; ++z)
        for (std::uint32_t x = 0; x < b.size; ++x) {
            const double e = height(b, x, z), w = e * e;
            const double dx = (x + 0.5) - c, dz = (z + 0.5) - c;
            num += w * std::sqrt(dx * dx + dz * dz) * cell;
            den += w;
        }
    return den > 0.0 ? num / den : 0.0;
}

// Ring speed between frames 120 and 120 + frames of a centred impulse.
double ring_speed(InteractionDesc d, float sigma, int frames)
{
    InteractionField f{d};
    f.add(rock(32.0f, 32.0f, sigma, 0.2f));
    f.update(kFrame);
    run(f, 119);
    const double r0 = centroid_radius(f.buffers());
    run(f, frames);
    const double r1 = centroid_radius(f.buffers());
    return (r1 - r0) / (frames / 60.0);
}

template <class F>
bool desc_refused(F mutate)
{
    InteractionDesc d = base_desc();
    mutate(d);
    try {
        InteractionField f{d};
    } catch (const std::invalid_argument&) {
        return true;
    }
    return false;
}

void iwave_desc_validation()  // tests/test_interaction.cpp:166-193
{
    EXPECT(!desc_refused([](InteractionDesc&) {}));
    EXPECT(desc_refused([](InteractionDesc& d) { d.size = 100; }));
    EXPECT(desc_refused([](InteractionDesc& d) { d.size = 16; }));
    EXPECT(desc_refused([](InteractionDesc& d) { d.size = 2048; }));
    EXPECT(desc_refused([](InteractionDesc& d) { d.extent = 0.0f; }));
    EXPECT(desc_refused([](InteractionDesc& d) { d.kernel_radius = 0; }));
    EXPECT(desc_refused([](InteractionDesc& d) { d.kernel_radius = 13; }));
    EXPECT(desc_refused([](InteractionDesc& d) { d.damping = -1.0f; }));
    EXPECT(desc_refused([](InteractionDesc& d) { d.max_substeps = 0; }));
    EXPECT(desc_refused([](InteractionDesc& d) { d.gravity = 0.0f; }));
    EXPECT(desc_refused([](InteractionDesc& d) { d.absorb_cells = 200; }));
    EXPECT(desc_refused([](InteractionDesc& d) { d.fixed_dt = 10.0f; }));
}

void iwave_stability_limit()  // tests/test_interaction.cpp:195-223
{
    InteractionField f{base_desc()};
    const float limit = f.stable_dt_limit(), dt = f.fixed_dt();
    EXPECT(limit > 0.0f && dt > 0.0f && dt < 0.5f * limit);
    const double a = 0.5 * (2.0 / limit) * dt;
    const double phase_err = std::asin(a) / a - 1.0;
    info("dt %.5f s, limit %.5f s, kernel dispersion error %.2f%%, leapfrog phase error %.4f%%", dt, limit,
         100.0 * f.kernel_dispersion_error(), 100.0 * phase_err);
    EXPECT(phase_err < 0.02);
}

void iwave_impulse_crater()  // tests/test_interaction.cpp:229-270
{
    InteractionDesc d = base_desc();
    d.absorb_cells = 0;
    InteractionField f{d};
    f.add(rock(32.0f, 32.0f, 0.8f, 0.3f));
    f.update(kFrame);
    const InteractionBuffers b = f.buffers();
    const std::uint32_t c = b.size / 2;
    EXPECT(height(b, c, c) < 0.0f);
    float rim = 0.0f;
    for (std::uint32_t r = 3; r < 16; ++r) rim = std::max(rim, height(b, c + r, c));
    EXPECT(rim > 0.0f && rim < 0.5f * std::fabs(height(b, c, c)));
    double scale = 0.0;
    for (std::size_t i = 0; i < std::size_t(b.size) * b.size; ++i) scale += std::fabs(double(b.field[4 * i]));
    info("net volume %.3e of %.3e", net_volume(b), scale);
    EXPECT(std::fabs(net_volume(b)) < 1e-3 * scale);
}

void iwave_radius_selects_wavelength()  // tests/test_interaction.cpp:272-301
{
    const double small = ring_speed(base_desc(), 0.30f, 240), large = ring_speed(base_desc(), 0.50f, 240);
    const double expected = std::sqrt(0.50 / 0.30);
    info("ring speed %.3f / %.3f m/s, ratio %.3f (expected %.3f)", small, large, large / small, expected);
    EXPECT(large > small);
    EXPECT(approx(large / small, expected, 0.15));
}

void iwave_ring_group_velocity()  // tests/test_interaction.cpp:305-369
{
    const double cell = 64.0 / 256.0;
    for (float sigma : {0.30f, 0.40f, 0.50f, 0.60f, 0.80f}) {
        const InteractionDesc d = base_desc();
        const double measured = ring_speed(d, sigma, 360);
        const double realised = realised_group_velocity(sigma, d.gravity, cell, d.kernel_radius,
                                                        double(InteractionField{d}.fixed_dt()));
        const double ideal = ideal_group_velocity(sigma, d.gravity, kPi / cell);
        const double err_r = std::fabs(measured / realised - 1.0), err_i = std::fabs(measured / ideal - 1.0);
        info("sigma %.2f: measured %.4f, realised %.4f (%.2f%%), ideal %.4f (%.2f%%)", sigma, measured, realised,
             100.0 * err_r, ideal, 100.0 * err_i);
        EXPECT(measured > 0.0);
        EXPECT(err_r < 0.12);
        if (sigma <= 0.5f) EXPECT(err_i < 0.06);
    }
}

void iwave_larger_radius_improves()  // tests/test_interaction.cpp:371-400
{
    auto error_at = [](std::uint32_t p) {
        InteractionDesc d = base_desc();
        d.kernel_radius = p;
        const double theory = ideal_group_velocity(0.5, d.gravity, kPi / (d.extent / d.size));
        return std::fabs(ring_speed(d, 0.5f, 360) / theory - 1.0);
    };
    const double e4 = error_at(4), e8 = error_at(8);
    info("ring-speed error P=4 %.2f%%, P=8 %.2f%%", 100.0 * e4, 100.0 * e8);
    EXPECT(e8 < e4);
}

void iwave_energy_monotone()  // tests/test_interaction.cpp:402-421
{
    InteractionDesc d = base_desc();
    d.damping = 0.5f;
    InteractionField f{d};
    f.add(rock(32.0f, 32.0f, 0.6f, 0.3f));
    f.update(kFrame);
    double prev = f.energy();
    EXPECT(prev > 0.0);
    for (int i = 0; i < 200; ++i) {
        run(f, 3);
        const double e = f.energy();
        if (!(e <= prev * (1.0 + 1e-6) + 1e-12)) fail("energy rose at step %d: %.9e -> %.9e", i, prev, e);
        prev = e;
    }
    EXPECT(prev > 0.0);
}

void iwave_undamped_energy()  // tests/test_interaction.cpp:423-442
{
    InteractionDesc d = base_desc(128, 64.0f);
    d.absorb_cells = 0;
    InteractionField f{d};
    f.add(rock(32.0f, 32.0f, 1.0f, 0.1f));
    f.update(kFrame);
    const double e0 = f.energy();
    run(f, 120);
    const double e1 = f.energy();
    info("undamped energy drift over 2 s: %.4f%%", 100.0 * std::fabs(e1 / e0 - 1.0));
    EXPECT(approx(e1, e0, 0.02));
}

double central_energy(const InteractionBuffers& b, double frac)
{
    const double c = 0.5 * b.size, lim = frac * b.size;
    double e = 0.0;
    for (std::uint32_t z = 0; z < b.size; ++z)
        for (std::uint32_t x = 0; x < b.size; ++x) {
            const double dx = (x + 0.5) - c, dz = (z + 0.5) - c;
            if (std::sqrt(dx * dx + dz * dz) > lim) continue;
            e += double(height(b, x, z)) * height(b, x, z);
        }
    return e;
}

void iwave_absorbing_layer()  // tests/test_interaction.cpp:444-492
{
    InteractionDesc d = base_desc(128, 32.0f);
    InteractionField late{d};
    late.add(rock(16.0f, 16.0f, 0.5f, 0.2f));
    late.update(kFrame);
    run(late, 60 * 30);
    InteractionField early{d};
    early.add(rock(16.0f, 16.0f, 0.5f, 0.2f));
    early.update(kFrame);
    const double ratio = central_energy(late.buffers(), 0.20) / central_energy(early.buffers(), 0.20);
    info("energy back in the central 20%% after a round trip: %.4f%% of the initial", 100.0 * ratio);
    EXPECT(ratio < 0.01);
}

void iwave_reflects_without_layer()  // tests/test_interaction.cpp:494-524
{
    auto left = [](std::uint32_t absorb) {
        InteractionDesc d = base_desc(128, 32.0f);
        d.absorb_cells = absorb;
        InteractionField f{d};
        f.add(rock(16.0f, 16.0f, 0.5f, 0.2f));
        f.update(kFrame);
        run(f, 60 * 30);
        return sum_squares(f.buffers());
    };
    const double with_layer = left(16), without = left(0);
    info("energy left after 30 s: absorb=16 %.3e, absorb=0 %.3e", with_layer, without);
    EXPECT(without > 50.0 * with_layer);
}

void iwave_hammered()  // tests/test_interaction.cpp:526-546
{
    InteractionDesc d = base_desc();
    d.damping = 0.1f;
    InteractionField f{d};
    for (int i = 0; i < 600; ++i) {
        f.add(rock(32.0f + 6.0f * std::sin(i * 0.7f), 32.0f + 6.0f * std::cos(i * 0.5f), 0.4f, 50.0f));
        f.add(rock(32.0f, 32.0f, 0.3f, -80.0f));
        f.update(kFrame);
    }
    const InteractionBuffers b = f.buffers();
    for (std::size_t i = 0; i < words(b); ++i)
        if (!std::isfinite(b.field[i])) fail("word %zu not finite", i);
}

void iwave_stable_at_limit()  // tests/test_interaction.cpp:548-574
{
    const float limit = InteractionField{base_desc()}.stable_dt_limit();
    InteractionDesc d = base_desc();
    d.max_substeps = 64;
    d.fixed_dt = limit * 0.995f;
    InteractionField f{d};
    f.add(rock(32.0f, 32.0f, 0.5f, 1.0f));
    for (int i = 0; i < 400; ++i) f.update(d.fixed_dt);
    const InteractionBuffers b = f.buffers();
    double peak = 0.0;
    for (std::size_t i = 0; i < std::size_t(b.size) * b.size; ++i) {
        if (!std::isfinite(b.field[4 * i])) fail("cell %zu not finite", i);
        peak = std::max(peak, std::fabs(double(b.field[4 * i])));
    }
    info("dt %.4f s (99.5%% of the limit): peak |eta| %.4f m", d.fixed_dt, peak);
    EXPECT(peak < 100.0);
}

void iwave_huge_dt_clamped()  // tests/test_interaction.cpp:576-598
{
    InteractionDesc d = base_desc();
    d.max_substeps = 4;
    InteractionField f{d};
    f.add(rock(32.0f, 32.0f));
    f.update(10.0f);
    EXPECT(f.last_substeps() == 4 && f.last_update_clamped());
    const InteractionBuffers b = f.buffers();
    for (std::size_t i = 0; i < std::size_t(b.size) * b.size; ++i)
        if (!std::isfinite(b.field[4 * i])) fail("cell %zu not finite", i);
}

std::vector<float> frames_at(float dt, int frames)
{
    InteractionField f{base_desc(128, 64.0f)};
    f.add(rock(32.0f, 32.0f, 0.6f, 0.2f));
    run(f, frames, dt);
    const InteractionBuffers b = f.buffers();
    return copy_of(b.field, words(b));
}

void iwave_frame_rate_independent()  // tests/test_interaction.cpp:600-623
{
    const std::vector<float> a = frames_at(1.0f / 60.0f, 120), b = frames_at(1.0f / 30.0f, 60);
    EXPECT(same_bits(a.data(), b.data(), a.size()));
}

struct Event {
    int frame;
    Disturbance d;
};

std::vector<Event> event_log()  // tests/test_interaction.cpp:668-686
{
    std::vector<Event> log;
    for (int i = 0; i < 40; ++i) {
        Disturbance d = rock(20.0f + 0.7f * i, 25.0f + 0.4f * i, 0.35f + 0.01f * i, 0.15f);
        d.velocity_y = -1.5f - 0.05f * i;
        log.push_back({i * 5, d});
        Disturbance c = rock(30.0f, 30.0f + 0.2f * i, 0.5f, 0.05f);
        c.kind = SourceKind::Continuous;
        c.velocity_x = 1.2f;
        c.velocity_z = -0.4f;
        log.push_back({i * 5 + 1, c});
    }
    return log;
}

std::vector<float> replay(const std::vector<Event>& log, ParallelForFn hook, std::uint32_t threads)
{
    InteractionDesc d = base_desc();
    d.damping = 0.05f;
    d.parallel_for = hook;
    d.thread_count = threads;
    InteractionField f{d};
    std::size_t next = 0;
    for (int i = 0; i < 220; ++i) {
        while (next < log.size() && log[next].frame == i) f.add(log[next++].d);
        f.update(kFrame);
    }
    const InteractionBuffers b = f.buffers();
    return copy_of(b.field, words(b));
}

void iwave_replay_bitidentical()  // tests/test_interaction.cpp:690-724 (the hooks are never called)
{
    g_hook_calls = 0;
    const std::vector<Event> log = event_log();
    const std::vector<float> ref = replay(log, nullptr, 1);
    EXPECT(same_bits(ref.data(), replay(log, nullptr, 1).data(), ref.size()));
    for (std::uint32_t t : {2u, 3u, 4u, 8u, 16u}) EXPECT(same_bits(ref.data(), replay(log, nullptr, t).data(), ref.size()));
    EXPECT(same_bits(ref.data(), replay(log, &reverse_hook, 0).data(), ref.size()));
    EXPECT(same_bits(ref.data(), replay(log, &scattering_hook, 0).data(), ref.size()));
    EXPECT(g_hook_calls.load() == 0);
}

void iwave_queue_overflow()  // tests/test_interaction.cpp:726-743
{
    InteractionDesc d = base_desc();
    d.max_sources = 4;
    InteractionField f{d}, g{d};
    for (int i = 0; i < 10; ++i) f.add(rock(30.0f + i, 30.0f));
    EXPECT(f.dropped_sources() == 6);
    for (int i = 0; i < 10; ++i) g.add(rock(30.0f + i, 30.0f));
    f.update(kFrame);
    g.update(kFrame);
    EXPECT(same_bits(f.buffers().field, g.buffers().field, words(f.buffers())));
}

void iwave_impulse_not_lost()  // tests/test_interaction.cpp:745-767
{
    InteractionField f{base_desc()};
    f.add(rock(32.0f, 32.0f, 0.5f, 0.3f));
    f.update(0.001f);
    EXPECT(f.last_substeps() == 0);
    f.update(1.0f / 30.0f);
    EXPECT(f.last_substeps() > 0);
    const InteractionBuffers b = f.buffers();
    double sum = 0.0;
    for (std::size_t i = 0; i < std::size_t(b.size) * b.size; ++i) sum += std::fabs(double(b.field[4 * i]));
    EXPECT(sum > 0.0);
}

void iwave_recentre_exact()  // tests/test_interaction.cpp:858-913
{
    InteractionField f{base_desc()};
    f.add(rock(32.0f, 32.0f, 0.8f, 0.3f));
    f.update(kFrame);
    run(f, 60);
    const InteractionBuffers b0 = f.buffers();
    const std::uint32_t n = b0.size;
    const float cell = b0.extent / n, ox = b0.origin_x, oz = b0.origin_z;
    const std::vector<float> before = copy_of(b0.field, words(b0));
    f.recenter(32.0f + 10.0f * cell, 32.0f + 7.0f * cell);
    EXPECT(f.recentered() && f.last_shift_x() == 10 && f.last_shift_z() == 7);
    const InteractionBuffers b1 = f.buffers();
    int compared = 0, differing = 0;
    for (std::uint32_t z = 0; z + 7 < n; ++z)
        for (std::uint32_t x = 0; x + 10 < n; ++x) {
            ++compared;
            differing += height(b1, x, z) != before[4 * (std::size_t(z + 7) * n + (x + 10))];
        }
    info("shift (10, 7): %d cells compared, %d differing", compared, differing);
    EXPECT(compared > 50000 && differing == 0);
    for (std::uint32_t z = 0; z < n; ++z)
        for (std::uint32_t x = n - 10; x < n; ++x)
            if (height(b1, x, z) != 0.0f) fail("uncovered cell (%u, %u) not zero", x, z);
    EXPECT(approx(b1.origin_x, ox + 10.0f * cell, 1e-6) && approx(b1.origin_z, oz + 7.0f * cell, 1e-6));
}

void iwave_recentre_subcell()  // tests/test_interaction.cpp:915-945
{
    InteractionDesc d = base_desc();
    InteractionField f{d};
    f.recenter(32.0f, 32.0f);
    const float ox = f.buffers().origin_x;
    f.recenter(32.0f + 0.1f * (d.extent / d.size), 32.0f);
    EXPECT(!f.recentered() && f.buffers().origin_x == ox);
    f.add(rock(32.0f, 32.0f, 0.8f, 0.5f));
    f.update(kFrame);
    const float before = f.height_at(32.0f, 32.0f);
    EXPECT(before < 0.0f);
    f.recenter(40.0f, 36.0f);
    f.update(kFrame);
    const float after = f.height_at(32.0f, 32.0f);
    info("height at the impact %.5f m, after recentring %.5f m", before, after);
    EXPECT(after < 0.0f && approx(after, before, 0.5));
}

double wall_response(Obstruction mode)  // tests/test_interaction.cpp:947-995
{
    InteractionDesc d = base_desc(128, 32.0f);
    d.absorb_cells = 8;
    d.obstruction = mode;
    InteractionField f{d};
    std::vector<std::uint8_t> mask(std::size_t(d.size) * d.size, 0);
    const std::uint32_t wall = d.size * 3 / 4;
    for (std::uint32_t z = 0; z < d.size; ++z)
        for (std::uint32_t x = wall; x < d.size; ++x) mask[std::size_t(z) * d.size + x] = 1;
    f.set_obstruction(mask.data());
    f.add(rock((d.extent / d.size) * float(wall - 20), 16.0f, 0.5f, 0.2f));
    f.update(kFrame);
    double best = 0.0;
    for (int i = 0; i < 400; ++i) {
        f.update(kFrame);
        if (i < 40) continue;
        const InteractionBuffers b = f.buffers();
        double s = 0.0;
        for (std::uint32_t z = d.size / 2 - 6; z < d.size / 2 + 6; ++z)
            for (std::uint32_t x = wall - 10; x < wall - 2; ++x) s += height(b, x, z);
        if (std::fabs(s) > std::fabs(best)) best = s;
    }
    return best;
}

void iwave_neumann_vs_dirichlet()  // tests/test_interaction.cpp:947-1004
{
    const double neumann = wall_response(Obstruction::Neumann), dirichlet = wall_response(Obstruction::Dirichlet);
    info("peak signed field before the wall: Neumann %+.5f, Dirichlet %+.5f", neumann, dirichlet);
    EXPECT(std::fabs(neumann) > 0.0 && std::fabs(dirichlet) > 0.0 && neumann != dirichlet);
}

void iwave_obstruction_changes_field()  // tests/test_interaction.cpp:1006-1032
{
    const InteractionDesc d = base_desc(128, 32.0f);
    InteractionField f{d}, g{d};
    std::vector<std::uint8_t> mask(std::size_t(d.size) * d.size, 0);
    for (std::uint32_t z = 40; z < 88; ++z) mask[std::size_t(z) * d.size + 80] = 1;
    f.set_obstruction(mask.data());
    g.set_obstruction(nullptr);
    for (InteractionField* h : {&f, &g}) {
        h->add(rock(16.0f, 16.0f, 0.5f, 0.3f));
        h->update(kFrame);
        run(*h, 120);
    }
    EXPECT(!same_bits(f.buffers().field, g.buffers().field, words(f.buffers())));
}

void iwave_queries_agree_with_buffer()  // tests/test_interaction.cpp:1034-1064
{
    InteractionField f{base_desc()};
    f.add(rock(32.0f, 32.0f, 0.8f, 0.3f));
    f.update(kFrame);
    run(f, 40);
    const InteractionBuffers b = f.buffers();
    const float cell = b.extent / b.size;
    double worst = 0.0;
    for (std::uint32_t z = 40; z < 200; z += 13)
        for (std::uint32_t x = 40; x < 200; x += 11) {
            const InteractionSample s = f.sample_at(b.origin_x + (x + 0.5f) * cell, b.origin_z + (z + 0.5f) * cell);
            worst = std::max(worst, std::fabs(double(s.height - height(b, x, z))));
        }
    info("query vs buffer at cell centres: worst %.3e m", worst);
    EXPECT(worst < 1e-6);
    EXPECT(f.height_at(10000.0f, 10000.0f) == 0.0f && f.height_at(-10000.0f, 0.0f) == 0.0f);
}

void iwave_update_allocation_free()  // tests/test_interaction.cpp:1066-1084
{
    InteractionField f{base_desc()};
    std::vector<std::uint8_t> mask(std::size_t(256) * 256, 0);
    for (std::size_t z = 100; z < 120; ++z)
        for (std::size_t x = 100; x < 140; ++x) mask[z * 256 + x] = 1;
    f.set_obstruction(mask.data());
    f.add(rock(32.0f, 32.0f));
    f.update(kFrame);
    const std::size_t before = library_allocations();
    for (int i = 0; i < 30; ++i) {
        f.add(rock(30.0f + 0.1f * i, 33.0f, 0.4f, 0.1f));
        Disturbance c = rock(31.0f, 31.0f, 0.5f, 0.05f);
        c.kind = SourceKind::Continuous;
        f.add(c);
        f.update(kFrame);
        f.recenter(32.0f + 0.5f * i, 32.0f);  // whole-cell shifts, mask included
    }
    EXPECT(library_allocations() == before);
}

void iwave_exact_one_substep()  // tests/test_interaction.cpp:1086-1100
{
    InteractionField f{base_desc()};
    const float dt = f.fixed_dt();
    f.add(rock(32.0f, 32.0f));
    for (int i = 0; i < 240; ++i) {
        f.update(dt);
        if (f.last_substeps() != 1) fail("frame %d ran %u substeps", i, f.last_substeps());
    }
}

// ---- tests/test_c_interaction.cpp ----------------------------------------------

void cinter_defaults()  // tests/test_c_interaction.cpp:16-40
{
    ocean_interaction_desc c;
    ocean_interaction_desc_init(&c);
    const InteractionDesc d{};
    EXPECT(c.struct_size == sizeof(ocean_interaction_desc) && c.size == d.size && c.extent == d.extent);
    EXPECT(c.kernel_radius == d.kernel_radius && c.damping == d.damping && c.absorb_cells == d.absorb_cells);
    EXPECT(c.fixed_dt == d.fixed_dt && c.max_substeps == d.max_substeps && c.max_sources == d.max_sources);
    EXPECT(c.gravity == d.gravity && c.water_depth == d.depth && c.thread_count == d.thread_count);
    EXPECT(c.obstruction == OCEAN_OBSTRUCTION_NEUMANN && d.obstruction == Obstruction::Neumann);
}

void cinter_c_equals_cpp()  // tests/test_c_interaction.cpp:42-107
{
    ocean_interaction_desc cd;
    ocean_interaction_desc_init(&cd);
    cd.size = 128;
    cd.extent = 32.0f;
    cd.damping = 0.1f;
    cd.thread_count = 1;
    ocean_status st = OCEAN_ERROR_UNKNOWN;
    ocean_interaction* cf = ocean_interaction_create(&cd, &st);
    EXPECT(cf != nullptr && st == OCEAN_OK);
    InteractionDesc pd;
    pd.size = 128;
    pd.extent = 32.0f;
    pd.damping = 0.1f;
    pd.thread_count = 1;
    InteractionField pf{pd};
    for (int i = 0; i < 40; ++i) {
        if (i % 7 == 0) {
            ocean_disturbance d{};
            d.world_x = 12.0f + 0.3f * i;
            d.world_z = 14.0f;
            d.radius = 0.5f;
            d.strength = 0.2f;
            d.velocity_y = -1.0f;
            d.kind = OCEAN_SOURCE_IMPULSE;
            ocean_interaction_add(cf, &d);
            Disturbance p = rock(d.world_x, d.world_z, d.radius, d.strength);
            p.velocity_y = d.velocity_y;
            pf.add(p);
        }
        ocean_interaction_update(cf, kFrame);
        pf.update(kFrame);
    }
    const ocean_interaction_buffers cb = ocean_interaction_get_buffers(cf);
    const InteractionBuffers pb = pf.buffers();
    EXPECT(cb.size == pb.size && same_bits(cb.field, pb.field, words(pb)));
    EXPECT(ocean_interaction_fixed_dt(cf) == pf.fixed_dt() && ocean_interaction_dt_limit(cf) == pf.stable_dt_limit());
    ocean_interaction_sample sc{};
    ocean_interaction_sample_at(cf, 13.0f, 14.5f, &sc);
    const InteractionSample sp = pf.sample_at(13.0f, 14.5f);
    EXPECT(sc.height == sp.height && sc.slope_x == sp.slope_x && sc.velocity_y == sp.velocity_y);
    ocean_interaction_destroy(cf);
}

void cinter_legacy_desc()  // tests/test_c_interaction.cpp:109-182
{
    struct DescBeforeObstruction {  // ocean_interaction_desc up to water_depth
        std::size_t struct_size;
        std::uint32_t size;
        float extent;
        std::uint32_t kernel_radius;
        float damping;
        std::uint32_t absorb_cells;
        float fixed_dt;
        std::uint32_t max_substeps;
        std::uint32_t max_sources;
        float gravity;
        float water_depth;
    };
    static_assert(offsetof(DescBeforeObstruction, water_depth) == offsetof(ocean_interaction_desc, water_depth));
    DescBeforeObstruction legacy;
    std::memset(&legacy, 0, sizeof legacy);
    legacy.struct_size = sizeof legacy;
    legacy.size = 128;
    legacy.extent = 32.0f;
    legacy.kernel_radius = 6;
    legacy.damping = 0.1f;
    legacy.absorb_cells = 16;
    legacy.max_substeps = 8;
    legacy.max_sources = 64;
    legacy.gravity = 9.81f;
    ocean_status st = OCEAN_ERROR_UNKNOWN;
    ocean_interaction* old_caller =
        ocean_interaction_create(reinterpret_cast<const ocean_interaction_desc*>(&legacy), &st);
    EXPECT(old_caller != nullptr && st == OCEAN_OK);
    ocean_interaction_desc current;
    ocean_interaction_desc_init(&current);
    current.size = 128;
    current.extent = 32.0f;
    current.damping = 0.1f;
    ocean_interaction* new_caller = ocean_interaction_create(&current, &st);
    EXPECT(new_caller != nullptr);
    ocean_disturbance d{};
    d.world_x = 16.0f;
    d.world_z = 16.0f;
    d.radius = 0.5f;
    d.strength = 0.3f;
    d.kind = OCEAN_SOURCE_IMPULSE;
    ocean_interaction_add(old_caller, &d);
    ocean_interaction_add(new_caller, &d);
    for (int i = 0; i < 20; ++i) {
        ocean_interaction_update(old_caller, kFrame);
        ocean_interaction_update(new_caller, kFrame);
    }
    const ocean_interaction_buffers a = ocean_interaction_get_buffers(old_caller);
    const std::vector<float> fa = copy_of(a.field, std::size_t(a.size) * a.size * 4);
    const ocean_interaction_buffers b = ocean_interaction_get_buffers(new_caller);
    EXPECT(same_bits(fa.data(), b.field, fa.size()));
    ocean_interaction_destroy(old_caller);
    ocean_interaction_destroy(new_caller);
}

void cinter_future_struct_size()  // tests/test_c_interaction.cpp:184-198
{
    ocean_interaction_desc d;
    ocean_interaction_desc_init(&d);
    d.struct_size = sizeof(ocean_interaction_desc) + 16;
    ocean_status st = OCEAN_OK;
    EXPECT(ocean_interaction_create(&d, &st) == nullptr && st == OCEAN_ERROR_INVALID_ARG);
    d.struct_size = 0;
    EXPECT(ocean_interaction_create(&d, &st) == nullptr && st == OCEAN_ERROR_INVALID_ARG);
}

void cinter_surface_frozen_velocity_ex()  // tests/test_c_interaction.cpp:200-247
{
    EXPECT(sizeof(ocean_surface) == 7 * sizeof(float));
    ocean_desc d;
    ocean_desc_init(&d);
    d.size = 64;
    d.seed = 99;
    d.compute_velocity = 1;
    ocean_status st = OCEAN_ERROR_UNKNOWN;
    ocean_sim* sim = ocean_create(&d, &st);
    EXPECT(sim != nullptr);
    ocean_update(sim, 5.0);
    const ocean_surface base = ocean_sample_at(sim, 3.0f, 4.0f);
    ocean_surface_ex ex{};
    ocean_sample_ex(sim, 3.0f, 4.0f, &ex);
    EXPECT(ex.height == base.height && ex.offset_x == base.offset_x && ex.normal_y == base.normal_y &&
           ex.foam == base.foam);
    EXPECT(ocean_get_velocity(sim) != nullptr);
    OceanDesc pd;
    pd.size = 64;
    pd.seed = 99;
    pd.compute_velocity = true;
    Ocean psim{pd};
    psim.update(5.0);
    const Surface ps = psim.sample_at(3.0f, 4.0f);
    EXPECT(ex.velocity_x == ps.velocity_x && ex.velocity_y == ps.velocity_y && ex.velocity_z == ps.velocity_z);
    ocean_destroy(sim);
}

void cinter_velocity_absent()  // tests/test_c_interaction.cpp:249-267
{
    ocean_desc d;
    ocean_desc_init(&d);
    d.size = 64;
    EXPECT(d.compute_velocity == 0);
    ocean_sim* sim = ocean_create(&d, nullptr);
    EXPECT(sim != nullptr);
    ocean_update(sim, 1.0);
    EXPECT(ocean_get_velocity(sim) == nullptr);
    ocean_surface_ex ex{};
    ocean_sample_ex(sim, 1.0f, 1.0f, &ex);
    EXPECT(ex.velocity_x == 0.0f && ex.velocity_y == 0.0f && ex.velocity_z == 0.0f);
    ocean_destroy(sim);
}

void cinter_combined_matches_water_surface()  // tests/test_c_interaction.cpp:269-330
{
    ocean_desc od;
    ocean_desc_init(&od);
    od.size = 64;
    od.seed = 7;
    od.compute_velocity = 1;
    ocean_sim* sim = ocean_create(&od, nullptr);
    EXPECT(sim != nullptr);
    ocean_update(sim, 3.0);
    ocean_interaction_desc id;
    ocean_interaction_desc_init(&id);
    id.size = 128;
    id.extent = 32.0f;
    id.thread_count = 1;
    ocean_interaction* field = ocean_interaction_create(&id, nullptr);
    EXPECT(field != nullptr);
    ocean_disturbance dd{};
    dd.world_x = 16.0f;
    dd.world_z = 16.0f;
    dd.radius = 0.6f;
    dd.strength = 0.4f;
    dd.kind = OCEAN_SOURCE_IMPULSE;
    ocean_interaction_add(field, &dd);
    for (int i = 0; i < 25; ++i) ocean_interaction_update(field, kFrame);
    ocean_surface_ex c{};
    ocean_sample_combined(sim, field, 16.5f, 16.2f, &c);

    OceanDesc pod;
    pod.size = 64;
    pod.seed = 7;
    pod.compute_velocity = true;
    Ocean psim{pod};
    psim.update(3.0);
    InteractionDesc pid;
    pid.size = 128;
    pid.extent = 32.0f;
    pid.thread_count = 1;
    InteractionField pfield{pid};
    pfield.add(rock(16.0f, 16.0f, 0.6f, 0.4f));
    for (int i = 0; i < 25; ++i) pfield.update(kFrame);
    const WaterSurface water{psim, pfield};
    const Surface ps = water.sample_at(16.5f, 16.2f);
    EXPECT(c.height == ps.height && c.normal_x == ps.normal_x && c.normal_y == ps.normal_y &&
           c.velocity_y == ps.velocity_y);
    EXPECT(std::fabs(c.height - psim.sample_at(16.5f, 16.2f).height) > 1e-6f);
    ocean_interaction_destroy(field);
    ocean_destroy(sim);
}

void cinter_null_handles()  // tests/test_c_interaction.cpp:332-358
{
    ocean_interaction_destroy(nullptr);
    ocean_interaction_update(nullptr, 0.016f);
    ocean_interaction_add(nullptr, nullptr);
    ocean_interaction_set_obstruction(nullptr, nullptr);
    EXPECT(ocean_interaction_dropped(nullptr) == 0u && ocean_interaction_recenter(nullptr, 0.0f, 0.0f) == 0);
    EXPECT(ocean_interaction_height_at(nullptr, 0.0f, 0.0f) == 0.0f && ocean_interaction_fixed_dt(nullptr) == 0.0f);
    EXPECT(ocean_interaction_dt_limit(nullptr) == 0.0f && ocean_interaction_dispersion_error(nullptr) == 0.0f);
    EXPECT(ocean_interaction_get_buffers(nullptr).field == nullptr && ocean_get_velocity(nullptr) == nullptr);
    ocean_interaction_sample s{};
    ocean_interaction_sample_at(nullptr, 0.0f, 0.0f, &s);
    EXPECT(s.height == 0.0f);
    ocean_surface_ex ex{};
    ocean_sample_ex(nullptr, 0.0f, 0.0f, &ex);
    EXPECT(ex.normal_y == 1.0f);
    ocean_sample_combined(nullptr, nullptr, 0.0f, 0.0f, &ex);
    EXPECT(ex.normal_y == 1.0f);
}

void cinter_overflow_and_recentre()  // tests/test_c_interaction.cpp:360-380
{
    ocean_interaction_desc d;
    ocean_interaction_desc_init(&d);
    d.size = 64;
    d.extent = 16.0f;
    d.max_sources = 2;
    ocean_interaction* f = ocean_interaction_create(&d, nullptr);
    EXPECT(f != nullptr);
    ocean_disturbance dd{};
    dd.radius = 0.4f;
    dd.strength = 0.1f;
    for (int i = 0; i < 5; ++i) ocean_interaction_add(f, &dd);
    EXPECT(ocean_interaction_dropped(f) == 3u);
    EXPECT(ocean_interaction_recenter(f, 0.0f, 0.0f) == 1);
    EXPECT(ocean_interaction_recenter(f, 0.0f, 0.0f) == 0);
    ocean_interaction_destroy(f);
}

}  // namespace

namespace cases {

const Case kInteractionCases[] = {
    {"iwave.desc_validation", iwave_desc_validation},
    {"iwave.stability_limit", iwave_stability_limit},
    {"iwave.impulse_crater", iwave_impulse_crater},
    {"iwave.radius_selects_wavelength", iwave_radius_selects_wavelength},
    {"iwave.ring_group_velocity", iwave_ring_group_velocity},
    {"iwave.larger_radius_improves", iwave_larger_radius_improves},
    {"iwave.energy_monotone", iwave_energy_monotone},
    {"iwave.undamped_energy", iwave_undamped_energy},
    {"iwave.absorbing_layer", iwave_absorbing_layer},
    {"iwave.reflects_without_layer", iwave_reflects_without_layer},
    {"iwave.hammered", iwave_hammered},
    {"iwave.stable_at_limit", iwave_stable_at_limit},
    {"iwave.huge_dt_clamped", iwave_huge_dt_clamped},
    {"iwave.frame_rate_independent", iwave_frame_rate_independent},
    {"iwave.replay_bitidentical", iwave_replay_bitidentical},
    {"iwave.queue_overflow", iwave_queue_overflow},
    {"iwave.impulse_not_lost", iwave_impulse_not_lost},
    {"iwave.recentre_exact", iwave_recentre_exact},
    {"iwave.recentre_subcell", iwave_recentre_subcell},
    {"iwave.neumann_vs_dirichlet", iwave_neumann_vs_dirichlet},
    {"iwave.obstruction_changes_field", iwave_obstruction_changes_field},
    {"iwave.queries_agree_with_buffer", iwave_queries_agree_with_buffer},
    {"iwave.update_allocation_free", iwave_update_allocation_free},
    {"iwave.exact_one_substep", iwave_exact_one_substep},
    {"cinter.defaults", cinter_defaults},
    {"cinter.c_equals_cpp", cinter_c_equals_cpp},
    {"cinter.legacy_desc", cinter_legacy_desc},
    {"cinter.future_struct_size", cinter_future_struct_size},
    {"cinter.surface_frozen_velocity_ex", cinter_surface_frozen_velocity_ex},
    {"cinter.velocity_absent", cinter_velocity_absent},
    {"cinter.combined_matches_water_surface", cinter_combined_matches_water_surface},
    {"cinter.null_handles", cinter_null_handles},
    {"cinter.overflow_and_recentre", cinter_overflow_and_recentre},
};
const std::size_t kInteractionCaseCount = sizeof(kInteractionCases) / sizeof(kInteractionCases[0]);

}  // namespace cases

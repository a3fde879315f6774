// The reference's behavioural tests of the per-frame path, restated against
// this library's C++ surface (ocean::Ocean, CascadeStack, FoamField) and its
// C ABI, on the GPU; interaction_cases.cpp does the same for the interaction
// field, and this file's main runs both tables. Each case names the reference
// test it follows (tests/<file>.cpp:<lines>) and checks the same property at
// the same tolerance. Two contracts differ by design and the cases say so where they
// meet them:
//   - the host scheduler hook and thread_count are accepted and never used
//     (the frame runs on the GPU), so where the reference counts the hook's
//     dispatches these cases check that it is never called and that the
//     results do not change;
//   - ocean_simd_level() names the one device path ("gfx950"), and forcing a
//     CPU level leaves it unchanged.
// The reference's allocation probe counts every operator new of the process;
// here it counts those made from this library's own code (the HIP runtime
// allocates its own command objects inside every launch), see
// library_allocations().
//
// Usage: reference_cases [--list | name-prefix]. Prints "RUN <case>" before
// and "PASS <case>" or "FAIL <case>: <reason>" after each case, and
// "INFO <case>: ..." for measured values; exits 1 if any case failed. Built
// and run by tests/test_gpu_reference_cases.py.
#include "harness.hpp"

#include "ocean/batch.hpp"
#include "ocean/cascade.hpp"
#include "ocean/foam.hpp"
#include "ocean/interaction.hpp"
#include "ocean/ocean.h"
#include "ocean/ocean.hpp"
#include "ocean/ocean_amd.h"

#include <dlfcn.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <new>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

using namespace ocean;
using namespace cases;

// The replacement pairs operator new with malloc and operator delete with
// free; g++ cannot see that pairing where it inlines the library's deletes.
#pragma GCC diagnostic ignored "-Wmismatched-new-delete"

namespace {

std::atomic<std::size_t> g_library_allocations{0};

bool called_from_library(void* ret) noexcept
{
    Dl_info info;
    return dladdr(ret, &info) != 0 && info.dli_fname != nullptr &&
           std::strstr(info.dli_fname, "liboceanlib_amd") != nullptr;
}

void* counted_alloc(std::size_t n, void* ret)
{
    if (called_from_library(ret)) g_library_allocations.fetch_add(1);
    void* p = std::malloc(n != 0 ? n : 1);
    if (p == nullptr) throw std::bad_alloc();
    return p;
}

}  // namespace

void* operator new(std::size_t n) { return counted_alloc(n, __builtin_return_address(0)); }
void* operator new[](std::size_t n) { return counted_alloc(n, __builtin_return_address(0)); }
void operator delete(void* p) noexcept { std::free(p); }
void operator delete[](void* p) noexcept { std::free(p); }
void operator delete(void* p, std::size_t) noexcept { std::free(p); }
void operator delete[](void* p, std::size_t) noexcept { std::free(p); }

std::size_t cases::library_allocations() { return g_library_allocations.load(); }

namespace {

std::size_t texels(const Buffers& b) { return std::size_t(b.size) * b.size * 4; }

bool same_surface(const Surface& a, const Surface& b)
{
    return std::memcmp(&a, &b, sizeof(Surface)) == 0;
}

double rms_height(const Buffers& b)
{
    const std::size_t cells = std::size_t(b.size) * b.size;
    double s = 0.0;
    for (std::size_t i = 0; i < cells; ++i) s += double(b.displacement[4 * i + 1]) * b.displacement[4 * i + 1];
    return std::sqrt(s / double(cells));
}

float normal_len2_error(const float* n)
{
    return std::fabs(n[0] * n[0] + n[1] * n[1] + n[2] * n[2] - 1.0f);
}

// Descriptors of the reference test files' fixtures.
OceanDesc make_desc(std::uint32_t n, float length, std::uint64_t seed, float chop, float wind,
                    float direction = 0.0f)
{
    OceanDesc d;
    d.size = n;
    d.patch_length = length;
    d.seed = seed;
    d.choppiness = chop;
    d.spectrum.wind_speed = wind;
    d.spectrum.wind_direction = direction;
    return d;
}

OceanDesc simulation_desc(std::uint32_t n)  // tests/test_simulation.cpp:14-26
{
    return make_desc(n, 200.0f, 20260917, 1.0f, 12.0f, 0.4f);
}

OceanDesc query_desc(std::uint32_t n, float chop)  // tests/test_query.cpp:13-23
{
    return make_desc(n, 200.0f, 4242, chop, 12.0f, 0.3f);
}

OceanDesc velocity_desc(std::uint32_t n)  // tests/test_velocity.cpp:24-34
{
    OceanDesc d = make_desc(n, 200.0f, 4242, 1.0f, 12.0f);
    d.compute_velocity = true;
    return d;
}

OceanDesc foam_ocean(std::uint32_t n)  // tests/test_foam.cpp:25-36
{
    OceanDesc d = make_desc(n, 200.0f, 31337, 1.4f, 16.0f);
    d.foam_threshold = 0.9f;
    d.compute_velocity = true;
    return d;
}

std::vector<OceanDesc> three_levels()  // tests/test_cascade.cpp:16-45
{
    std::vector<OceanDesc> v;
    const float length[3] = {800.0f, 150.0f, 25.0f};
    const std::uint64_t seed[3] = {1001, 2002, 3003};
    const float cutoff[3] = {15.0f, 4.0f, 0.5f};
    for (int i = 0; i < 3; ++i) {
        OceanDesc d = make_desc(128, length[i], seed[i], 1.0f, 12.0f, 0.3f);
        d.spectrum.small_wave_cutoff = cutoff[i];
        v.push_back(d);
    }
    return v;
}

ocean_desc c_desc(std::uint32_t n)  // tests/test_c_api.cpp:14-25
{
    ocean_desc d;
    ocean_desc_init(&d);
    d.size = n;
    d.patch_length = 200.0f;
    d.seed = 555777;
    d.choppiness = 1.0f;
    d.spectrum.wind_speed = 11.0f;
    d.spectrum.wind_direction = 0.25f;
    return d;
}

OceanDesc cpp_desc(std::uint32_t n)  // tests/test_c_api.cpp:27-37
{
    return make_desc(n, 200.0f, 555777, 1.0f, 11.0f, 0.25f);
}

// ---- tests/test_api.cpp ------------------------------------------------------

void api_buffers_aligned()  // tests/test_api.cpp:43-59
{
    OceanDesc d;
    d.size = 64;
    d.patch_length = 137.0f;
    Ocean sim{d};
    const Buffers b = sim.buffers();
    EXPECT(b.displacement != nullptr && b.normal != nullptr);
    EXPECT(b.size == 64 && b.patch_length == 137.0f);
    EXPECT(reinterpret_cast<std::uintptr_t>(b.displacement) % kBufferAlignment == 0);
    EXPECT(reinterpret_cast<std::uintptr_t>(b.normal) % kBufferAlignment == 0);
}

void api_constructed_at_t0()  // tests/test_api.cpp:61-86
{
    OceanDesc d;
    d.size = 32;
    Ocean sim{d};
    EXPECT(sim.time() == 0.0);
    const Buffers b = sim.buffers();
    bool any_height = false;
    for (std::size_t i = 0; i < std::size_t(b.size) * b.size; ++i) {
        any_height |= b.displacement[4 * i + 1] != 0.0f;
        if (normal_len2_error(b.normal + 4 * i) >= 1e-4f) fail("texel %zu normal not unit", i);
    }
    EXPECT(any_height);
    // and it is the t = 0 frame, bit for bit
    const std::vector<float> first = copy_of(b.displacement, texels(b));
    sim.update(0.0);
    EXPECT(same_bits(first.data(), sim.buffers().displacement, first.size()));
}

void api_zero_wind_flat()  // tests/test_api.cpp:88-106
{
    OceanDesc d;
    d.size = 32;
    d.spectrum.wind_speed = 0.0f;
    Ocean sim{d};
    sim.update(3.0);
    const Buffers b = sim.buffers();
    for (std::size_t i = 0; i < std::size_t(b.size) * b.size; ++i) {
        const float* p = b.displacement + 4 * i;
        const float* n = b.normal + 4 * i;
        if (p[0] != 0.0f || p[1] != 0.0f || p[2] != 0.0f || p[3] != 0.0f) fail("texel %zu displaced", i);
        if (n[1] != 1.0f || n[3] != 1.0f) fail("texel %zu normal/jacobian", i);
    }
}

void api_absolute_time()  // tests/test_api.cpp:139-149
{
    OceanDesc d;
    d.size = 32;
    Ocean sim{d};
    sim.update(5.0);
    EXPECT(sim.time() == 5.0);
    sim.update(1.0);
    EXPECT(sim.time() == 1.0);
    Ocean fresh{d};
    fresh.update(1.0);
    EXPECT(same_bits(fresh.buffers().displacement, sim.buffers().displacement, texels(sim.buffers())));
}

void api_movable_desc_roundtrip()  // tests/test_api.cpp:151-167
{
    OceanDesc d;
    d.size = 64;
    d.seed = 99;
    d.spectrum.wind_speed = 14.5f;
    d.spectrum.wind_direction = 1.25f;
    Ocean a{d};
    const float* original = a.buffers().displacement;
    Ocean b{std::move(a)};
    EXPECT(b.buffers().displacement == original);
    EXPECT(b.desc().seed == 99);
    EXPECT(b.desc().spectrum.wind_speed == 14.5f);
    EXPECT(b.desc().spectrum.wind_direction == 1.25f);
    Ocean c{make_desc(32, 50.0f, 1, 1.0f, 5.0f)};
    c = std::move(b);
    EXPECT(c.buffers().displacement == original && c.desc().seed == 99);
    c.update(2.0);
    Ocean again{d};
    again.update(2.0);
    EXPECT(same_bits(again.buffers().displacement, c.buffers().displacement, texels(c.buffers())));
}

void api_allocation_probe_installed()  // tests/test_api.cpp:108-117
{
    const std::size_t before = library_allocations();
    OceanDesc d;
    d.size = 64;
    Ocean sim{d};
    EXPECT(library_allocations() > before);
}

void api_update_allocation_free()  // tests/test_api.cpp:120-137, test_simulation.cpp:360-371
{
    for (std::uint32_t n : {16u, 64u, 128u, 256u}) {
        OceanDesc d;
        d.size = n;
        Ocean sim{d};
        sim.update(0.0);
        const std::size_t before = library_allocations();
        for (int i = 0; i < 16; ++i) sim.update(double(i) * 0.016);
        (void)sim.buffers();  // the host mirror exists from creation
        if (library_allocations() != before) fail("N=%u: %zu allocations", n, library_allocations() - before);
    }
}

void api_bad_desc_throws()  // src/ocean.cpp:28-59
{
    auto throws = [](OceanDesc d) {
        try {
            Ocean sim{d};
        } catch (const std::invalid_argument&) {
            return true;
        }
        return false;
    };
    OceanDesc d;
    d.size = 100;
    EXPECT(throws(d));
    d = OceanDesc{};
    d.size = 8192;
    EXPECT(throws(d));
    d = OceanDesc{};
    d.patch_length = 0.0f;
    EXPECT(throws(d));
    d = OceanDesc{};
    d.choppiness = -1.0f;
    EXPECT(throws(d));
    d = OceanDesc{};
    d.spectrum.gravity = 0.0f;
    EXPECT(throws(d));
    d = OceanDesc{};
    d.spectrum.wind_speed = std::nanf("");
    EXPECT(throws(d));
}

// ---- tests/test_query.cpp ----------------------------------------------------

void query_agrees_with_drawn_grid()  // tests/test_query.cpp:71-96
{
    for (float chop : {0.0f, 0.5f, 1.0f, 1.5f}) {
        Ocean sim{query_desc(128, chop)};
        sim.update(5.0);
        const Buffers b = sim.buffers();
        const std::uint32_t n = b.size;
        const float cell = b.patch_length / float(n);
        std::vector<float> xz;
        xz.reserve(std::size_t(n) * n * 2);
        for (std::uint32_t z = 0; z < n; ++z)
            for (std::uint32_t x = 0; x < n; ++x) {
                const std::size_t i = std::size_t(z) * n + x;
                xz.push_back(float(x) * cell + b.displacement[4 * i + 0]);
                xz.push_back(float(z) * cell + b.displacement[4 * i + 2]);
            }
        std::vector<Surface> out(std::size_t(n) * n);
        EXPECT(sim.sample_batch(xz.data(), n * n, out.data()));
        double worst = 0.0, total = 0.0;
        for (std::size_t i = 0; i < out.size(); ++i) {
            const double err = std::fabs(double(out[i].height) - b.displacement[4 * i + 1]);
            worst = std::max(worst, err);
            total += err;
        }
        const double mean = total / double(out.size());
        const double rms = rms_height(b);
        info("chop %.1f: mean %.3e worst %.3e rms height %.3f", chop, mean, worst, rms);
        EXPECT(mean < 0.02 * rms);
        EXPECT(worst < 0.25 * rms);
        // the single-point query is the same function
        for (std::size_t i = 0; i < out.size(); i += 997) {
            EXPECT(sim.height_at(xz[2 * i], xz[2 * i + 1]) == out[i].height);
        }
    }
}

void query_chop0_exact_lookup()  // tests/test_query.cpp:98-118
{
    Ocean sim{query_desc(64, 0.0f)};
    sim.update(2.0);
    const Buffers b = sim.buffers();
    const float cell = b.patch_length / float(b.size);
    for (std::uint32_t z = 0; z < b.size; z += 7)
        for (std::uint32_t x = 0; x < b.size; x += 5) {
            const std::size_t i = std::size_t(z) * b.size + x;
            const float q = sim.height_at(float(x) * cell, float(z) * cell);
            if (std::fabs(q - b.displacement[4 * i + 1]) >= 1e-5f) fail("(%u,%u) %g vs %g", x, z, q, b.displacement[4 * i + 1]);
        }
}

void query_sample_consistent_with_buffers()  // tests/test_query.cpp:120-140
{
    Ocean sim{query_desc(64, 0.0f)};
    sim.update(3.0);
    const Buffers b = sim.buffers();
    const float cell = b.patch_length / float(b.size);
    for (std::uint32_t z = 0; z < b.size; z += 9)
        for (std::uint32_t x = 0; x < b.size; x += 11) {
            const std::size_t i = std::size_t(z) * b.size + x;
            const Surface s = sim.sample_at(float(x) * cell, float(z) * cell);
            EXPECT(std::fabs(s.normal_x - b.normal[4 * i + 0]) < 1e-4f);
            EXPECT(std::fabs(s.normal_y - b.normal[4 * i + 1]) < 1e-4f);
            EXPECT(std::fabs(s.normal_z - b.normal[4 * i + 2]) < 1e-4f);
            EXPECT(std::fabs(s.foam - b.displacement[4 * i + 3]) < 1e-4f);
        }
}

void query_unit_normals()  // tests/test_query.cpp:142-163
{
    Ocean sim{query_desc(64, 1.2f)};
    sim.update(8.0);
    for (int i = 0; i < 500; ++i) {
        const Surface s = sim.sample_at(-450.0f + 1.7f * float(i), 380.0f - 2.3f * float(i));
        const float n[3] = {s.normal_x, s.normal_y, s.normal_z};
        if (normal_len2_error(n) >= 1e-4f) fail("query %d normal length", i);
    }
}

void query_periodic()  // tests/test_query.cpp:165-184
{
    const OceanDesc d = query_desc(64, 1.0f);
    Ocean sim{d};
    sim.update(4.0);
    const float L = d.patch_length;
    for (int i = 0; i < 40; ++i) {
        const float x = -137.0f + 9.3f * float(i);
        const float z = 64.0f - 5.1f * float(i);
        const float base = sim.height_at(x, z);
        EXPECT(std::fabs(sim.height_at(x + L, z) - base) < 1e-4f);
        EXPECT(std::fabs(sim.height_at(x, z + L) - base) < 1e-4f);
        EXPECT(std::fabs(sim.height_at(x - 3.0f * L, z + 2.0f * L) - base) < 1e-4f);
    }
}

void query_far_outside_finite()  // tests/test_query.cpp:186-200
{
    Ocean sim{query_desc(64, 1.0f)};
    sim.update(6.0);
    for (float m : {1.0f, 1e3f, 1e5f, -1e5f}) {
        const Surface s = sim.sample_at(m * 3.1f, m * -7.9f);
        EXPECT(std::isfinite(s.height) && std::isfinite(s.normal_y) && std::isfinite(s.foam));
        EXPECT(std::fabs(s.height) < 100.0f);
    }
}

void query_extreme_chop()  // tests/test_query.cpp:202-220
{
    Ocean sim{query_desc(64, 4.0f)};
    sim.update(7.0);
    std::vector<float> xz;
    for (int i = 0; i < 2000; ++i) {
        xz.push_back(-100.0f + 0.13f * float(i));
        xz.push_back(77.0f - 0.09f * float(i));
    }
    std::vector<Surface> out(2000);
    EXPECT(sim.sample_batch(xz.data(), 2000, out.data()));
    for (int i = 0; i < 2000; ++i) {
        EXPECT(std::isfinite(out[i].height) && std::isfinite(out[i].offset_x));
        EXPECT(std::fabs(out[i].height) < 100.0f);
        if (i % 191 == 0) EXPECT(same_surface(out[i], sim.sample_at(xz[2 * i], xz[2 * i + 1])));
    }
}

void query_follows_surface()  // tests/test_query.cpp:222-236
{
    Ocean sim{query_desc(64, 1.0f)};
    sim.update(0.0);
    const float h0 = sim.height_at(37.0f, 91.0f);
    sim.update(2.5);
    EXPECT(sim.height_at(37.0f, 91.0f) != h0);
    sim.update(0.0);
    EXPECT(sim.height_at(37.0f, 91.0f) == h0);
}

void query_flat_ocean()  // tests/test_query.cpp:238-252
{
    OceanDesc d = query_desc(32, 1.0f);
    d.spectrum.wind_speed = 0.0f;
    Ocean sim{d};
    sim.update(12.0);
    for (int i = 0; i < 50; ++i) {
        const Surface s = sim.sample_at(3.7f * float(i), -8.1f * float(i));
        EXPECT(s.height == 0.0f && s.offset_x == 0.0f && s.offset_z == 0.0f);
        EXPECT(s.normal_y == 1.0f && s.foam == 0.0f);
    }
}

// ---- tests/test_simulation.cpp -----------------------------------------------

void sim_bitidentical_instances()  // tests/test_simulation.cpp:42-61
{
    Ocean a{simulation_desc(64)};
    Ocean b{simulation_desc(64)};
    for (double t : {0.0, 1.5, 97.25, 3600.0}) {
        a.update(t);
        b.update(t);
        EXPECT(same_bits(a.buffers().displacement, b.buffers().displacement, texels(a.buffers())));
        EXPECT(same_bits(a.buffers().normal, b.buffers().normal, texels(a.buffers())));
    }
}

void sim_seek_backwards()  // tests/test_simulation.cpp:63-81
{
    Ocean sim{simulation_desc(64)};
    sim.update(12.5);
    const std::vector<float> snap = copy_of(sim.buffers().displacement, texels(sim.buffers()));
    sim.update(400.0);
    sim.update(0.125);
    sim.update(12.5);
    EXPECT(same_bits(snap.data(), sim.buffers().displacement, snap.size()));
}

void sim_different_seeds_same_statistics()  // tests/test_simulation.cpp:83-116
{
    OceanDesc da = simulation_desc(64), db = simulation_desc(64);
    db.seed = da.seed + 1;
    Ocean a{da};
    Ocean b{db};
    a.update(2.0);
    b.update(2.0);
    EXPECT(!same_bits(a.buffers().displacement, b.buffers().displacement, texels(a.buffers())));
    const double ra = rms_height(a.buffers()), rb = rms_height(b.buffers());
    info("rms %.4f vs %.4f", ra, rb);
    EXPECT(approx(ra, rb, 0.25));
}

// Parseval for the unnormalised inverse transform: the mean of h^2 over the
// grid equals the sum over k of |h~(k, t)|^2, rebuilt here from the library's
// own tables (h0, omega; h0c is the conjugate mirror of h0).
void parseval_at(std::uint32_t n, double t)
{
    OceanDesc d = simulation_desc(n);
    Ocean sim{d};
    sim.update(t);
    ocean_desc cd;
    ocean_desc_init(&cd);
    cd.size = d.size;
    cd.patch_length = d.patch_length;
    cd.seed = d.seed;
    cd.choppiness = d.choppiness;
    cd.spectrum.wind_speed = d.spectrum.wind_speed;
    cd.spectrum.wind_direction = d.spectrum.wind_direction;
    const std::size_t cells = std::size_t(n) * n;
    std::vector<float> h0(2 * cells), omega(cells);
    EXPECT(ocean_amd_build_spectrum(&cd, h0.data(), omega.data()) == 0);
    const double two_pi = 6.283185307179586476925286766559;
    double spectral = 0.0;
    for (std::uint32_t z = 0; z < n; ++z)
        for (std::uint32_t x = 0; x < n; ++x) {
            const std::size_t i = std::size_t(z) * n + x;
            const std::size_t m = std::size_t((n - z) % n) * n + (n - x) % n;
            const double ph = std::fmod(double(omega[i]) * t, two_pi);
            const double c = std::cos(ph), s = std::sin(ph);
            const double ar = h0[2 * i], ai = h0[2 * i + 1];
            const double br = h0[2 * m], bi = -double(h0[2 * m + 1]);
            // h0 e^{i w t} + h0c e^{-i w t}
            const double hr = ar * c - ai * s + br * c + bi * s;
            const double hi = ar * s + ai * c + bi * c - br * s;
            spectral += hr * hr + hi * hi;
        }
    const Buffers b = sim.buffers();
    double mean = 0.0, spatial = 0.0;
    for (std::size_t i = 0; i < cells; ++i) {
        mean += b.displacement[4 * i + 1];
        spatial += double(b.displacement[4 * i + 1]) * b.displacement[4 * i + 1];
    }
    mean /= double(cells);
    spatial /= double(cells);
    info("N=%u mean %.3e spatial %.6e spectral %.6e (rel %.2e)", n, mean, spatial, spectral,
         rel_close(spatial, spectral));
    EXPECT(std::fabs(mean) < 1e-4);
    EXPECT(approx(spatial, spectral, 1e-4));
}

void sim_parseval()  // tests/test_simulation.cpp:118-185, at 128^2 and 512^2
{
    parseval_at(128, 7.75);
    parseval_at(512, 7.75);
}

void sim_hs_plausible()  // tests/test_simulation.cpp:187-207
{
    Ocean sim{simulation_desc(256)};
    sim.update(10.0);
    const double hs = 4.0 * rms_height(sim.buffers());
    info("Hs %.3f m", hs);
    EXPECT(hs > 0.3 && hs < 6.0);
}

void sim_stronger_wind()  // tests/test_simulation.cpp:209-234
{
    double previous = 0.0;
    for (float u : {5.0f, 10.0f, 15.0f, 20.0f}) {
        OceanDesc d = simulation_desc(128);
        d.spectrum.wind_speed = u;
        Ocean sim{d};
        sim.update(3.0);
        const double rms = rms_height(sim.buffers());
        EXPECT(rms > previous);
        previous = rms;
    }
}

void sim_surface_moves()  // tests/test_simulation.cpp:236-249
{
    Ocean sim{simulation_desc(64)};
    sim.update(0.0);
    const std::vector<float> first = copy_of(sim.buffers().displacement, texels(sim.buffers()));
    sim.update(1.0);
    EXPECT(!same_bits(first.data(), sim.buffers().displacement, first.size()));
}

void sim_finite_unit_normals()  // tests/test_simulation.cpp:251-273
{
    OceanDesc d = simulation_desc(128);
    d.choppiness = 2.0f;
    Ocean sim{d};
    for (double t : {0.0, 0.5, 11.0, 250.0}) {
        sim.update(t);
        const Buffers b = sim.buffers();
        for (std::size_t i = 0; i < texels(b); ++i)
            if (!std::isfinite(b.displacement[i]) || !std::isfinite(b.normal[i])) fail("t=%g word %zu", t, i);
        for (std::size_t i = 0; i < texels(b) / 4; ++i)
            if (normal_len2_error(b.normal + 4 * i) >= 1e-4f) fail("t=%g texel %zu normal", t, i);
    }
}

void sim_chop0_flat_jacobian()  // tests/test_simulation.cpp:275-296
{
    OceanDesc d = simulation_desc(64);
    d.choppiness = 0.0f;
    Ocean sim{d};
    sim.update(4.0);
    const Buffers b = sim.buffers();
    for (std::size_t i = 0; i < std::size_t(b.size) * b.size; ++i) {
        EXPECT(b.displacement[4 * i + 0] == 0.0f && b.displacement[4 * i + 2] == 0.0f);
        EXPECT(b.normal[4 * i + 3] == 1.0f && b.displacement[4 * i + 3] == 0.0f);
        EXPECT(b.normal[4 * i + 1] > 0.0f);
    }
}

void sim_chop_only_horizontal()  // tests/test_simulation.cpp:298-324
{
    OceanDesc flat = simulation_desc(64), chop = simulation_desc(64);
    flat.choppiness = 0.0f;
    chop.choppiness = 1.5f;
    Ocean a{flat};
    Ocean b{chop};
    a.update(6.0);
    b.update(6.0);
    bool moved = false;
    for (std::size_t i = 0; i < 64u * 64u; ++i) {
        EXPECT(a.buffers().displacement[4 * i + 1] == b.buffers().displacement[4 * i + 1]);
        moved |= b.buffers().displacement[4 * i + 0] != 0.0f;
    }
    EXPECT(moved);
}

void sim_foam_ramp()  // tests/test_simulation.cpp:326-358
{
    OceanDesc d = simulation_desc(128);
    d.choppiness = 2.5f;
    d.foam_threshold = 0.6f;
    Ocean sim{d};
    sim.update(9.0);
    const Buffers b = sim.buffers();
    std::size_t foamy = 0;
    for (std::size_t i = 0; i < 128u * 128u; ++i) {
        const float j = b.normal[4 * i + 3], f = b.displacement[4 * i + 3];
        EXPECT(f >= 0.0f && f <= 1.0f);
        const float want = std::min(1.0f, std::max(0.0f, (0.6f - j) / 0.6f));
        if (std::fabs(f - want) >= 1e-5f) fail("texel %zu foam %g vs ramp %g", i, f, want);
        foamy += f > 0.0f;
    }
    info("%zu foamy texels", foamy);
    EXPECT(foamy > 0);
}

// ---- tests/test_velocity.cpp -------------------------------------------------

void vel_absent_unless_asked()  // tests/test_velocity.cpp:38-51
{
    OceanDesc d = velocity_desc(128);
    d.compute_velocity = false;
    Ocean sim{d};
    sim.update(3.0);
    EXPECT(sim.buffers().velocity == nullptr);
    const Surface s = sim.sample_at(12.0f, -7.0f);
    EXPECT(s.velocity_x == 0.0f && s.velocity_y == 0.0f && s.velocity_z == 0.0f);
}

void vel_no_bit_change()  // tests/test_velocity.cpp:53-76
{
    OceanDesc off = velocity_desc(128);
    off.compute_velocity = false;
    Ocean a{off};
    Ocean b{velocity_desc(128)};
    for (double t : {0.0, 1.5, 40.0, 1000.0}) {
        a.update(t);
        b.update(t);
        EXPECT(same_bits(a.buffers().displacement, b.buffers().displacement, texels(a.buffers())));
        EXPECT(same_bits(a.buffers().normal, b.buffers().normal, texels(a.buffers())));
    }
}

void vel_uy_is_dhdt()  // tests/test_velocity.cpp:78-123
{
    Ocean sim{velocity_desc(128)};
    const double t = 21.0, dt = 0.004;
    sim.update(t + dt);
    const std::vector<float> hp = copy_of(sim.buffers().displacement, texels(sim.buffers()));
    sim.update(t - dt);
    const std::vector<float> hm = copy_of(sim.buffers().displacement, texels(sim.buffers()));
    sim.update(t);
    const Buffers b = sim.buffers();
    EXPECT(b.velocity != nullptr);
    double num = 0.0, den = 0.0;
    for (std::size_t i = 0; i < std::size_t(b.size) * b.size; ++i) {
        const double fd = (double(hp[4 * i + 1]) - double(hm[4 * i + 1])) / (2.0 * dt);
        const double uy = b.velocity[4 * i + 1];
        num += (fd - uy) * (fd - uy);
        den += fd * fd;
    }
    const double rel = std::sqrt(num / den);
    info("u_y vs central difference: RMS relative %.3e", rel);
    EXPECT(rel < 1e-3);
}

void vel_circular_orbits()  // tests/test_velocity.cpp:125-162
{
    Ocean sim{velocity_desc(256)};
    sim.update(17.0);
    const Buffers b = sim.buffers();
    EXPECT(b.velocity != nullptr);
    double hx = 0.0, hy = 0.0, hz = 0.0;
    const std::size_t cells = std::size_t(b.size) * b.size;
    for (std::size_t i = 0; i < cells; ++i) {
        hx += double(b.velocity[4 * i + 0]) * b.velocity[4 * i + 0];
        hy += double(b.velocity[4 * i + 1]) * b.velocity[4 * i + 1];
        hz += double(b.velocity[4 * i + 2]) * b.velocity[4 * i + 2];
    }
    const double ratio = (hx + hz) / hy;
    info("horizontal/vertical orbital variance %.4f", ratio);
    EXPECT(approx(ratio, 1.0, 0.02));
    EXPECT(hy / double(cells) > 1e-4);
}

void vel_query_agrees_with_buffer()  // tests/test_velocity.cpp:164-186
{
    OceanDesc d = velocity_desc(128);
    d.choppiness = 0.0f;
    Ocean sim{d};
    sim.update(9.0);
    const Buffers b = sim.buffers();
    const float cell = b.patch_length / float(b.size);
    double worst = 0.0;
    for (std::uint32_t z = 10; z < 110; z += 7)
        for (std::uint32_t x = 10; x < 110; x += 5) {
            const Surface s = sim.sample_at(float(x) * cell, float(z) * cell);
            worst = std::max(worst, std::fabs(double(s.velocity_y) - b.velocity[4 * (std::size_t(z) * b.size + x) + 1]));
        }
    info("worst %.3e m/s", worst);
    EXPECT(worst < 1e-5);
}

void vel_deterministic_seek()  // tests/test_velocity.cpp:188-206
{
    Ocean a{velocity_desc(128)};
    Ocean b{velocity_desc(128)};
    a.update(55.0);
    b.update(55.0);
    const std::size_t n = texels(a.buffers());
    EXPECT(same_bits(a.buffers().velocity, b.buffers().velocity, n));
    a.update(12.0);
    a.update(55.0);
    EXPECT(same_bits(a.buffers().velocity, b.buffers().velocity, n));
}

void vel_cascade_sums()  // tests/test_velocity.cpp:208-245
{
    std::vector<OceanDesc> levels;
    for (int i = 0; i < 3; ++i) {
        OceanDesc d = velocity_desc(128);
        d.patch_length = 400.0f / (1.0f + float(i) * 2.0f);
        d.seed = 900 + std::uint64_t(i);
        levels.push_back(d);
    }
    CascadeStack stack{std::span<const OceanDesc>(levels)};
    stack.update(6.0);
    std::vector<Ocean> singles;
    for (const auto& d : levels) singles.emplace_back(d);
    for (auto& o : singles) o.update(6.0);
    double worst = 0.0;
    for (float x = -80.0f; x < 80.0f; x += 13.7f)
        for (float z = -80.0f; z < 80.0f; z += 11.3f) {
            const Surface c = stack.sample_at(x, z);
            float sx = 0.0f, sy = 0.0f, sz = 0.0f;
            for (auto& o : singles) {
                const Surface s = o.sample_at(x, z);
                sx += s.velocity_x;
                sy += s.velocity_y;
                sz += s.velocity_z;
            }
            worst = std::max({worst, std::fabs(double(c.velocity_x - sx)), std::fabs(double(c.velocity_y - sy)),
                              std::fabs(double(c.velocity_z - sz))});
        }
    info("worst %.3e m/s", worst);
    EXPECT(worst < 1e-5);
}

// ---- tests/test_foam.cpp -----------------------------------------------------

template <class F>
bool foam_throws(const Ocean& sim, F mutate)
{
    FoamDesc fd;
    mutate(fd);
    try {
        FoamField f{sim, fd};
    } catch (const std::invalid_argument&) {
        return true;
    }
    return false;
}

void foam_update_allocation_free()  // tests/test_foam.cpp:395-409
{
    Ocean sim{foam_ocean(128)};
    FoamField foam{sim, FoamDesc{}};
    sim.update(1.0);
    foam.update(1.0f / 60.0f);
    const std::size_t before = library_allocations();
    for (int i = 0; i < 60; ++i) {
        sim.update(2.0 + i / 60.0);
        foam.update(1.0f / 60.0f);
    }
    EXPECT(library_allocations() == before);
}

void foam_refuses()  // tests/test_foam.cpp:47-69
{
    OceanDesc d = foam_ocean(128);
    d.compute_velocity = false;
    Ocean no_velocity{d};
    EXPECT(foam_throws(no_velocity, [](FoamDesc&) {}));
    Ocean ok{foam_ocean(128)};
    EXPECT(!foam_throws(ok, [](FoamDesc&) {}));
    EXPECT(foam_throws(ok, [](FoamDesc& f) { f.decay = -1.0f; }));
    EXPECT(foam_throws(ok, [](FoamDesc& f) { f.source_gain = -1.0f; }));
    EXPECT(foam_throws(ok, [](FoamDesc& f) { f.fixed_dt = 0.0f; }));
    EXPECT(foam_throws(ok, [](FoamDesc& f) { f.max_substeps = 0; }));
}

void foam_decay_rate()  // tests/test_foam.cpp:71-113
{
    Ocean sim{foam_ocean(128)};
    sim.update(20.0);
    FoamDesc fd;
    fd.decay = 0.5f;
    fd.source_gain = 6.0f;
    fd.advect_scale = 0.0f;
    FoamField foam{sim, fd};
    for (int i = 0; i < 120; ++i) foam.update(1.0f / 60.0f);
    const double built = foam.coverage();
    EXPECT(built > 1e-4);
    fd.source_gain = 0.0f;
    foam.set_desc(fd);
    for (int i = 0; i < 120; ++i) foam.update(1.0f / 60.0f);
    const double after = foam.coverage();
    const double expected = built * std::exp(-double(fd.decay) * 2.0);
    info("%.6f -> %.6f (expected %.6f)", built, after, expected);
    EXPECT(approx(after, expected, 1e-3));
    EXPECT(after > 0.0);
}

// Builds foam with no decay and no drift, then switches the source off and
// drifts one whole cell in +x per step.
std::pair<std::vector<float>, FoamDesc> foam_for_shift(FoamField& foam, float cell, float dt)
{
    for (int i = 0; i < 60; ++i) foam.update(dt);
    const std::uint32_t n = foam.size();
    std::vector<float> before = copy_of(foam.data(), std::size_t(n) * n);
    FoamDesc fd = foam.desc();
    fd.source_gain = 0.0f;
    fd.wind_drift_x = cell / dt;
    foam.set_desc(fd);
    return {before, fd};
}

FoamDesc still_foam(float dt)
{
    FoamDesc fd;
    fd.decay = 0.0f;
    fd.source_gain = 5.0f;
    fd.advect_scale = 0.0f;
    fd.fixed_dt = dt;
    return fd;
}

void foam_exact_advection()  // tests/test_foam.cpp:115-162
{
    Ocean sim{foam_ocean(128)};
    sim.update(20.0);
    const float cell = 200.0f / 128.0f, dt = 1.0f / 60.0f;
    FoamField foam{sim, still_foam(dt)};
    auto [before, fd] = foam_for_shift(foam, cell, dt);
    double mean = 0.0;
    for (float v : before) mean += v;
    EXPECT(mean / double(before.size()) > 1e-4);
    const int steps = 5;
    for (int i = 0; i < steps; ++i) foam.update(dt);
    const float* after = foam.data();
    const std::uint32_t n = foam.size();
    std::size_t bad = 0;
    for (std::uint32_t z = 0; z < n; ++z)
        for (std::uint32_t x = 0; x < n; ++x) {
            const std::uint32_t sx = (x + n - std::uint32_t(steps) % n) % n;
            bad += after[std::size_t(z) * n + x] != before[std::size_t(z) * n + sx];
        }
    info("%zu of %u cells differ from the hand shift", bad, n * n);
    EXPECT(bad == 0);
}

void foam_wraps()  // tests/test_foam.cpp:164-194
{
    Ocean sim{foam_ocean(128)};
    sim.update(12.0);
    const float cell = 200.0f / 128.0f, dt = 1.0f / 60.0f;
    FoamField foam{sim, still_foam(dt)};
    auto [before, fd] = foam_for_shift(foam, cell, dt);
    for (std::uint32_t i = 0; i < foam.size(); ++i) foam.update(dt);
    EXPECT(same_bits(before.data(), foam.data(), before.size()));
}

void foam_outlives_jacobian()  // tests/test_foam.cpp:196-234
{
    Ocean sim{foam_ocean(256)};
    FoamDesc fd;
    fd.decay = 0.3f;
    fd.source_gain = 4.0f;
    FoamField foam{sim, fd};
    for (int i = 0; i < 400; ++i) {
        sim.update(30.0 + i / 60.0);
        foam.update(1.0f / 60.0f);
    }
    double worst = 1e9;
    for (int i = 0; i < 200; ++i) {
        sim.update(36.666 + i / 60.0);
        foam.update(1.0f / 60.0f);
        const Buffers b = sim.buffers();
        const std::size_t cells = std::size_t(b.size) * b.size;
        double instant = 0.0;
        for (std::size_t c = 0; c < cells; ++c) instant += b.displacement[4 * c + 3];
        instant /= double(cells);
        if (instant > 1e-6) worst = std::min(worst, foam.coverage() / instant);
    }
    info("persistent/instantaneous coverage, worst frame %.2fx", worst);
    EXPECT(worst > 1.0);
}

void foam_saturating_clamp()  // tests/test_foam.cpp:236-257
{
    Ocean sim{foam_ocean(128)};
    FoamDesc fd;
    fd.decay = 0.01f;
    fd.source_gain = 500.0f;
    FoamField foam{sim, fd};
    for (int i = 0; i < 300; ++i) {
        sim.update(5.0 + i / 60.0);
        foam.update(1.0f / 60.0f);
    }
    const float* f = foam.data();
    for (std::size_t i = 0; i < std::size_t(foam.size()) * foam.size(); ++i)
        if (!std::isfinite(f[i]) || f[i] < 0.0f || f[i] > 1.0f) fail("cell %zu = %g", i, f[i]);
    EXPECT(foam.coverage() > 0.5);
}

std::vector<float> foam_run(ParallelForFn hook, std::uint32_t threads)
{
    Ocean sim{foam_ocean(128)};
    FoamDesc fd;
    fd.parallel_for = hook;
    fd.thread_count = threads;
    FoamField foam{sim, fd};
    for (int i = 0; i < 90; ++i) {
        sim.update(10.0 + i / 60.0);
        foam.update(1.0f / 60.0f);
    }
    return copy_of(foam.data(), std::size_t(foam.size()) * foam.size());
}

void foam_deterministic()  // tests/test_foam.cpp:259-307 (the hooks are never called)
{
    g_hook_calls = 0;
    const std::vector<float> ref = foam_run(nullptr, 1);
    for (std::uint32_t t : {1u, 2u, 4u, 8u, 16u}) EXPECT(same_bits(ref.data(), foam_run(nullptr, t).data(), ref.size()));
    EXPECT(same_bits(ref.data(), foam_run(&reverse_hook, 0).data(), ref.size()));
    EXPECT(same_bits(ref.data(), foam_run(&scattering_hook, 0).data(), ref.size()));
    EXPECT(g_hook_calls.load() == 0);
}

double foam_coverage_at(float dt, int frames)
{
    Ocean sim{foam_ocean(128)};
    FoamField foam{sim, FoamDesc{}};
    for (int i = 0; i < frames; ++i) {
        sim.update(10.0 + double(i + 1) * dt);
        foam.update(dt);
    }
    return foam.coverage();
}

void foam_frame_rate_independent()  // tests/test_foam.cpp:309-333
{
    const double a = foam_coverage_at(1.0f / 60.0f, 120);
    const double b = foam_coverage_at(1.0f / 30.0f, 60);
    info("coverage at 60 fps %.6f, at 30 fps %.6f", a, b);
    EXPECT(approx(a, b, 0.05));
}

void foam_query_and_clear()  // tests/test_foam.cpp:335-393
{
    Ocean sim{foam_ocean(128)};
    sim.update(14.0);
    FoamDesc fd;
    fd.source_gain = 6.0f;
    FoamField foam{sim, fd};
    for (int i = 0; i < 60; ++i) foam.update(1.0f / 60.0f);
    EXPECT(foam.coverage() > 0.0);
    const Buffers b = sim.buffers();
    const float cell = b.patch_length / float(b.size);
    const float* data = foam.data();
    double worst = 0.0, worst_naive = 0.0;
    for (std::uint32_t z = 5; z < 120; z += 9)
        for (std::uint32_t x = 5; x < 120; x += 7) {
            const std::size_t i = std::size_t(z) * b.size + x;
            const float wx = float(x) * cell + b.displacement[4 * i + 0];
            const float wz = float(z) * cell + b.displacement[4 * i + 2];
            worst = std::max(worst, std::fabs(double(foam.foam_at(wx, wz) - data[i])));
            const int nx = int(wx / cell + 0.5f) & 127, nz = int(wz / cell + 0.5f) & 127;
            worst_naive = std::max(worst_naive, std::fabs(double(data[std::size_t(nz) * 128 + nx] - data[i])));
        }
    info("foam_at at displaced positions: worst %.4f (naive lookup %.4f)", worst, worst_naive);
    EXPECT(worst < 0.15);
    EXPECT(worst_naive > worst);
    foam.clear();
    EXPECT(foam.coverage() == 0.0);
}

// ---- tests/test_cascade.cpp --------------------------------------------------

void cascade_empty_rejected()  // tests/test_cascade.cpp:50-54
{
    std::vector<OceanDesc> none;
    bool threw = false;
    try {
        CascadeStack s{std::span<const OceanDesc>(none)};
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    EXPECT(threw);
}

void cascade_invalid_rejected()  // tests/test_cascade.cpp:56-61
{
    std::vector<OceanDesc> levels = three_levels();
    levels[1].size = 100;
    bool threw = false;
    try {
        CascadeStack s{std::span<const OceanDesc>(levels)};
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    EXPECT(threw);
}

void cascade_single_level_equals_ocean()  // tests/test_cascade.cpp:63-99
{
    OceanDesc level = make_desc(64, 200.0f, 555, 1.0f, 12.0f, 0.3f);
    Ocean solo{level};
    CascadeStack stack{level};
    solo.update(4.0);
    stack.update(4.0);
    EXPECT(stack.level_count() == 1);
    const std::size_t n = texels(solo.buffers());
    EXPECT(same_bits(solo.buffers().displacement, stack.buffers(0).displacement, n));
    EXPECT(same_bits(solo.buffers().normal, stack.buffers(0).normal, n));
    for (int i = 0; i < 30; ++i) {
        const float x = -50.0f + 3.3f * float(i), z = 20.0f - 1.7f * float(i);
        const Surface a = solo.sample_at(x, z), b = stack.sample_at(x, z);
        // One term: the sums are that term exactly; foam passes through
        // 1 - (1 - f) and the normal is renormalised once more.
        EXPECT(a.height == b.height && a.offset_x == b.offset_x && a.offset_z == b.offset_z);
        EXPECT(approx(a.foam, b.foam, 1e-5) && approx(a.normal_x, b.normal_x, 1e-5));
        EXPECT(approx(a.normal_y, b.normal_y, 1e-5) && approx(a.normal_z, b.normal_z, 1e-5));
    }
}

void cascade_exact_sum()  // tests/test_cascade.cpp:101-132
{
    const std::vector<OceanDesc> levels = three_levels();
    CascadeStack stack{std::span<const OceanDesc>(levels)};
    stack.update(6.0);
    std::vector<Ocean> solo;
    for (const auto& l : levels) solo.emplace_back(l);
    for (auto& o : solo) o.update(6.0);
    for (int i = 0; i < 25; ++i) {
        const float x = -300.0f + 11.0f * float(i), z = 150.0f - 6.0f * float(i);
        float h = 0.0f, ox = 0.0f, oz = 0.0f;
        for (auto& o : solo) {
            const Surface s = o.sample_at(x, z);
            h += s.height;
            ox += s.offset_x;
            oz += s.offset_z;
        }
        const Surface got = stack.sample_at(x, z);
        // the same float additions in level order: exact
        EXPECT(got.height == h && got.offset_x == ox && got.offset_z == oz);
    }
}

void cascade_unit_normal()  // tests/test_cascade.cpp:134-147
{
    CascadeStack stack{three_levels()[0], three_levels()[1], three_levels()[2]};
    stack.update(9.5);
    for (int i = 0; i < 300; ++i) {
        const Surface s = stack.sample_at(-400.0f + 2.9f * float(i), 200.0f - 1.3f * float(i));
        const float n[3] = {s.normal_x, s.normal_y, s.normal_z};
        if (normal_len2_error(n) >= 1e-4f) fail("query %d", i);
        if (!(s.foam >= 0.0f && s.foam <= 1.0f)) fail("query %d foam %g", i, s.foam);
    }
}

void cascade_variances_add()  // tests/test_cascade.cpp:173-246
{
    const std::vector<OceanDesc> levels = three_levels();
    double sum_var = 0.0;
    for (const auto& l : levels) {
        Ocean o{l};
        o.update(11.0);
        const Buffers b = o.buffers();
        const std::size_t cells = std::size_t(b.size) * b.size;
        double mean = 0.0, var = 0.0;
        for (std::size_t i = 0; i < cells; ++i) mean += b.displacement[4 * i + 1];
        mean /= double(cells);
        for (std::size_t i = 0; i < cells; ++i) var += (b.displacement[4 * i + 1] - mean) * (b.displacement[4 * i + 1] - mean);
        sum_var += var / double(cells);
    }
    CascadeStack stack{std::span<const OceanDesc>(levels)};
    stack.update(11.0);
    const int samples = 6000;
    std::vector<float> xz;
    for (int i = 0; i < samples; ++i) {
        xz.push_back(-900.0f + 0.4173f * float(i));
        xz.push_back(500.0f - 0.2591f * float(i));
    }
    std::vector<Surface> out(samples);
    EXPECT(stack.sample_batch(xz.data(), samples, out.data()));
    double mean = 0.0, var = 0.0;
    for (const Surface& s : out) mean += s.height;
    mean /= samples;
    for (const Surface& s : out) var += (s.height - mean) * (s.height - mean);
    var /= samples;
    info("composite variance %.5f, sum of level variances %.5f", var, sum_var);
    EXPECT(approx(var, sum_var, 0.08));
    EXPECT(stack.height_at(xz[0], xz[1]) == out[0].height);
}

void cascade_deterministic()  // tests/test_cascade.cpp:248-267
{
    const std::vector<OceanDesc> levels = three_levels();
    CascadeStack a{std::span<const OceanDesc>(levels)};
    CascadeStack b{std::span<const OceanDesc>(levels)};
    for (double t : {0.0, 2.5, 88.25}) {
        a.update(t);
        b.update(t);
        for (std::size_t l = 0; l < a.level_count(); ++l) {
            const std::size_t n = texels(a.buffers(l));
            EXPECT(same_bits(a.buffers(l).displacement, b.buffers(l).displacement, n));
            EXPECT(same_bits(a.buffers(l).normal, b.buffers(l).normal, n));
        }
    }
}

void cascade_update_allocation_free()  // tests/test_cascade.cpp:285-300
{
    CascadeStack stack{std::span<const OceanDesc>(three_levels())};
    stack.update(0.0);
    const std::size_t before = library_allocations();
    for (int i = 0; i < 16; ++i) stack.update(double(i) * 0.016);
    EXPECT(library_allocations() == before);
}

void cascade_seek_backwards()  // tests/test_cascade.cpp:269-283
{
    const std::vector<OceanDesc> levels = three_levels();
    CascadeStack stack{std::span<const OceanDesc>(levels)};
    stack.update(15.0);
    const Surface snap = stack.sample_at(12.0f, -7.0f);
    stack.update(500.0);
    stack.update(0.1);
    stack.update(15.0);
    EXPECT(same_surface(snap, stack.sample_at(12.0f, -7.0f)));
}

// ---- tests/test_c_api.cpp, tests/test_threading.cpp ----------------------------

void capi_c_equals_cpp()  // tests/test_c_api.cpp:51-95
{
    Ocean cpp{cpp_desc(64)};
    ocean_desc cd = c_desc(64);
    ocean_status st = OCEAN_ERROR_UNKNOWN;
    ocean_sim* c = ocean_create(&cd, &st);
    EXPECT(c != nullptr && st == OCEAN_OK);
    for (double t : {0.0, 1.25, 60.5}) {
        cpp.update(t);
        ocean_update(c, t);
        const Buffers a = cpp.buffers();
        const ocean_buffers b = ocean_get_buffers(c);
        EXPECT(b.size == a.size && b.patch_length == a.patch_length);
        EXPECT(same_bits(a.displacement, b.displacement, texels(a)));
        EXPECT(same_bits(a.normal, b.normal, texels(a)));
        EXPECT(ocean_get_time(c) == t);
        for (int i = 0; i < 20; ++i) {
            const float x = -90.0f + 7.3f * float(i), z = 45.0f - 3.1f * float(i);
            const Surface s = cpp.sample_at(x, z);
            const ocean_surface cs = ocean_sample_at(c, x, z);
            EXPECT(s.height == cs.height && s.normal_y == cs.normal_y && s.foam == cs.foam);
            ocean_surface_ex ex;
            ocean_sample_ex(c, x, z, &ex);
            EXPECT(std::memcmp(&ex, &s, sizeof(Surface)) == 0);
            EXPECT(ocean_height_at(c, x, z) == cpp.height_at(x, z));
        }
    }
    ocean_destroy(c);
}

void capi_invalid_desc()  // tests/test_c_api.cpp:97-135
{
    auto refused = [](ocean_desc d) {
        ocean_status s = OCEAN_OK;
        ocean_sim* sim = ocean_create(&d, &s);
        if (sim != nullptr) ocean_destroy(sim);
        return sim == nullptr && s == OCEAN_ERROR_INVALID_ARG;
    };
    EXPECT(refused(c_desc(100)));
    ocean_desc d = c_desc(64);
    d.patch_length = 0.0f;
    EXPECT(refused(d));
    d = c_desc(64);
    d.struct_size = 0;
    EXPECT(refused(d));
    d = c_desc(64);
    d.struct_size = sizeof(ocean_desc) + 64;
    EXPECT(refused(d));
    ocean_status s = OCEAN_OK;
    EXPECT(ocean_create(nullptr, &s) == nullptr && s == OCEAN_ERROR_INVALID_ARG);
    d = c_desc(100);
    EXPECT(ocean_create(&d, nullptr) == nullptr);
}

// ocean_desc as it stood before the scheduler fields (tests/test_c_api.cpp:146-156)
// and before water_depth (tests/test_c_api.cpp:295-313): callers built
// against older headers.
struct DescBeforeThreads {
    std::size_t struct_size;
    std::uint32_t size;
    float patch_length;
    ocean_spectrum_desc spectrum;
    float choppiness;
    float foam_threshold;
    std::uint64_t seed;
};

struct DescBeforeDepth {
    std::size_t struct_size;
    std::uint32_t size;
    float patch_length;
    ocean_spectrum_desc spectrum;
    float choppiness;
    float foam_threshold;
    std::uint64_t seed;
    ocean_parallel_for_fn parallel_for;
    void* parallel_for_user;
    std::uint32_t thread_count;
};

template <class Legacy>
void legacy_caller_case()
{
    static_assert(sizeof(Legacy) <= sizeof(ocean_desc));
    Legacy legacy{};
    legacy.struct_size = sizeof(Legacy);
    legacy.size = 64;
    legacy.patch_length = 200.0f;
    legacy.spectrum.wind_speed = 11.0f;
    legacy.spectrum.fetch = 100000.0f;
    legacy.spectrum.wind_direction = 0.25f;
    legacy.spectrum.peak_enhancement = 3.3f;
    legacy.spectrum.swell = 0.0f;
    legacy.spectrum.small_wave_cutoff = 0.5f;
    legacy.spectrum.gravity = 9.81f;
    legacy.choppiness = 1.0f;
    legacy.foam_threshold = 0.5f;
    legacy.seed = 555777;
    ocean_status st = OCEAN_ERROR_UNKNOWN;
    ocean_sim* sim = ocean_create(reinterpret_cast<const ocean_desc*>(&legacy), &st);
    EXPECT(sim != nullptr && st == OCEAN_OK);
    ocean_update(sim, 1.25);
    Ocean ref{cpp_desc(64)};
    ref.update(1.25);
    EXPECT(same_bits(ref.buffers().displacement, ocean_get_buffers(sim).displacement, texels(ref.buffers())));
    ocean_destroy(sim);
}

void capi_legacy_struct_size()  // tests/test_c_api.cpp:160-203
{
    legacy_caller_case<DescBeforeThreads>();
}

void capi_legacy_before_depth()  // tests/test_c_api.cpp:316-369
{
    legacy_caller_case<DescBeforeDepth>();
}

void capi_water_depth()  // tests/test_c_api.cpp:252-314
{
    ocean_desc d = c_desc(32);
    d.water_depth = 6.5f;
    ocean_sim* sim = ocean_create(&d, nullptr);
    EXPECT(sim != nullptr);
    ocean_update(sim, 3.0);
    OceanDesc cd = cpp_desc(32);
    cd.spectrum.depth = 6.5f;
    Ocean ref{cd};
    ref.update(3.0);
    EXPECT(same_bits(ref.buffers().displacement, ocean_get_buffers(sim).displacement, texels(ref.buffers())));
    Ocean deep{cpp_desc(32)};
    deep.update(3.0);
    EXPECT(!same_bits(deep.buffers().displacement, ref.buffers().displacement, texels(ref.buffers())));
    ocean_destroy(sim);
}

void capi_hook_never_called()  // tests/test_c_api.cpp:371-394, tests/test_threading.cpp:139-199
{
    g_hook_calls = 0;
    Ocean builtin{cpp_desc(64)};
    builtin.update(1.0);
    for (ocean_parallel_for_fn hook : {reinterpret_cast<ocean_parallel_for_fn>(&counting_hook),
                                       reinterpret_cast<ocean_parallel_for_fn>(&reverse_hook),
                                       reinterpret_cast<ocean_parallel_for_fn>(&scattering_hook)}) {
        ocean_desc d = c_desc(64);
        d.parallel_for = hook;
        ocean_sim* sim = ocean_create(&d, nullptr);
        EXPECT(sim != nullptr);
        ocean_update(sim, 1.0);
        EXPECT(same_bits(builtin.buffers().displacement, ocean_get_buffers(sim).displacement,
                         texels(builtin.buffers())));
        ocean_destroy(sim);
        OceanDesc cd = cpp_desc(64);
        cd.parallel_for = reinterpret_cast<ParallelForFn>(hook);
        Ocean cpp{cd};
        cpp.update(1.0);
        EXPECT(same_bits(builtin.buffers().displacement, cpp.buffers().displacement, texels(builtin.buffers())));
    }
    EXPECT(g_hook_calls.load() == 0);
}

void capi_status_version_simd()  // tests/test_c_api.cpp:396-458
{
    EXPECT(std::strcmp(ocean_status_string(OCEAN_OK), "ok") == 0);
    EXPECT(std::strlen(ocean_status_string(OCEAN_ERROR_INVALID_ARG)) > 0);
    std::uint32_t major = 99, minor = 99, patch = 99;
    ocean_version(&major, &minor, &patch);
    EXPECT(major == 1 && minor == 0 && patch == 0);
    ocean_version(nullptr, nullptr, nullptr);
    const std::string level = ocean_simd_level();
    EXPECT(!level.empty());
    // one device path: forcing a CPU level is a no-op that still returns a
    // valid name, and a simulation built meanwhile gives the same bits
    EXPECT(ocean_force_simd_level("scalar") != nullptr);
    EXPECT(level == ocean_simd_level());
    ocean_desc d = c_desc(32);
    d.seed = 555;
    ocean_sim* forced = ocean_create(&d, nullptr);
    ocean_force_simd_level(nullptr);
    ocean_sim* native = ocean_create(&d, nullptr);
    EXPECT(forced != nullptr && native != nullptr);
    ocean_update(forced, 4.0);
    ocean_update(native, 4.0);
    const ocean_buffers a = ocean_get_buffers(forced);
    const std::vector<float> da = copy_of(a.displacement, 32u * 32u * 4u);
    const std::vector<float> na = copy_of(a.normal, 32u * 32u * 4u);
    const ocean_buffers b = ocean_get_buffers(native);
    EXPECT(same_bits(da.data(), b.displacement, da.size()) && same_bits(na.data(), b.normal, na.size()));
    ocean_destroy(forced);
    ocean_destroy(native);
}

void threads_bitidentical_thread_count()  // tests/test_threading.cpp:35-61
{
    OceanDesc d = make_desc(128, 200.0f, 90210, 1.0f, 13.0f, 0.7f);
    d.thread_count = 1;
    Ocean ref{d};
    ref.update(6.25);
    for (std::uint32_t threads : {1u, 2u, 3u, 4u, 8u, 16u, 32u}) {
        d.thread_count = threads;
        Ocean sim{d};
        sim.update(6.25);
        EXPECT(same_bits(ref.buffers().displacement, sim.buffers().displacement, texels(ref.buffers())));
        EXPECT(same_bits(ref.buffers().normal, sim.buffers().normal, texels(ref.buffers())));
    }
}

void threads_repeated_stable()  // tests/test_threading.cpp:63-87
{
    OceanDesc d = make_desc(128, 200.0f, 90210, 1.0f, 13.0f, 0.7f);
    d.thread_count = 16;
    Ocean sim{d};
    sim.update(3.5);
    const std::vector<float> first = copy_of(sim.buffers().displacement, texels(sim.buffers()));
    for (int i = 0; i < 30; ++i) {
        sim.update(100.0 + i);
        sim.update(3.5);
        if (!same_bits(first.data(), sim.buffers().displacement, first.size())) fail("round %d", i);
    }
}

// No reference counterpart as a test: the reference's queries are const and
// may run concurrently on one handle. Here they share a handle's staging
// (and its stream), so they take a lock; four threads querying one ocean and
// one cascade at once, single points and batches, get the single-threaded
// bits.
void threads_concurrent_queries()
{
    Ocean sim{make_desc(128, 200.0f, 4242, 1.0f, 13.0f, 0.9f)};
    sim.update(2.25);
    std::vector<OceanDesc> levels{make_desc(64, 300.0f, 7, 1.0f, 12.0f, 0.8f),
                                  make_desc(64, 60.0f, 8, 1.0f, 12.0f, 0.8f)};
    CascadeStack stack{std::span<const OceanDesc>(levels)};
    stack.update(2.25);
    constexpr int kPoints = 600;
    std::vector<float> xz(2 * kPoints);
    for (int i = 0; i < kPoints; ++i) {
        xz[2 * i] = 0.37f * float(i) - 40.0f;
        xz[2 * i + 1] = 25.0f - 0.21f * float(i);
    }
    std::vector<float> want_h(kPoints), want_c(kPoints);
    for (int i = 0; i < kPoints; ++i) {
        want_h[i] = sim.height_at(xz[2 * i], xz[2 * i + 1]);
        want_c[i] = stack.sample_at(xz[2 * i], xz[2 * i + 1]).height;
    }
    std::vector<Surface> want_b(kPoints);
    EXPECT(sim.sample_batch(xz.data(), kPoints, want_b.data()));
    std::atomic<int> bad{0};
    auto worker = [&](int t) {
        std::vector<Surface> got(kPoints);
        for (int rep = 0; rep < 4; ++rep) {
            for (int i = t; i < kPoints; i += 3) {
                const float h = sim.height_at(xz[2 * i], xz[2 * i + 1]);
                const float c = stack.sample_at(xz[2 * i], xz[2 * i + 1]).height;
                if (!same_bits(&h, &want_h[i], 1) || !same_bits(&c, &want_c[i], 1)) ++bad;
            }
            if (!sim.sample_batch(xz.data(), kPoints, got.data()) ||
                std::memcmp(got.data(), want_b.data(), sizeof(Surface) * kPoints) != 0)
                ++bad;
        }
    };
    std::vector<std::thread> threads;
    for (int t = 0; t < 4; ++t) threads.emplace_back(worker, t);
    for (std::thread& th : threads) th.join();
    EXPECT(bad.load() == 0);
}

// The other const queries on shared handles, four threads at once against
// the single-threaded bits: FoamField::foam_at / coverage, InteractionField
// and WaterSurface sample_at, and OceanBatch::sample over one level and over
// a cascade of two. Every one of them shares its handle's staging and stream.
void threads_concurrent_other_queries()
{
    OceanDesc od = make_desc(64, 80.0f, 4243, 1.2f, 18.0f, 0.9f);
    od.compute_velocity = true;
    Ocean sim{od};
    sim.update(3.0);
    FoamDesc fd;
    fd.source_gain = 6.0f;
    FoamField foam{sim, fd};
    for (int i = 0; i < 20; ++i) foam.update(1.0f / 60.0f);
    InteractionDesc idc;
    idc.size = 64;
    idc.extent = 12.8f;
    InteractionField field{idc};
    Disturbance dist;
    dist.world_x = 6.0f;
    dist.world_z = 6.5f;
    dist.radius = 0.8f;
    dist.strength = 0.2f;
    dist.velocity_y = -1.0f;
    for (int i = 0; i < 12; ++i) {
        field.add(dist);
        field.update(1.0f / 60.0f);
    }
    WaterSurface surface{sim, field};
    std::vector<OceanDesc> levels{make_desc(64, 300.0f, 7, 1.0f, 12.0f, 0.8f),
                                  make_desc(64, 60.0f, 8, 1.0f, 12.0f, 0.8f)};
    OceanBatch batch{std::span<const OceanDesc>(levels)};
    batch.update(2.5);
    constexpr int kPoints = 300;
    std::vector<float> xz(2 * kPoints);
    for (int i = 0; i < kPoints; ++i) {
        xz[2 * i] = 0.053f * float(i) - 1.0f;
        xz[2 * i + 1] = 13.0f - 0.041f * float(i);
    }
    std::vector<float> want_f(kPoints);
    std::vector<InteractionSample> want_i(kPoints);
    std::vector<Surface> want_w(kPoints), want_b1(kPoints), want_b2(kPoints);
    for (int i = 0; i < kPoints; ++i) {
        want_f[i] = foam.foam_at(xz[2 * i], xz[2 * i + 1]);
        want_i[i] = field.sample_at(xz[2 * i], xz[2 * i + 1]);
        want_w[i] = surface.sample_at(xz[2 * i], xz[2 * i + 1]);
    }
    const double want_cov = foam.coverage();
    EXPECT(want_cov > 0.0);
    EXPECT(batch.sample(0, 1, xz.data(), kPoints, want_b1.data()));
    EXPECT(batch.sample(0, 2, xz.data(), kPoints, want_b2.data()));
    std::atomic<int> bad{0};
    auto worker = [&](int t) {
        std::vector<Surface> got(kPoints);
        for (int rep = 0; rep < 3; ++rep) {
            for (int i = t; i < kPoints; i += 5) {
                const float f = foam.foam_at(xz[2 * i], xz[2 * i + 1]);
                const InteractionSample e = field.sample_at(xz[2 * i], xz[2 * i + 1]);
                const Surface w = surface.sample_at(xz[2 * i], xz[2 * i + 1]);
                if (!same_bits(&f, &want_f[i], 1) ||
                    std::memcmp(&e, &want_i[i], sizeof e) != 0 ||
                    std::memcmp(&w, &want_w[i], sizeof w) != 0)
                    ++bad;
            }
            const double c = foam.coverage();
            if (std::memcmp(&c, &want_cov, sizeof c) != 0) ++bad;
            if (!batch.sample(0, 1, xz.data(), kPoints, got.data()) ||
                std::memcmp(got.data(), want_b1.data(), sizeof(Surface) * kPoints) != 0)
                ++bad;
            if (!batch.sample(0, 2, xz.data(), kPoints, got.data()) ||
                std::memcmp(got.data(), want_b2.data(), sizeof(Surface) * kPoints) != 0)
                ++bad;
        }
    };
    std::vector<std::thread> threads;
    for (int t = 0; t < 4; ++t) threads.emplace_back(worker, t);
    for (std::thread& th : threads) th.join();
    EXPECT(bad.load() == 0);
    EXPECT(sim.last_error() == 0 && field.last_error() == 0);
}

// A FoamField on a cascade's level and on a batch's level (the reference
// builds foam fields on stack.level(i): include/ocean/cascade.hpp:72-79).
// Those levels hold no host query staging of their own; foam_at, coverage
// and the field must equal a FoamField on a plain Ocean of the same desc.
void foam_on_cascade_and_batch_levels()
{
    OceanDesc a = make_desc(64, 120.0f, 31, 1.3f, 20.0f, 0.9f);
    OceanDesc b = make_desc(128, 30.0f, 32, 1.1f, 20.0f, 0.9f);
    a.compute_velocity = b.compute_velocity = true;
    std::vector<OceanDesc> levels{a, b};
    CascadeStack stack{std::span<const OceanDesc>(levels)};
    OceanBatch batch{std::span<const OceanDesc>(levels)};
    Ocean plain_a{a}, plain_b{b};
    FoamDesc fd;
    fd.source_gain = 5.0f;
    FoamField on_stack{stack.level(0), fd}, on_batch{batch.level(1), fd};
    FoamField ref_a{plain_a, fd}, ref_b{plain_b, fd};
    for (int k = 0; k < 30; ++k) {
        const double t = 1.0 + k / 60.0;
        stack.update(t);
        batch.update(t);
        plain_a.update(t);
        plain_b.update(t);
        on_stack.update(1.0f / 60.0f);
        on_batch.update(1.0f / 60.0f);
        ref_a.update(1.0f / 60.0f);
        ref_b.update(1.0f / 60.0f);
    }
    const std::size_t ca = std::size_t(a.size) * a.size, cb = std::size_t(b.size) * b.size;
    EXPECT(same_bits(on_stack.data(), ref_a.data(), ca));
    EXPECT(same_bits(on_batch.data(), ref_b.data(), cb));
    const double cs = on_stack.coverage(), cra = ref_a.coverage();
    const double cbt = on_batch.coverage(), crb = ref_b.coverage();
    EXPECT(cs > 0.0 && std::memcmp(&cs, &cra, sizeof cs) == 0);
    EXPECT(cbt > 0.0 && std::memcmp(&cbt, &crb, sizeof cbt) == 0);
    for (int i = 0; i < 200; ++i) {
        const float x = 0.91f * float(i) - 40.0f, z = 70.0f - 0.63f * float(i);
        const float fs = on_stack.foam_at(x, z), fa = ref_a.foam_at(x, z);
        const float fb = on_batch.foam_at(x, z), fr = ref_b.foam_at(x, z);
        if (!same_bits(&fs, &fa, 1) || !same_bits(&fb, &fr, 1)) fail("foam_at differs at point %d", i);
    }
    EXPECT(stack.level(0).last_error() == 0 && batch.level(1).last_error() == 0);
}

}  // namespace

namespace cases {

const Case kFrameCases[] = {
    {"api.buffers_aligned", api_buffers_aligned},
    {"api.constructed_at_t0", api_constructed_at_t0},
    {"api.zero_wind_flat", api_zero_wind_flat},
    {"api.absolute_time", api_absolute_time},
    {"api.movable_desc_roundtrip", api_movable_desc_roundtrip},
    {"api.allocation_probe_installed", api_allocation_probe_installed},
    {"api.update_allocation_free", api_update_allocation_free},
    {"api.bad_desc_throws", api_bad_desc_throws},
    {"query.agrees_with_drawn_grid", query_agrees_with_drawn_grid},
    {"query.chop0_exact_lookup", query_chop0_exact_lookup},
    {"query.sample_consistent_with_buffers", query_sample_consistent_with_buffers},
    {"query.unit_normals", query_unit_normals},
    {"query.periodic", query_periodic},
    {"query.far_outside_finite", query_far_outside_finite},
    {"query.extreme_chop", query_extreme_chop},
    {"query.follows_surface", query_follows_surface},
    {"query.flat_ocean", query_flat_ocean},
    {"sim.bitidentical_instances", sim_bitidentical_instances},
    {"sim.seek_backwards", sim_seek_backwards},
    {"sim.different_seeds_same_statistics", sim_different_seeds_same_statistics},
    {"sim.parseval", sim_parseval},
    {"sim.hs_plausible", sim_hs_plausible},
    {"sim.stronger_wind", sim_stronger_wind},
    {"sim.surface_moves", sim_surface_moves},
    {"sim.finite_unit_normals", sim_finite_unit_normals},
    {"sim.chop0_flat_jacobian", sim_chop0_flat_jacobian},
    {"sim.chop_only_horizontal", sim_chop_only_horizontal},
    {"sim.foam_ramp", sim_foam_ramp},
    {"velocity.absent_unless_asked", vel_absent_unless_asked},
    {"velocity.no_bit_change", vel_no_bit_change},
    {"velocity.uy_is_dhdt", vel_uy_is_dhdt},
    {"velocity.circular_orbits", vel_circular_orbits},
    {"velocity.query_agrees_with_buffer", vel_query_agrees_with_buffer},
    {"velocity.deterministic_seek", vel_deterministic_seek},
    {"velocity.cascade_sums", vel_cascade_sums},
    {"foam.update_allocation_free", foam_update_allocation_free},
    {"foam.refuses", foam_refuses},
    {"foam.decay_rate", foam_decay_rate},
    {"foam.exact_advection", foam_exact_advection},
    {"foam.wraps", foam_wraps},
    {"foam.outlives_jacobian", foam_outlives_jacobian},
    {"foam.saturating_clamp", foam_saturating_clamp},
    {"foam.deterministic", foam_deterministic},
    {"foam.frame_rate_independent", foam_frame_rate_independent},
    {"foam.query_and_clear", foam_query_and_clear},
    {"cascade.empty_rejected", cascade_empty_rejected},
    {"cascade.invalid_rejected", cascade_invalid_rejected},
    {"cascade.single_level_equals_ocean", cascade_single_level_equals_ocean},
    {"cascade.exact_sum", cascade_exact_sum},
    {"cascade.unit_normal", cascade_unit_normal},
    {"cascade.variances_add", cascade_variances_add},
    {"cascade.deterministic", cascade_deterministic},
    {"cascade.update_allocation_free", cascade_update_allocation_free},
    {"cascade.seek_backwards", cascade_seek_backwards},
    {"capi.c_equals_cpp", capi_c_equals_cpp},
    {"capi.invalid_desc", capi_invalid_desc},
    {"capi.legacy_struct_size", capi_legacy_struct_size},
    {"capi.legacy_before_depth", capi_legacy_before_depth},
    {"capi.water_depth", capi_water_depth},
    {"capi.hook_never_called", capi_hook_never_called},
    {"capi.status_version_simd", capi_status_version_simd},
    {"threads.bitidentical_thread_count", threads_bitidentical_thread_count},
    {"threads.repeated_stable", threads_repeated_stable},
    {"threads.concurrent_queries", threads_concurrent_queries},
    {"threads.concurrent_other_queries", threads_concurrent_other_queries},
    {"foam.on_cascade_and_batch_levels", foam_on_cascade_and_batch_levels},
};
const std::size_t kFrameCaseCount = sizeof(kFrameCases) / sizeof(kFrameCases[0]);

}  // namespace cases

int main(int argc, char** argv)
{
    const std::string prefix = argc > 1 ? argv[1] : "";
    const std::pair<const Case*, std::size_t> tables[] = {{kFrameCases, kFrameCaseCount},
                                                          {kInteractionCases, kInteractionCaseCount}};
    int failed = 0;
    for (const auto& [table, count] : tables) {
        for (std::size_t k = 0; k < count; ++k) {
            const Case& c = table[k];
            if (prefix == "--list") {
                std::printf("%s\n", c.name);
                continue;
            }
            if (std::string(c.name).rfind(prefix, 0) != 0) continue;
            g_case = c.name;
            std::printf("RUN %s\n", c.name);
            std::fflush(stdout);
            try {
                c.run();
                std::printf("PASS %s\n", c.name);
            } catch (const CaseFailed& f) {
                std::printf("FAIL %s: %s\n", c.name, f.why.c_str());
                ++failed;
            } catch (const std::exception& e) {
                std::printf("FAIL %s: exception %s\n", c.name, e.what());
                ++failed;
            }
            std::fflush(stdout);
        }
    }
    return failed == 0 ? 0 : 1;
}

// Fixture hooks into the reference library, compiled together with the
// reference's own, unmodified sources into oracle/_ref/liboceanref.so.
//
// Test infrastructure only. Every symbol here is prefixed `oref_` and is the
// only thing the .so exports (the reference objects are built with hidden
// visibility), so it can sit in one process beside the shipped library, which
// exports the real `ocean_*` names.
//
// The hooks expose internals the reference's public API hides (spectrum
// tables, twiddles, single FFTs, the evolve/finalize stages) so that the C
// restatement in ocean_oracle.c can be compared against the reference stage
// by stage, not only end to end.

#include "core/evolve.hpp"
#include "core/fft.hpp"
#include "core/iwave_kernel.hpp"
#include "core/iwave_step.hpp"
#include "core/rng.hpp"
#include "core/spectrum.hpp"
#include "core/trig.hpp"
#include "ocean/cascade.hpp"
#include "ocean/foam.hpp"
#include "ocean/interaction.hpp"
#include "ocean/ocean.h"
#include "ocean/ocean.hpp"

#include <algorithm>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

#define OREF extern "C" __attribute__((visibility("default")))

using namespace ocean;
using namespace ocean::detail;

namespace {

// Same field-by-field mapping as the reference's own c_api.cpp translate().
OceanDesc to_desc(const ocean_desc* c)
{
    OceanDesc d;
    d.size = c->size;
    d.patch_length = c->patch_length;
    d.spectrum.wind_speed = c->spectrum.wind_speed;
    d.spectrum.fetch = c->spectrum.fetch;
    d.spectrum.wind_direction = c->spectrum.wind_direction;
    d.spectrum.peak_enhancement = c->spectrum.peak_enhancement;
    d.spectrum.swell = c->spectrum.swell;
    d.spectrum.small_wave_cutoff = c->spectrum.small_wave_cutoff;
    d.spectrum.gravity = c->spectrum.gravity;
    d.spectrum.depth = c->water_depth;
    d.choppiness = c->choppiness;
    d.foam_threshold = c->foam_threshold;
    d.seed = c->seed;
    d.thread_count = c->thread_count;
    d.compute_velocity = c->compute_velocity != 0;
    return d;
}

}  // namespace

// --- C API passthrough ------------------------------------------------------

OREF ocean_sim* oref_create(const ocean_desc* d, ocean_status* st) { return ocean_create(d, st); }
OREF void oref_destroy(ocean_sim* s) { ocean_destroy(s); }
OREF void oref_update(ocean_sim* s, double t) { ocean_update(s, t); }
OREF ocean_buffers oref_get_buffers(const ocean_sim* s) { return ocean_get_buffers(s); }
OREF const float* oref_get_velocity(const ocean_sim* s) { return ocean_get_velocity(s); }
OREF void oref_sample_ex(const ocean_sim* s, float x, float z, ocean_surface_ex* o) { ocean_sample_ex(s, x, z, o); }
OREF void oref_desc_init(ocean_desc* d) { ocean_desc_init(d); }

// --- primitives -----------------------------------------------------------

OREF void oref_pcg32_draws(std::uint64_t seed, std::uint64_t stream, std::uint32_t count,
                           std::uint32_t* out)
{
    Pcg32 r = seed_pcg32(seed, stream);
    for (std::uint32_t i = 0; i < count; ++i) out[i] = next_u32(r);
}

OREF void oref_row_gaussians(std::uint64_t seed, std::uint32_t row, std::uint32_t pairs,
                             float* out)
{
    Pcg32 r = row_rng(seed, row);
    for (std::uint32_t i = 0; i < pairs; ++i) {
        const GaussianPair g = next_gaussian_pair(r);
        out[2 * i] = g.a;
        out[2 * i + 1] = g.b;
    }
}

OREF void oref_sincos(const float* x, std::uint32_t count, float* s, float* c)
{
    for (std::uint32_t i = 0; i < count; ++i) sincos_f32(x[i], s[i], c[i]);
}

OREF void oref_reduced_phase(const float* omega, std::uint32_t count, double time, float* out)
{
    const double t2 = time * kInvTwoPiD;
    for (std::uint32_t i = 0; i < count; ++i) out[i] = reduced_phase(omega[i], t2);
}

OREF void oref_spectrum(const ocean_desc* c, float* h0re, float* h0im, float* h0cre,
                        float* h0cim, float* omega, float* kx, float* kz, float* kinv)
{
    SpectrumTables t;
    build_spectrum(to_desc(c), t);
    const std::size_t cells = static_cast<std::size_t>(c->size) * c->size;
    std::memcpy(h0re, t.h0_re.data(), cells * 4);
    std::memcpy(h0im, t.h0_im.data(), cells * 4);
    std::memcpy(h0cre, t.h0c_re.data(), cells * 4);
    std::memcpy(h0cim, t.h0c_im.data(), cells * 4);
    std::memcpy(omega, t.omega.data(), cells * 4);
    std::memcpy(kx, t.kx.data(), cells * 4);
    std::memcpy(kz, t.kz.data(), cells * 4);
    std::memcpy(kinv, t.k_inv.data(), cells * 4);
}

OREF double oref_wave_density(const ocean_desc* c, double kx, double kz)
{
    return wave_density(to_desc(c).spectrum, kx, kz);
}

// Inverse 1-D transform of one contiguous sequence, through the scalar stage
// kernel (all SIMD levels are bit-identical to it by the reference's contract).
OREF void oref_fft_inverse(std::uint32_t n, float* re, float* im)
{
    FftPlan plan(n, SimdLevel::Scalar);
    std::vector<float> scratch(plan.scratch_floats());
    plan.transform(re, im, 1, scratch.data(), FftSign::Inverse);
}

OREF void oref_fft2d_inverse(std::uint32_t n, float* re, float* im)
{
    FftPlan plan(n);
    std::vector<float> scratch(plan.scratch_floats());
    plan.transform_2d(re, im, scratch.data(), FftSign::Inverse);
}

// Evolve for all rows into `planes` (8 or 12 planes of n*n, split re/im per
// complex field, the reference's FieldSet layout).
OREF void oref_evolve(const ocean_desc* c, double time, float* planes)
{
    const OceanDesc d = to_desc(c);
    SpectrumTables t;
    build_spectrum(d, t);
    FieldSet f;
    f.allocate(d.size, d.compute_velocity);
    evolve_rows(t, time, f, 0, d.size);
    if (d.compute_velocity) evolve_velocity_rows(t, time, f, 0, d.size);
    std::memcpy(planes, f.storage.data(), f.storage.size() * 4);
}

// --- cascades ---------------------------------------------------------------

OREF void* oref_cascade_create(const ocean_desc* levels, std::uint32_t count)
{
    try {
        std::vector<CascadeLevel> v;
        for (std::uint32_t i = 0; i < count; ++i) v.push_back(to_desc(&levels[i]));
        return new CascadeStack(std::span<const CascadeLevel>(v.data(), v.size()));
    } catch (...) {
        return nullptr;
    }
}
OREF void oref_cascade_destroy(void* s) { delete static_cast<CascadeStack*>(s); }
OREF void oref_cascade_update(void* s, double t) { static_cast<CascadeStack*>(s)->update(t); }
OREF const float* oref_cascade_displacement(void* s, std::uint32_t level)
{
    return static_cast<CascadeStack*>(s)->buffers(level).displacement;
}
OREF const float* oref_cascade_normal(void* s, std::uint32_t level)
{
    return static_cast<CascadeStack*>(s)->buffers(level).normal;
}
OREF void oref_cascade_sample(void* s, float x, float z, ocean_surface_ex* o)
{
    const Surface r = static_cast<CascadeStack*>(s)->sample_at(x, z);
    o->height = r.height; o->offset_x = r.offset_x; o->offset_z = r.offset_z;
    o->normal_x = r.normal_x; o->normal_y = r.normal_y; o->normal_z = r.normal_z;
    o->foam = r.foam; o->velocity_x = r.velocity_x; o->velocity_y = r.velocity_y;
    o->velocity_z = r.velocity_z;
}

// --- persistent foam --------------------------------------------------------

struct FoamParams {
    float decay, source_gain, advect_scale, wind_drift_x, wind_drift_z, fixed_dt;
    std::uint32_t max_substeps;
};

OREF void* oref_foam_create(ocean_sim* sim, const FoamParams* p)
{
    try {
        FoamDesc d;
        d.decay = p->decay; d.source_gain = p->source_gain; d.advect_scale = p->advect_scale;
        d.wind_drift_x = p->wind_drift_x; d.wind_drift_z = p->wind_drift_z;
        d.fixed_dt = p->fixed_dt; d.max_substeps = p->max_substeps;
        return new FoamField(*reinterpret_cast<const Ocean*>(sim), d);
    } catch (...) {
        return nullptr;
    }
}
OREF void oref_foam_destroy(void* f) { delete static_cast<FoamField*>(f); }
OREF void oref_foam_update(void* f, float dt) { static_cast<FoamField*>(f)->update(dt); }
OREF const float* oref_foam_data(void* f) { return static_cast<FoamField*>(f)->data(); }
OREF std::uint32_t oref_foam_last_substeps(void* f) { return static_cast<FoamField*>(f)->last_substeps(); }
OREF float oref_foam_at(void* f, float x, float z) { return static_cast<FoamField*>(f)->foam_at(x, z); }
OREF double oref_foam_coverage(void* f) { return static_cast<FoamField*>(f)->coverage(); }

// --- interaction (iWave) ----------------------------------------------------

OREF void oref_interaction_desc_init(ocean_interaction_desc* d) { ocean_interaction_desc_init(d); }
OREF ocean_interaction* oref_interaction_create(const ocean_interaction_desc* d, ocean_status* st)
{
    return ocean_interaction_create(d, st);
}
OREF void oref_interaction_destroy(ocean_interaction* f) { ocean_interaction_destroy(f); }
OREF void oref_interaction_add(ocean_interaction* f, const ocean_disturbance* d) { ocean_interaction_add(f, d); }
OREF void oref_interaction_update(ocean_interaction* f, float dt) { ocean_interaction_update(f, dt); }
OREF std::int32_t oref_interaction_recenter(ocean_interaction* f, float x, float z)
{
    return ocean_interaction_recenter(f, x, z);
}
OREF void oref_interaction_set_obstruction(ocean_interaction* f, const std::uint8_t* m)
{
    ocean_interaction_set_obstruction(f, m);
}
OREF ocean_interaction_buffers oref_interaction_get_buffers(const ocean_interaction* f)
{
    return ocean_interaction_get_buffers(f);
}
OREF void oref_interaction_sample_at(const ocean_interaction* f, float x, float z,
                                     ocean_interaction_sample* o)
{
    ocean_interaction_sample_at(f, x, z, o);
}
OREF float oref_interaction_fixed_dt(const ocean_interaction* f) { return ocean_interaction_fixed_dt(f); }
OREF float oref_interaction_dt_limit(const ocean_interaction* f) { return ocean_interaction_dt_limit(f); }
OREF std::uint64_t oref_interaction_dropped(const ocean_interaction* f) { return ocean_interaction_dropped(f); }
// The C handle is the InteractionField itself (reference src/c_api.cpp:289-296).
OREF double oref_interaction_energy(const ocean_interaction* f)
{
    return reinterpret_cast<const InteractionField*>(f)->energy();
}
OREF void oref_sample_combined(const ocean_sim* s, const ocean_interaction* f, float x, float z,
                               ocean_surface_ex* o)
{
    ocean_sample_combined(s, f, x, z, o);
}

// iWave convolution taps as the reference builds them (LeastSquares, the
// solver's default method). Writes (2P+1)^2 floats, row-major j-outer.
OREF int oref_iwave_taps(std::uint32_t radius, double cell, double depth, float* taps)
{
    try {
        IWaveKernel k;
        build_iwave_kernel(radius, cell, depth, KernelMethod::LeastSquares, k);
        const std::size_t w = 2u * radius + 1u;
        for (std::size_t i = 0; i < w * w; ++i) taps[i] = k.taps[i];
        return static_cast<int>(w * w);
    } catch (...) {
        return -1;
    }
}

// The same fit with its scan summary: sym = max_symbol, min_symbol, band_low,
// peak_relative_error.
OREF int oref_iwave_stencil(std::uint32_t radius, double cell, double depth, float* taps,
                            double* sym)
{
    try {
        IWaveKernel k;
        build_iwave_kernel(radius, cell, depth, KernelMethod::LeastSquares, k);
        const std::size_t w = 2u * radius + 1u;
        for (std::size_t i = 0; i < w * w; ++i) taps[i] = k.taps[i];
        sym[0] = k.max_symbol;
        sym[1] = k.min_symbol;
        sym[2] = k.band_low;
        sym[3] = k.peak_relative_error;
        return static_cast<int>(w * w);
    } catch (...) {
        return -1;
    }
}

// A caller's loop of sample_at over an array of points, kept in C++ so the
// CPU baseline times the reference's query, not the FFI call.
OREF void oref_sample_batch(const ocean_sim* s, const float* xz, std::uint32_t count,
                            ocean_surface_ex* out)
{
    const Ocean& o = *reinterpret_cast<const Ocean*>(s);
    for (std::uint32_t i = 0; i < count; ++i) {
        const Surface r = o.sample_at(xz[2 * i], xz[2 * i + 1]);
        ocean_surface_ex& e = out[i];
        e.height = r.height; e.offset_x = r.offset_x; e.offset_z = r.offset_z;
        e.normal_x = r.normal_x; e.normal_y = r.normal_y; e.normal_z = r.normal_z;
        e.foam = r.foam; e.velocity_x = r.velocity_x; e.velocity_y = r.velocity_y;
        e.velocity_z = r.velocity_z;
    }
}

OREF void oref_height_batch(const ocean_sim* s, const float* xz, std::uint32_t count,
                            float* heights)
{
    const Ocean& o = *reinterpret_cast<const Ocean*>(s);
    for (std::uint32_t i = 0; i < count; ++i) heights[i] = o.height_at(xz[2 * i], xz[2 * i + 1]);
}

// The same query loops over `threads` host threads, each taking a contiguous
// slice of the points (the reference's queries are const and may run
// concurrently on one handle, include/ocean/ocean.hpp:17-31). threads == 0:
// std::thread::hardware_concurrency(), as the reference's thread_count = 0.
namespace {
template <typename Fn>
void split_threads(std::uint32_t count, std::uint32_t threads, Fn fn)
{
    if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    std::vector<std::thread> pool;
    const std::uint32_t per = (count + threads - 1) / threads;
    for (std::uint32_t t = 0; t < threads; ++t) {
        const std::uint32_t a = t * per, b = std::min(count, a + per);
        if (a >= b) break;
        pool.emplace_back([=] { fn(a, b); });
    }
    for (std::thread& th : pool) th.join();
}
}  // namespace

OREF std::uint32_t oref_hardware_concurrency() { return std::thread::hardware_concurrency(); }

OREF void oref_height_batch_mt(const ocean_sim* s, const float* xz, std::uint32_t count,
                               float* heights, std::uint32_t threads)
{
    const Ocean& o = *reinterpret_cast<const Ocean*>(s);
    split_threads(count, threads, [&](std::uint32_t a, std::uint32_t b) {
        for (std::uint32_t i = a; i < b; ++i) heights[i] = o.height_at(xz[2 * i], xz[2 * i + 1]);
    });
}

OREF void oref_cascade_height_batch_mt(void* c, const float* xz, std::uint32_t count,
                                       float* heights, std::uint32_t threads)
{
    const CascadeStack& st = *static_cast<const CascadeStack*>(c);
    split_threads(count, threads, [&](std::uint32_t a, std::uint32_t b) {
        for (std::uint32_t i = a; i < b; ++i) heights[i] = st.height_at(xz[2 * i], xz[2 * i + 1]);
    });
}

// ---- the viewer's half conversion (examples/viewer/half.hpp, header-only;
// its scalar path, which the header proves equal to its F16C path) ---------

#include "../examples/viewer/half.hpp"

OREF std::uint16_t oref_float_to_half(float v) { return viewer::float_to_half(v); }

OREF void oref_half_range(std::uint32_t first, std::uint32_t count, std::uint16_t* out)
{
    for (std::uint32_t i = 0; i < count; ++i) {
        const std::uint32_t bits = first + i;
        float f;
        std::memcpy(&f, &bits, sizeof f);
        out[i] = viewer::float_to_half(f);
    }
}

OREF void oref_pack_rgba16(std::uint16_t* dst, const float* src, const float* alpha,
                           std::size_t cells)
{
    viewer::pack_rgba16_scalar(dst, src, alpha, cells);
}

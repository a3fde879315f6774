// C ABI over the C++ classes: include/ocean/ocean.h (the reference's ABI) and
// include/ocean/ocean_amd.h (additions).
//
// Same firewall rules as the reference's src/c_api.cpp: the handle is the C++
// object itself, no exception crosses, NULL is tolerated everywhere, and
// struct_size decides how much of a caller's descriptor is read.
#include "ocean/ocean.h"
#include "ocean/ocean_amd.h"
#include "ocean/ocean.hpp"

#include "ocean_device.hpp"
#include "spectrum_host.hpp"

#include <hip/hip_runtime_api.h>

#include <cstring>
#include <limits>
#include <new>
#include <stdexcept>

namespace {

ocean::Ocean* to_cpp(ocean_sim* h) noexcept { return reinterpret_cast<ocean::Ocean*>(h); }
const ocean::Ocean* to_cpp(const ocean_sim* h) noexcept
{
    return reinterpret_cast<const ocean::Ocean*>(h);
}

// reference src/c_api.cpp:46-60
bool copy_desc(const ocean_desc* src, ocean_desc& dst) noexcept
{
    if (src == nullptr) return false;
    if (src->struct_size == 0 || src->struct_size > sizeof(ocean_desc)) return false;
    ocean_desc_init(&dst);
    std::memcpy(&dst, src, src->struct_size);
    dst.struct_size = sizeof(ocean_desc);
    return true;
}

// reference src/c_api.cpp:62-95
ocean::OceanDesc translate(const ocean_desc& c) noexcept
{
    ocean::OceanDesc d;
    d.size = c.size;
    d.patch_length = c.patch_length;
    d.spectrum.wind_speed = c.spectrum.wind_speed;
    d.spectrum.fetch = c.spectrum.fetch;
    d.spectrum.wind_direction = c.spectrum.wind_direction;
    d.spectrum.peak_enhancement = c.spectrum.peak_enhancement;
    d.spectrum.swell = c.spectrum.swell;
    d.spectrum.small_wave_cutoff = c.spectrum.small_wave_cutoff;
    d.spectrum.gravity = c.spectrum.gravity;
    d.spectrum.depth = c.water_depth;
    d.choppiness = c.choppiness;
    d.foam_threshold = c.foam_threshold;
    d.seed = c.seed;
    d.parallel_for = reinterpret_cast<ocean::ParallelForFn>(c.parallel_for);
    d.parallel_for_user = c.parallel_for_user;
    d.thread_count = c.thread_count;
    d.compute_velocity = (c.compute_velocity != 0);
    return d;
}

void to_surface_ex(const ocean::Surface& s, ocean_surface_ex* out) noexcept
{
    out->height = s.height;
    out->offset_x = s.offset_x;
    out->offset_z = s.offset_z;
    out->normal_x = s.normal_x;
    out->normal_y = s.normal_y;
    out->normal_z = s.normal_z;
    out->foam = s.foam;
    out->velocity_x = s.velocity_x;
    out->velocity_y = s.velocity_y;
    out->velocity_z = s.velocity_z;
}

}  // namespace

extern "C" {

const char* ocean_status_string(ocean_status status)
{
    switch (status) {
        case OCEAN_OK: return "ok";
        case OCEAN_ERROR_INVALID_ARG: return "invalid argument";
        case OCEAN_ERROR_OUT_OF_MEMORY: return "out of memory";
        case OCEAN_ERROR_UNKNOWN: return "unknown error";
    }
    return "unrecognised status";
}

void ocean_desc_init(ocean_desc* desc)
{
    if (desc == nullptr) return;
    const ocean::OceanDesc d{};
    std::memset(desc, 0, sizeof(*desc));
    desc->struct_size = sizeof(ocean_desc);
    desc->size = d.size;
    desc->patch_length = d.patch_length;
    desc->spectrum.wind_speed = d.spectrum.wind_speed;
    desc->spectrum.fetch = d.spectrum.fetch;
    desc->spectrum.wind_direction = d.spectrum.wind_direction;
    desc->spectrum.peak_enhancement = d.spectrum.peak_enhancement;
    desc->spectrum.swell = d.spectrum.swell;
    desc->spectrum.small_wave_cutoff = d.spectrum.small_wave_cutoff;
    desc->spectrum.gravity = d.spectrum.gravity;
    desc->choppiness = d.choppiness;
    desc->foam_threshold = d.foam_threshold;
    desc->seed = d.seed;
    desc->thread_count = d.thread_count;
    desc->water_depth = d.spectrum.depth;
    desc->compute_velocity = d.compute_velocity ? 1 : 0;
}

ocean_sim* ocean_create(const ocean_desc* desc, ocean_status* out_status)
{
    auto fail = [&](ocean_status s) -> ocean_sim* {
        if (out_status != nullptr) *out_status = s;
        return nullptr;
    };
    ocean_desc resolved;
    if (!copy_desc(desc, resolved)) return fail(OCEAN_ERROR_INVALID_ARG);
    try {
        auto* sim = new ocean::Ocean(translate(resolved));
        if (out_status != nullptr) *out_status = OCEAN_OK;
        return reinterpret_cast<ocean_sim*>(sim);
    } catch (const std::invalid_argument&) {
        return fail(OCEAN_ERROR_INVALID_ARG);
    } catch (const std::bad_alloc&) {
        return fail(OCEAN_ERROR_OUT_OF_MEMORY);
    } catch (...) {
        return fail(OCEAN_ERROR_UNKNOWN);
    }
}

void ocean_destroy(ocean_sim* sim) { delete to_cpp(sim); }

void ocean_update(ocean_sim* sim, double time)
{
    if (sim == nullptr) return;
    to_cpp(sim)->update(time);
}

ocean_buffers ocean_get_buffers(const ocean_sim* sim)
{
    ocean_buffers out;
    std::memset(&out, 0, sizeof(out));
    if (sim == nullptr) return out;
    const ocean::Buffers b = to_cpp(sim)->buffers();
    out.displacement = b.displacement;
    out.normal = b.normal;
    out.size = b.size;
    out.patch_length = b.patch_length;
    return out;
}

double ocean_get_time(const ocean_sim* sim) { return sim == nullptr ? 0.0 : to_cpp(sim)->time(); }

float ocean_height_at(const ocean_sim* sim, float world_x, float world_z)
{
    return sim == nullptr ? 0.0f : to_cpp(sim)->height_at(world_x, world_z);
}

ocean_surface ocean_sample_at(const ocean_sim* sim, float world_x, float world_z)
{
    ocean_surface out;
    std::memset(&out, 0, sizeof(out));
    out.normal_y = 1.0f;
    if (sim == nullptr) return out;
    const ocean::Surface s = to_cpp(sim)->sample_at(world_x, world_z);
    out.height = s.height;
    out.offset_x = s.offset_x;
    out.offset_z = s.offset_z;
    out.normal_x = s.normal_x;
    out.normal_y = s.normal_y;
    out.normal_z = s.normal_z;
    out.foam = s.foam;
    return out;
}

void ocean_sample_ex(const ocean_sim* sim, float world_x, float world_z, ocean_surface_ex* out)
{
    if (out == nullptr) return;
    std::memset(out, 0, sizeof(*out));
    out->normal_y = 1.0f;
    if (sim == nullptr) return;
    to_surface_ex(to_cpp(sim)->sample_at(world_x, world_z), out);
}

const float* ocean_get_velocity(const ocean_sim* sim)
{
    if (sim == nullptr) return nullptr;
    return to_cpp(sim)->buffers().velocity;
}

const char* ocean_simd_level(void) { return "gfx950"; }

const char* ocean_force_simd_level(const char* name)
{
    (void)name;
    return ocean_simd_level();
}

void ocean_version(uint32_t* major, uint32_t* minor, uint32_t* patch)
{
    if (major != nullptr) *major = 1;
    if (minor != nullptr) *minor = 0;
    if (patch != nullptr) *patch = 0;
}

// ---- ocean_amd.h: device buffers, batched queries, host tables ------------

int ocean_amd_get_device_buffers(const ocean_sim* sim, ocean_amd_device_buffers* out)
{
    if (sim == nullptr || out == nullptr) return OCEAN_ERROR_INVALID_ARG;
    const ocean::DeviceBuffers b = to_cpp(sim)->device_buffers();
    out->displacement = b.displacement;
    out->normal = b.normal;
    out->velocity = b.velocity;
    out->size = b.size;
    out->patch_length = b.patch_length;
    out->stream = b.stream;
    return OCEAN_OK;
}

int ocean_amd_bind_outputs(ocean_sim* sim, float* displacement_dev, float* normal_dev,
                           float* velocity_dev)
{
    if (sim == nullptr) return OCEAN_ERROR_INVALID_ARG;
    ocean::Ocean& o = *to_cpp(sim);
    if (o.bind_outputs(displacement_dev, normal_dev, velocity_dev)) return OCEAN_OK;
    return o.last_error() != 0 ? OCEAN_ERROR_UNKNOWN : OCEAN_ERROR_INVALID_ARG;
}

int ocean_amd_profile_frame(ocean_sim* sim, double t0, double dt, uint32_t iters, float* row_ms,
                            float* col_ms)
{
    if (sim == nullptr || iters == 0) return OCEAN_ERROR_INVALID_ARG;
    ocean::Ocean& o = *to_cpp(sim);
    auto& dev = o.device();
    dev.make_current();
    hipEvent_t ev[3];
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) return OCEAN_ERROR_UNKNOWN;
    // Back-to-back launches of each pass between two events, so the mean
    // excludes per-launch event overhead (comparable to a profiler's kernel
    // duration plus the dispatch gap).
    hipEventRecord(ev[0], dev.stream);
    for (uint32_t i = 0; i < iters; ++i)
        dev.note(ocean_hip_frame_passes(&dev.level, (t0 + dt * i) * 0.159154943091895335769,
                                        dev.stream, 1u));
    hipEventRecord(ev[1], dev.stream);
    for (uint32_t i = 0; i < iters; ++i)
        dev.note(ocean_hip_frame_passes(&dev.level, 0.0, dev.stream, 2u));
    hipEventRecord(ev[2], dev.stream);
    int rc = hipEventSynchronize(ev[2]) == hipSuccess ? OCEAN_OK : OCEAN_ERROR_UNKNOWN;
    float a = 0.0f, b = 0.0f;
    hipEventElapsedTime(&a, ev[0], ev[1]);
    hipEventElapsedTime(&b, ev[1], ev[2]);
    for (auto& e : ev) hipEventDestroy(e);
    o.update(t0 + dt * (iters - 1));  // a complete, consistent frame again
    o.synchronize();
    if (row_ms) *row_ms = a / static_cast<float>(iters);
    if (col_ms) *col_ms = b / static_cast<float>(iters);
    return dev.status == 0 ? rc : OCEAN_ERROR_UNKNOWN;
}

int ocean_amd_frame_timing(ocean_sim* sim, uint32_t every)
{
    if (sim == nullptr) return OCEAN_ERROR_INVALID_ARG;
    return to_cpp(sim)->device().set_frame_timing(every) == 0 ? OCEAN_OK : OCEAN_ERROR_UNKNOWN;
}

int ocean_amd_frame_timing_read(ocean_sim* sim, float* row_ms, float* col_ms, uint32_t* samples)
{
    if (sim == nullptr) return OCEAN_ERROR_INVALID_ARG;
    auto& dev = to_cpp(sim)->device();
    if (dev.read_frame_timing(row_ms, col_ms, samples) != 0) return OCEAN_ERROR_UNKNOWN;
    return dev.status == 0 ? OCEAN_OK : OCEAN_ERROR_UNKNOWN;
}

int ocean_amd_get_tables(const ocean_sim* sim, float* h0, float* omega, uint32_t* repaired)
{
    if (sim == nullptr) return OCEAN_ERROR_INVALID_ARG;
    auto& dev = to_cpp(sim)->device();
    dev.make_current();
    const std::size_t cells = std::size_t(dev.level.n) * dev.level.n;
    if (h0 != nullptr)
        dev.note(hipMemcpyAsync(h0, dev.level.h0, cells * 8, hipMemcpyDeviceToHost, dev.stream));
    if (omega != nullptr)
        dev.note(hipMemcpyAsync(omega, dev.level.omega, cells * 4, hipMemcpyDeviceToHost, dev.stream));
    dev.note(hipStreamSynchronize(dev.stream));
    if (repaired != nullptr) *repaired = dev.init_repaired;
    return dev.status == 0 ? OCEAN_OK : OCEAN_ERROR_UNKNOWN;
}

namespace {

// A level's shape word: row variant (bits 0-7), column variant (8-15),
// precision mode (16-23), layout of the fields between the passes (24-31).
uint32_t shape_mode(uint32_t shape) { return (shape >> 16) & 0xffu; }
uint32_t shape_layout(uint32_t shape) { return (shape >> 24) & 0xffu; }

int shape_counts(const ocean::detail::OceanDevice& dev, uint32_t layout, int& r, int& c)
{
    return ocean_hip_frame_shapes_layout(dev.level.n, uint32_t(dev.fields), shape_mode(dev.level.shape),
                                         layout, &r, &c);
}

void upload_shape(ocean::detail::OceanDevice& dev, uint32_t shape)
{
    dev.make_current();
    dev.level.shape = shape;
    dev.note(hipMemcpyAsync(dev.d_level, &dev.level, sizeof(dev.level), hipMemcpyHostToDevice, dev.stream));
    dev.note(hipStreamSynchronize(dev.stream));
}

}  // namespace

int ocean_amd_frame_shapes(const ocean_sim* sim, uint32_t* row_variants, uint32_t* col_variants)
{
    if (sim == nullptr) return OCEAN_ERROR_INVALID_ARG;
    const auto& dev = to_cpp(sim)->device();
    int r = 0, c = 0;
    if (shape_counts(dev, shape_layout(dev.level.shape), r, c) != 0) return OCEAN_ERROR_INVALID_ARG;
    if (row_variants) *row_variants = static_cast<uint32_t>(r);
    if (col_variants) *col_variants = static_cast<uint32_t>(c);
    return OCEAN_OK;
}

int ocean_amd_set_frame_layout(ocean_sim* sim, uint32_t layout)
{
    if (sim == nullptr) return OCEAN_ERROR_INVALID_ARG;
    auto& dev = to_cpp(sim)->device();
    int r = 0, c = 0;
    if (shape_counts(dev, layout, r, c) != 0) return OCEAN_ERROR_INVALID_ARG;
    // the default shapes of the new layout
    upload_shape(dev, (shape_mode(dev.level.shape) << 16) | (layout << 24));
    return dev.status == 0 ? OCEAN_OK : OCEAN_ERROR_UNKNOWN;
}

uint32_t ocean_amd_get_frame_layout(const ocean_sim* sim)
{
    return sim == nullptr ? 0u : shape_layout(to_cpp(sim)->device().level.shape);
}

int ocean_amd_set_frame_shape(ocean_sim* sim, uint32_t row_variant, uint32_t col_variant)
{
    if (sim == nullptr) return OCEAN_ERROR_INVALID_ARG;
    ocean::Ocean& o = *to_cpp(sim);
    auto& dev = o.device();
    int r = 0, c = 0;
    const uint32_t mode_bits = dev.level.shape & 0xffff0000u;
    if (shape_counts(dev, shape_layout(dev.level.shape), r, c) != 0 || row_variant >= uint32_t(r) ||
        col_variant >= uint32_t(c))
        return OCEAN_ERROR_INVALID_ARG;
    dev.make_current();
    dev.level.shape = mode_bits | row_variant | (col_variant << 8);
    dev.note(hipMemcpyAsync(dev.d_level, &dev.level, sizeof(dev.level), hipMemcpyHostToDevice, dev.stream));
    dev.note(hipStreamSynchronize(dev.stream));
    return dev.status == 0 ? OCEAN_OK : OCEAN_ERROR_UNKNOWN;
}

int ocean_amd_autotune_frame(ocean_sim* sim, uint32_t reps, uint32_t* row_variant,
                             uint32_t* col_variant, float* row_ms, float* col_ms)
{
    if (sim == nullptr || reps == 0 || reps > 256) return OCEAN_ERROR_INVALID_ARG;
    ocean::Ocean& o = *to_cpp(sim);
    auto& dev = o.device();
    dev.make_current();
    // Each candidate is timed inside whole frames (row pass then column pass,
    // the other pass at its current pick), by the dispatches' own events, so
    // it sees the caches a frame leaves: the column pass reads fields the
    // row pass just wrote on other XCDs, the row pass follows a column pass.
    // Every layout of the fields between the passes is tuned this way, and
    // the layout with the smallest row + column time is kept (row-major
    // unless a tiled layout is at least 5% faster).
    std::vector<hipEvent_t> ev(std::size_t(reps) * 4, nullptr);
    bool ok = true;
    for (auto& e : ev) ok = ok && hipEventCreate(&e) == hipSuccess;
    auto release = [&] {
        for (auto& e : ev)
            if (e) hipEventDestroy(e);
    };
    if (!ok) {
        release();
        return OCEAN_ERROR_UNKNOWN;
    }
    const double t2 = o.time() * 0.159154943091895335769;
    const uint32_t saved = dev.level.shape;
    const uint32_t mode = shape_mode(saved);
    float best_total = 1e30f, best_out[2] = {0.0f, 0.0f};
    uint32_t best_shape = saved;
    for (uint32_t layout = 0; layout < 4u && ok; ++layout) {
        int nr = 0, nc = 0;
        if (shape_counts(dev, layout, nr, nc) != 0) continue;
        const uint32_t base = (mode << 16) | (layout << 24);
        float best[2] = {1e30f, 1e30f};
        uint32_t pick[2] = {0u, 0u};
        if (layout == shape_layout(saved)) {
            pick[0] = saved & 0xffu;
            pick[1] = (saved >> 8) & 0xffu;
        }
        for (unsigned pass = 0; pass < 2 && ok; ++pass) {
            const int count = pass == 0 ? nr : nc;
            for (int v = 0; v < count && ok; ++v) {
                dev.level.shape =
                    base | (pass == 0 ? (uint32_t(v) | (pick[1] << 8)) : (pick[0] | (uint32_t(v) << 8)));
                dev.note(ocean_hip_frame(&dev.level, t2, dev.stream));  // warm
                for (uint32_t i = 0; i < reps; ++i)
                    dev.note(ocean_hip_frame_timed(&dev.level, t2, dev.stream, &ev[std::size_t(i) * 4]));
                ok = hipStreamSynchronize(dev.stream) == hipSuccess;
                double sum = 0.0;
                for (uint32_t i = 0; i < reps && ok; ++i) {
                    float ms = 0.0f;
                    hipEvent_t* e = &ev[std::size_t(i) * 4 + 2 * pass];
                    ok = hipEventElapsedTime(&ms, e[0], e[1]) == hipSuccess;
                    sum += ms;
                }
                const float ms = float(sum / reps);
                if (ok && ms < best[pass]) {
                    best[pass] = ms;
                    pick[pass] = uint32_t(v);
                }
            }
        }
        // a tiled layout must beat the row-major one by 5%: these back-to-back
        // frames flatter the tiled row pass, whose pieced stores cost more
        // after other kernels' traffic (at 512^2 with velocity: 2.8% ahead
        // here, even in a frame loop with the foam step and queries)
        const float margin = layout == 0 ? 1.0f : 0.95f;
        if (ok && best[0] + best[1] < best_total * margin) {
            best_total = best[0] + best[1];
            best_out[0] = best[0];
            best_out[1] = best[1];
            best_shape = base | pick[0] | (pick[1] << 8);
        }
    }
    release();
    if (!ok) {
        dev.level.shape = saved;
        return OCEAN_ERROR_UNKNOWN;
    }
    dev.level.shape = best_shape;
    dev.note(hipMemcpyAsync(dev.d_level, &dev.level, sizeof(dev.level), hipMemcpyHostToDevice, dev.stream));
    dev.note(ocean_hip_frame(&dev.level, t2, dev.stream));  // a complete frame at the current time
    dev.note(hipStreamSynchronize(dev.stream));
    if (row_variant) *row_variant = best_shape & 0xffu;
    if (col_variant) *col_variant = (best_shape >> 8) & 0xffu;
    if (row_ms) *row_ms = best_out[0];
    if (col_ms) *col_ms = best_out[1];
    return dev.status == 0 ? OCEAN_OK : OCEAN_ERROR_UNKNOWN;
}

int ocean_amd_set_precision(ocean_sim* sim, uint32_t mode)
{
    if (sim == nullptr || mode > OCEAN_AMD_PRECISION_FAST_HALF) return OCEAN_ERROR_INVALID_ARG;
    ocean::Ocean& o = *to_cpp(sim);
    if (!o.set_precision(mode)) return o.last_error() != 0 ? OCEAN_ERROR_UNKNOWN : OCEAN_ERROR_INVALID_ARG;
    return OCEAN_OK;
}

uint32_t ocean_amd_get_precision(const ocean_sim* sim)
{
    return sim == nullptr ? 0u : to_cpp(sim)->precision();
}

void ocean_amd_synchronize(const ocean_sim* sim)
{
    if (sim != nullptr) to_cpp(sim)->synchronize();
}

int ocean_amd_last_error(const ocean_sim* sim) { return sim == nullptr ? 0 : to_cpp(sim)->last_error(); }

int ocean_amd_sample_batch(const ocean_sim* sim, const float* xz, uint32_t count,
                           ocean_surface_ex* out)
{
    static_assert(sizeof(ocean_surface_ex) == sizeof(ocean::Surface));
    if (sim == nullptr) return OCEAN_ERROR_INVALID_ARG;
    if (count == 0) return OCEAN_OK;
    if (xz == nullptr || out == nullptr) return OCEAN_ERROR_INVALID_ARG;
    return to_cpp(sim)->sample_batch(xz, count, reinterpret_cast<ocean::Surface*>(out))
               ? OCEAN_OK
               : OCEAN_ERROR_UNKNOWN;
}

int ocean_amd_sample_batch_device(const ocean_sim* sim, const float* xz_device, uint32_t count,
                                  ocean_surface_ex* out_device, float* heights_device)
{
    if (sim == nullptr) return OCEAN_ERROR_INVALID_ARG;
    if (count == 0) return OCEAN_OK;
    if (xz_device == nullptr || (out_device == nullptr && heights_device == nullptr))
        return OCEAN_ERROR_INVALID_ARG;
    return to_cpp(sim)->sample_batch_device(xz_device, count,
                                            reinterpret_cast<ocean::Surface*>(out_device),
                                            heights_device)
               ? OCEAN_OK
               : OCEAN_ERROR_UNKNOWN;
}

int ocean_amd_build_spectrum(const ocean_desc* desc, float* h0, float* omega)
{
    ocean_desc resolved;
    if (!copy_desc(desc, resolved) || h0 == nullptr || omega == nullptr)
        return OCEAN_ERROR_INVALID_ARG;
    const ocean::OceanDesc d = translate(resolved);
    if (d.size < 16 || d.size > 4096 || (d.size & (d.size - 1)) != 0)
        return OCEAN_ERROR_INVALID_ARG;
    oceanamd::build_spectrum_tables(d, h0, omega, 0, d.size);
    return OCEAN_OK;
}

int ocean_amd_build_twiddles(uint32_t n, float* tw)
{
    if (tw == nullptr || n < 2 || (n & (n - 1)) != 0) return OCEAN_ERROR_INVALID_ARG;
    oceanamd::build_twiddles(n, tw);
    return OCEAN_OK;
}

}  // extern "C"

// ---- ocean_amd.h: foam and cascades ---------------------------------------

#include "ocean/cascade.hpp"
#include "ocean/foam.hpp"

#include <vector>

namespace {

ocean::FoamField* to_foam(ocean_amd_foam* f) noexcept
{
    return reinterpret_cast<ocean::FoamField*>(f);
}
const ocean::FoamField* to_foam(const ocean_amd_foam* f) noexcept
{
    return reinterpret_cast<const ocean::FoamField*>(f);
}
ocean::CascadeStack* to_stack(ocean_amd_cascade* s) noexcept
{
    return reinterpret_cast<ocean::CascadeStack*>(s);
}
const ocean::CascadeStack* to_stack(const ocean_amd_cascade* s) noexcept
{
    return reinterpret_cast<const ocean::CascadeStack*>(s);
}

}  // namespace

extern "C" {

void ocean_amd_foam_desc_init(ocean_amd_foam_desc* desc)
{
    if (desc == nullptr) return;
    const ocean::FoamDesc d{};
    std::memset(desc, 0, sizeof(*desc));
    desc->struct_size = sizeof(*desc);
    desc->decay = d.decay;
    desc->source_gain = d.source_gain;
    desc->advect_scale = d.advect_scale;
    desc->wind_drift_x = d.wind_drift_x;
    desc->wind_drift_z = d.wind_drift_z;
    desc->fixed_dt = d.fixed_dt;
    desc->max_substeps = d.max_substeps;
}

ocean_amd_foam* ocean_amd_foam_create(const ocean_sim* sim, const ocean_amd_foam_desc* desc,
                                      ocean_status* out_status)
{
    auto fail = [&](ocean_status s) -> ocean_amd_foam* {
        if (out_status != nullptr) *out_status = s;
        return nullptr;
    };
    if (sim == nullptr || desc == nullptr || desc->struct_size == 0 ||
        desc->struct_size > sizeof(ocean_amd_foam_desc))
        return fail(OCEAN_ERROR_INVALID_ARG);
    ocean_amd_foam_desc c;
    ocean_amd_foam_desc_init(&c);
    std::memcpy(&c, desc, desc->struct_size);
    ocean::FoamDesc d;
    d.decay = c.decay;
    d.source_gain = c.source_gain;
    d.advect_scale = c.advect_scale;
    d.wind_drift_x = c.wind_drift_x;
    d.wind_drift_z = c.wind_drift_z;
    d.fixed_dt = c.fixed_dt;
    d.max_substeps = c.max_substeps;
    try {
        auto* f = new ocean::FoamField(*to_cpp(sim), d);
        if (out_status != nullptr) *out_status = OCEAN_OK;
        return reinterpret_cast<ocean_amd_foam*>(f);
    } catch (const std::invalid_argument&) {
        return fail(OCEAN_ERROR_INVALID_ARG);
    } catch (const std::bad_alloc&) {
        return fail(OCEAN_ERROR_OUT_OF_MEMORY);
    } catch (...) {
        return fail(OCEAN_ERROR_UNKNOWN);
    }
}

void ocean_amd_foam_destroy(ocean_amd_foam* foam) { delete to_foam(foam); }
void ocean_amd_foam_update(ocean_amd_foam* foam, float dt)
{
    if (foam != nullptr) to_foam(foam)->update(dt);
}
int ocean_amd_foam_update_sample_device(ocean_amd_foam* foam, float dt, const float* xz_device,
                                        uint32_t count, ocean_surface_ex* out_device,
                                        float* heights_device)
{
    if (foam == nullptr) return OCEAN_ERROR_INVALID_ARG;
    ocean::FoamField* f = to_foam(foam);
    if (f->update_and_sample(dt, xz_device, count, reinterpret_cast<ocean::Surface*>(out_device),
                             heights_device))
        return OCEAN_OK;
    return f->ocean().last_error() != 0 ? OCEAN_ERROR_UNKNOWN : OCEAN_ERROR_INVALID_ARG;
}
void ocean_amd_foam_clear(ocean_amd_foam* foam)
{
    if (foam != nullptr) to_foam(foam)->clear();
}
const float* ocean_amd_foam_data(const ocean_amd_foam* foam)
{
    return foam == nullptr ? nullptr : to_foam(foam)->data();
}
const float* ocean_amd_foam_device_data(const ocean_amd_foam* foam)
{
    return foam == nullptr ? nullptr : to_foam(foam)->device_data();
}
uint32_t ocean_amd_foam_last_substeps(const ocean_amd_foam* foam)
{
    return foam == nullptr ? 0u : to_foam(foam)->last_substeps();
}
float ocean_amd_foam_at(const ocean_amd_foam* foam, float world_x, float world_z)
{
    return foam == nullptr ? 0.0f : to_foam(foam)->foam_at(world_x, world_z);
}
double ocean_amd_foam_coverage(const ocean_amd_foam* foam)
{
    return foam == nullptr ? 0.0 : to_foam(foam)->coverage();
}

ocean_amd_cascade* ocean_amd_cascade_create(const ocean_desc* levels, uint32_t count,
                                            ocean_status* out_status)
{
    auto fail = [&](ocean_status s) -> ocean_amd_cascade* {
        if (out_status != nullptr) *out_status = s;
        return nullptr;
    };
    if (levels == nullptr || count == 0) return fail(OCEAN_ERROR_INVALID_ARG);
    std::vector<ocean::CascadeLevel> descs;
    descs.reserve(count);
    for (uint32_t i = 0; i < count; ++i) {
        ocean_desc r;
        if (!copy_desc(&levels[i], r)) return fail(OCEAN_ERROR_INVALID_ARG);
        descs.push_back(translate(r));
    }
    try {
        auto* s = new ocean::CascadeStack(std::span<const ocean::CascadeLevel>(descs));
        if (out_status != nullptr) *out_status = OCEAN_OK;
        return reinterpret_cast<ocean_amd_cascade*>(s);
    } catch (const std::invalid_argument&) {
        return fail(OCEAN_ERROR_INVALID_ARG);
    } catch (const std::bad_alloc&) {
        return fail(OCEAN_ERROR_OUT_OF_MEMORY);
    } catch (...) {
        return fail(OCEAN_ERROR_UNKNOWN);
    }
}

void ocean_amd_cascade_destroy(ocean_amd_cascade* stack) { delete to_stack(stack); }
void ocean_amd_cascade_update(ocean_amd_cascade* stack, double time)
{
    if (stack != nullptr) to_stack(stack)->update(time);
}
uint32_t ocean_amd_cascade_level_count(const ocean_amd_cascade* stack)
{
    return stack == nullptr ? 0u : static_cast<uint32_t>(to_stack(stack)->level_count());
}
ocean_buffers ocean_amd_cascade_get_buffers(const ocean_amd_cascade* stack, uint32_t level)
{
    ocean_buffers out;
    std::memset(&out, 0, sizeof(out));
    if (stack == nullptr || level >= to_stack(stack)->level_count()) return out;
    const ocean::Buffers b = to_stack(stack)->buffers(level);
    out.displacement = b.displacement;
    out.normal = b.normal;
    out.size = b.size;
    out.patch_length = b.patch_length;
    return out;
}
void ocean_amd_cascade_sample(const ocean_amd_cascade* stack, float world_x, float world_z,
                              ocean_surface_ex* out)
{
    if (out == nullptr) return;
    std::memset(out, 0, sizeof(*out));
    out->normal_y = 1.0f;
    if (stack == nullptr) return;
    to_surface_ex(to_stack(stack)->sample_at(world_x, world_z), out);
}
int ocean_amd_cascade_sample_batch(const ocean_amd_cascade* stack, const float* xz,
                                   uint32_t count, ocean_surface_ex* out)
{
    if (stack == nullptr) return OCEAN_ERROR_INVALID_ARG;
    if (count == 0) return OCEAN_OK;
    if (xz == nullptr || out == nullptr) return OCEAN_ERROR_INVALID_ARG;
    return to_stack(stack)->sample_batch(xz, count, reinterpret_cast<ocean::Surface*>(out))
               ? OCEAN_OK
               : OCEAN_ERROR_UNKNOWN;
}

}  // extern "C"

// ---- ocean_amd.h: batches of seas ------------------------------------------

#include "ocean/batch.hpp"

namespace {

ocean::OceanBatch* to_batch(ocean_amd_batch* b) noexcept
{
    return reinterpret_cast<ocean::OceanBatch*>(b);
}
const ocean::OceanBatch* to_batch(const ocean_amd_batch* b) noexcept
{
    return reinterpret_cast<const ocean::OceanBatch*>(b);
}

bool batch_range(const ocean_amd_batch* b, uint32_t first, uint32_t count) noexcept
{
    if (b == nullptr || count == 0) return false;
    const std::size_t n = to_batch(b)->level_count();
    return first < n && count <= n - first;
}

}  // namespace

extern "C" {

ocean_amd_batch* ocean_amd_batch_create(const ocean_desc* levels, uint32_t count,
                                        ocean_status* out_status)
{
    auto fail = [&](ocean_status s) -> ocean_amd_batch* {
        if (out_status != nullptr) *out_status = s;
        return nullptr;
    };
    if (levels == nullptr || count == 0) return fail(OCEAN_ERROR_INVALID_ARG);
    std::vector<ocean::OceanDesc> descs;
    descs.reserve(count);
    for (uint32_t i = 0; i < count; ++i) {
        ocean_desc r;
        if (!copy_desc(&levels[i], r)) return fail(OCEAN_ERROR_INVALID_ARG);
        descs.push_back(translate(r));
    }
    try {
        auto* b = new ocean::OceanBatch(std::span<const ocean::OceanDesc>(descs));
        if (out_status != nullptr) *out_status = OCEAN_OK;
        return reinterpret_cast<ocean_amd_batch*>(b);
    } catch (const std::invalid_argument&) {
        return fail(OCEAN_ERROR_INVALID_ARG);
    } catch (const std::bad_alloc&) {
        return fail(OCEAN_ERROR_OUT_OF_MEMORY);
    } catch (...) {
        return fail(OCEAN_ERROR_UNKNOWN);
    }
}

void ocean_amd_batch_destroy(ocean_amd_batch* batch) { delete to_batch(batch); }
void ocean_amd_batch_update(ocean_amd_batch* batch, double time)
{
    if (batch != nullptr) to_batch(batch)->update(time);
}
uint32_t ocean_amd_batch_level_count(const ocean_amd_batch* batch)
{
    return batch == nullptr ? 0u : static_cast<uint32_t>(to_batch(batch)->level_count());
}
const ocean_sim* ocean_amd_batch_level(const ocean_amd_batch* batch, uint32_t level)
{
    if (!batch_range(batch, level, 1)) return nullptr;
    return reinterpret_cast<const ocean_sim*>(&to_batch(batch)->level(level));
}
int ocean_amd_batch_sample(const ocean_amd_batch* batch, uint32_t first, uint32_t level_count,
                           const float* xz, uint32_t count, ocean_surface_ex* out)
{
    if (!batch_range(batch, first, level_count)) return OCEAN_ERROR_INVALID_ARG;
    if (count == 0) return OCEAN_OK;
    if (xz == nullptr || out == nullptr) return OCEAN_ERROR_INVALID_ARG;
    return to_batch(batch)->sample(first, level_count, xz, count,
                                   reinterpret_cast<ocean::Surface*>(out))
               ? OCEAN_OK
               : OCEAN_ERROR_UNKNOWN;
}
int ocean_amd_batch_sample_device(const ocean_amd_batch* batch, uint32_t first,
                                  uint32_t level_count, const float* xz_dev, uint32_t count,
                                  ocean_surface_ex* out_dev, float* heights_dev)
{
    if (!batch_range(batch, first, level_count)) return OCEAN_ERROR_INVALID_ARG;
    if (count == 0) return OCEAN_OK;
    if (xz_dev == nullptr || (out_dev == nullptr && heights_dev == nullptr))
        return OCEAN_ERROR_INVALID_ARG;
    return to_batch(batch)->sample_device(first, level_count, xz_dev, count,
                                          reinterpret_cast<ocean::Surface*>(out_dev), heights_dev)
               ? OCEAN_OK
               : OCEAN_ERROR_UNKNOWN;
}
void ocean_amd_batch_autotune(ocean_amd_batch* batch, uint32_t reps)
{
    if (batch != nullptr) to_batch(batch)->autotune(reps);
}

int ocean_amd_batch_set_precision(ocean_amd_batch* batch, uint32_t mode)
{
    if (batch == nullptr || mode > OCEAN_AMD_PRECISION_FAST_HALF) return OCEAN_ERROR_INVALID_ARG;
    return to_batch(batch)->set_precision(mode) ? OCEAN_OK : OCEAN_ERROR_UNKNOWN;
}
uint32_t ocean_amd_batch_group_count(const ocean_amd_batch* batch)
{
    return batch == nullptr ? 0u : static_cast<uint32_t>(to_batch(batch)->group_count());
}
int ocean_amd_batch_group_info(const ocean_amd_batch* batch, uint32_t group,
                               ocean_amd_batch_group* out)
{
    if (batch == nullptr || out == nullptr || group >= to_batch(batch)->group_count())
        return OCEAN_ERROR_INVALID_ARG;
    const ocean::BatchGroupInfo g = to_batch(batch)->group(group);
    out->size = g.size;
    out->velocity = g.velocity ? 1u : 0u;
    out->levels = g.levels;
    out->items = g.items;
    out->grid_row = g.grid_row;
    out->grid_col = g.grid_col;
    out->row_ms = g.row_ms;
    out->col_ms = g.col_ms;
    out->radix_row = g.radix_row;
    out->radix_col = g.radix_col;
    out->layout = g.layout;
    return OCEAN_OK;
}
int ocean_amd_batch_set_group_launch(ocean_amd_batch* batch, uint32_t group, uint32_t layout,
                                     uint32_t radix_row, uint32_t grid_row, uint32_t radix_col,
                                     uint32_t grid_col)
{
    if (batch == nullptr || group >= to_batch(batch)->group_count()) return OCEAN_ERROR_INVALID_ARG;
    return to_batch(batch)->set_group_launch(group, layout, radix_row, grid_row, radix_col, grid_col)
               ? OCEAN_OK
               : OCEAN_ERROR_INVALID_ARG;
}
int ocean_amd_batch_checksums(const ocean_amd_batch* batch, uint32_t first, uint32_t count,
                              uint64_t* out)
{
    if (!batch_range(batch, first, count) || out == nullptr) return OCEAN_ERROR_INVALID_ARG;
    return to_batch(batch)->checksums(first, count, out) ? OCEAN_OK : OCEAN_ERROR_UNKNOWN;
}
int ocean_amd_batch_synchronize(const ocean_amd_batch* batch)
{
    if (batch == nullptr) return OCEAN_ERROR_INVALID_ARG;
    to_batch(batch)->synchronize();
    return to_batch(batch)->last_error() == 0 ? OCEAN_OK : OCEAN_ERROR_UNKNOWN;
}
int ocean_amd_batch_last_error(const ocean_amd_batch* batch)
{
    return batch == nullptr ? 0 : to_batch(batch)->last_error();
}

}  // extern "C"

// ---- ocean.h: local interaction (reference src/c_api.cpp:256-518) ----------

#include "ocean/interaction.hpp"

namespace {

ocean::InteractionField* to_field(ocean_interaction* f) noexcept
{
    return reinterpret_cast<ocean::InteractionField*>(f);
}
const ocean::InteractionField* to_cfield(const ocean_interaction* f) noexcept
{
    return reinterpret_cast<const ocean::InteractionField*>(f);
}

// reference src/c_api.cpp:289-301
bool copy_interaction_desc(const ocean_interaction_desc* src, ocean_interaction_desc& dst) noexcept
{
    if (src == nullptr) return false;
    if (src->struct_size == 0 || src->struct_size > sizeof(ocean_interaction_desc)) return false;
    ocean_interaction_desc_init(&dst);
    std::memcpy(&dst, src, src->struct_size);
    dst.struct_size = sizeof(ocean_interaction_desc);
    return true;
}

// reference src/c_api.cpp:303-325
ocean::InteractionDesc translate(const ocean_interaction_desc& c) noexcept
{
    ocean::InteractionDesc d;
    d.size = c.size;
    d.extent = c.extent;
    d.kernel_radius = c.kernel_radius;
    d.damping = c.damping;
    d.absorb_cells = c.absorb_cells;
    d.fixed_dt = c.fixed_dt;
    d.max_substeps = c.max_substeps;
    d.max_sources = c.max_sources;
    d.gravity = c.gravity;
    d.depth = c.water_depth;
    d.obstruction = (c.obstruction == OCEAN_OBSTRUCTION_DIRICHLET) ? ocean::Obstruction::Dirichlet
                                                                  : ocean::Obstruction::Neumann;
    d.parallel_for = reinterpret_cast<ocean::ParallelForFn>(c.parallel_for);
    d.parallel_for_user = c.parallel_for_user;
    d.thread_count = c.thread_count;
    return d;
}

}  // namespace

extern "C" {

void ocean_interaction_desc_init(ocean_interaction_desc* desc)
{
    if (desc == nullptr) return;
    const ocean::InteractionDesc d{};
    std::memset(desc, 0, sizeof(*desc));
    desc->struct_size = sizeof(*desc);
    desc->size = d.size;
    desc->extent = d.extent;
    desc->kernel_radius = d.kernel_radius;
    desc->damping = d.damping;
    desc->absorb_cells = d.absorb_cells;
    desc->fixed_dt = d.fixed_dt;
    desc->max_substeps = d.max_substeps;
    desc->max_sources = d.max_sources;
    desc->gravity = d.gravity;
    desc->water_depth = d.depth;
    desc->obstruction = (d.obstruction == ocean::Obstruction::Dirichlet) ? OCEAN_OBSTRUCTION_DIRICHLET
                                                                        : OCEAN_OBSTRUCTION_NEUMANN;
    desc->thread_count = d.thread_count;
}

ocean_interaction* ocean_interaction_create(const ocean_interaction_desc* desc,
                                            ocean_status* out_status)
{
    auto fail = [&](ocean_status st) -> ocean_interaction* {
        if (out_status != nullptr) *out_status = st;
        return nullptr;
    };
    ocean_interaction_desc resolved;
    if (!copy_interaction_desc(desc, resolved)) return fail(OCEAN_ERROR_INVALID_ARG);
    try {
        auto* f = new ocean::InteractionField(translate(resolved));
        if (out_status != nullptr) *out_status = OCEAN_OK;
        return reinterpret_cast<ocean_interaction*>(f);
    } catch (const std::invalid_argument&) {
        return fail(OCEAN_ERROR_INVALID_ARG);
    } catch (const std::bad_alloc&) {
        return fail(OCEAN_ERROR_OUT_OF_MEMORY);
    } catch (...) {
        return fail(OCEAN_ERROR_UNKNOWN);
    }
}

void ocean_interaction_destroy(ocean_interaction* field) { delete to_field(field); }

void ocean_interaction_add(ocean_interaction* field, const ocean_disturbance* d)
{
    if (field == nullptr || d == nullptr) return;
    ocean::Disturbance s;
    s.world_x = d->world_x;
    s.world_z = d->world_z;
    s.radius = d->radius;
    s.strength = d->strength;
    s.velocity_y = d->velocity_y;
    s.velocity_x = d->velocity_x;
    s.velocity_z = d->velocity_z;
    s.kind = (d->kind == OCEAN_SOURCE_CONTINUOUS) ? ocean::SourceKind::Continuous
                                                  : ocean::SourceKind::Impulse;
    to_field(field)->add(s);
}

uint64_t ocean_interaction_dropped(const ocean_interaction* field)
{
    return field == nullptr ? 0u : to_cfield(field)->dropped_sources();
}

void ocean_interaction_update(ocean_interaction* field, float dt)
{
    if (field != nullptr) to_field(field)->update(dt);
}

int32_t ocean_interaction_recenter(ocean_interaction* field, float world_x, float world_z)
{
    if (field == nullptr) return 0;
    to_field(field)->recenter(world_x, world_z);
    return to_field(field)->recentered() ? 1 : 0;
}

void ocean_interaction_set_obstruction(ocean_interaction* field, const uint8_t* mask)
{
    if (field != nullptr) to_field(field)->set_obstruction(mask);
}

ocean_interaction_buffers ocean_interaction_get_buffers(const ocean_interaction* field)
{
    ocean_interaction_buffers b;
    std::memset(&b, 0, sizeof(b));
    if (field == nullptr) return b;
    const ocean::InteractionBuffers ib = to_cfield(field)->buffers();
    b.field = ib.field;
    b.size = ib.size;
    b.extent = ib.extent;
    b.origin_x = ib.origin_x;
    b.origin_z = ib.origin_z;
    return b;
}

void ocean_interaction_sample_at(const ocean_interaction* field, float world_x, float world_z,
                                 ocean_interaction_sample* out)
{
    if (out == nullptr) return;
    std::memset(out, 0, sizeof(*out));
    if (field == nullptr) return;
    const ocean::InteractionSample s = to_cfield(field)->sample_at(world_x, world_z);
    out->height = s.height;
    out->slope_x = s.slope_x;
    out->slope_z = s.slope_z;
    out->velocity_y = s.velocity_y;
}

float ocean_interaction_height_at(const ocean_interaction* field, float world_x, float world_z)
{
    return field == nullptr ? 0.0f : to_cfield(field)->height_at(world_x, world_z);
}

float ocean_interaction_fixed_dt(const ocean_interaction* field)
{
    return field == nullptr ? 0.0f : to_cfield(field)->fixed_dt();
}

float ocean_interaction_dt_limit(const ocean_interaction* field)
{
    return field == nullptr ? 0.0f : to_cfield(field)->stable_dt_limit();
}

float ocean_interaction_dispersion_error(const ocean_interaction* field)
{
    return field == nullptr ? 0.0f : to_cfield(field)->kernel_dispersion_error();
}

// reference src/c_api.cpp:494-516
void ocean_sample_combined(const ocean_sim* sim, const ocean_interaction* field, float world_x,
                           float world_z, ocean_surface_ex* out)
{
    if (out == nullptr) return;
    if (field == nullptr) {
        ocean_sample_ex(sim, world_x, world_z, out);
        return;
    }
    if (sim == nullptr) {
        std::memset(out, 0, sizeof(*out));
        out->normal_y = 1.0f;
        const ocean::InteractionSample e = to_cfield(field)->sample_at(world_x, world_z);
        out->height = e.height;
        out->velocity_y = e.velocity_y;
        return;
    }
    const ocean::WaterSurface water{*to_cpp(sim), *to_cfield(field)};
    to_surface_ex(water.sample_at(world_x, world_z), out);
}

}  // extern "C"

#include "iwave_kernel_host.hpp"

namespace {

int build_stencil(uint32_t radius, double cell, double depth, float* taps, double* sym)
{
    if (taps == nullptr || radius < 1 || radius > 12 || !(cell > 0.0)) return -1;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return -1;
    oceanamd::IWaveStencil k;
    const oceanamd::StencilStatus st = oceanamd::build_iwave_stencil(radius, cell, depth, s, k);
    (void)hipStreamDestroy(s);
    if (st != oceanamd::StencilStatus::Ok) return -1;
    std::memcpy(taps, k.taps.data(), k.taps.size() * sizeof(float));
    if (sym != nullptr) {
        sym[0] = k.max_symbol;
        sym[1] = k.min_symbol;
        sym[2] = k.band_low;
        sym[3] = k.peak_relative_error;
    }
    return static_cast<int>(k.taps.size());
}

}  // namespace

extern "C" int ocean_amd_build_iwave_taps(uint32_t radius, double cell, double depth, float* taps,
                                          double* max_symbol)
{
    double sym[4];
    const int r = build_stencil(radius, cell, depth, taps, sym);
    if (r > 0 && max_symbol != nullptr) *max_symbol = sym[0];
    return r;
}

extern "C" int ocean_amd_build_iwave_stencil(uint32_t radius, double cell, double depth, float* taps,
                                             double* sym)
{
    return build_stencil(radius, cell, depth, taps, sym);
}

// ---- ocean_amd.h: converged queries and the learned proposal ---------------

#include "ocean/proposal.hpp"

namespace {

ocean::HeightProposal* to_model(ocean_amd_proposal* p) noexcept
{
    return reinterpret_cast<ocean::HeightProposal*>(p);
}
const ocean::HeightProposal* to_model(const ocean_amd_proposal* p) noexcept
{
    return reinterpret_cast<const ocean::HeightProposal*>(p);
}

}  // namespace

extern "C" {

ocean_amd_proposal* ocean_amd_proposal_create(const float* params, uint32_t max_queries,
                                              ocean_status* out_status)
{
    auto fail = [&](ocean_status s) -> ocean_amd_proposal* {
        if (out_status != nullptr) *out_status = s;
        return nullptr;
    };
    if (max_queries == 0) return fail(OCEAN_ERROR_INVALID_ARG);
    try {
        auto* p = new ocean::HeightProposal(params, max_queries);
        if (out_status != nullptr) *out_status = OCEAN_OK;
        return reinterpret_cast<ocean_amd_proposal*>(p);
    } catch (const std::invalid_argument&) {
        return fail(OCEAN_ERROR_INVALID_ARG);
    } catch (const std::bad_alloc&) {
        return fail(OCEAN_ERROR_OUT_OF_MEMORY);
    } catch (...) {
        return fail(OCEAN_ERROR_UNKNOWN);
    }
}

void ocean_amd_proposal_destroy(ocean_amd_proposal* model) { delete to_model(model); }

int ocean_amd_proposal_set_params(ocean_amd_proposal* model, const float* params)
{
    if (model == nullptr) return OCEAN_ERROR_INVALID_ARG;
    return to_model(model)->set_params(params) ? OCEAN_OK : OCEAN_ERROR_UNKNOWN;
}

int ocean_amd_proposal_get_params(const ocean_amd_proposal* model, float* out)
{
    if (model == nullptr || out == nullptr) return OCEAN_ERROR_INVALID_ARG;
    return to_model(model)->get_params(out) ? OCEAN_OK : OCEAN_ERROR_UNKNOWN;
}

int ocean_amd_proposal_forward(const ocean_amd_proposal* model, const float* feats_dev,
                               uint32_t rows, float* out_dev)
{
    if (model == nullptr) return OCEAN_ERROR_INVALID_ARG;
    if (rows == 0) return OCEAN_OK;
    if (feats_dev == nullptr || out_dev == nullptr) return OCEAN_ERROR_INVALID_ARG;
    return to_model(model)->forward(feats_dev, rows, out_dev) ? OCEAN_OK : OCEAN_ERROR_UNKNOWN;
}

namespace {
bool batch_ok(uint32_t batch) noexcept
{
    return batch != 0 && batch % 128u == 0 && batch <= OCEAN_AMD_PROPOSAL_MAX_BATCH;
}
}  // namespace

int ocean_amd_proposal_train_step(ocean_amd_proposal* model, const float* feats_dev,
                                  const float* targets_dev, const uint32_t* rows_dev, uint32_t batch,
                                  float lr, float* loss)
{
    if (model == nullptr || feats_dev == nullptr || targets_dev == nullptr || rows_dev == nullptr ||
        !batch_ok(batch) || !(lr > 0.0f))
        return OCEAN_ERROR_INVALID_ARG;
    return to_model(model)->train_step(feats_dev, targets_dev, rows_dev, batch, lr, loss)
               ? OCEAN_OK
               : OCEAN_ERROR_UNKNOWN;
}

int ocean_amd_proposal_gradient(ocean_amd_proposal* model, const float* feats_dev,
                                const float* targets_dev, const uint32_t* rows_dev, uint32_t batch,
                                float* grad, float* loss)
{
    if (model == nullptr || feats_dev == nullptr || targets_dev == nullptr || rows_dev == nullptr ||
        grad == nullptr || !batch_ok(batch))
        return OCEAN_ERROR_INVALID_ARG;
    return to_model(model)->gradient(feats_dev, targets_dev, rows_dev, batch, grad, loss)
               ? OCEAN_OK
               : OCEAN_ERROR_UNKNOWN;
}

int ocean_amd_proposal_synchronize(const ocean_amd_proposal* model)
{
    if (model == nullptr) return OCEAN_ERROR_INVALID_ARG;
    to_model(model)->synchronize();
    return to_model(model)->last_error() == 0 ? OCEAN_OK : OCEAN_ERROR_UNKNOWN;
}

int ocean_amd_batch_height_converged(const ocean_amd_batch* batch, uint32_t first,
                                     uint32_t level_count, const float* xz_dev, uint32_t count,
                                     float tol, const ocean_amd_proposal* model, float* heights_dev,
                                     uint8_t* iters_dev, uint8_t* flags_dev)
{
    if (!batch_range(batch, first, level_count) || !(tol > 0.0f)) return OCEAN_ERROR_INVALID_ARG;
    if (count == 0) return OCEAN_OK;
    if (xz_dev == nullptr || heights_dev == nullptr) return OCEAN_ERROR_INVALID_ARG;
    if (model != nullptr && (level_count > 3u || count > to_model(model)->max_queries()))
        return OCEAN_ERROR_INVALID_ARG;
    return ocean::height_converged(*to_batch(batch), first, level_count, xz_dev, count, tol,
                                   to_model(model), heights_dev, iters_dev, flags_dev)
               ? OCEAN_OK
               : OCEAN_ERROR_UNKNOWN;
}

int ocean_amd_batch_height_converged_static(const ocean_amd_batch* batch, uint32_t first,
                                            uint32_t level_count, const float* xz_dev,
                                            uint32_t count, float tol,
                                            const ocean_amd_proposal* model, float* heights_dev,
                                            uint8_t* iters_dev, uint8_t* flags_dev)
{
    if (!batch_range(batch, first, level_count) || !(tol > 0.0f)) return OCEAN_ERROR_INVALID_ARG;
    if (count == 0) return OCEAN_OK;
    if (xz_dev == nullptr || heights_dev == nullptr) return OCEAN_ERROR_INVALID_ARG;
    if (count > OCEAN_AMD_STATIC_MAX_QUERIES || (model != nullptr && level_count > 3u))
        return OCEAN_ERROR_INVALID_ARG;
    int error = 0;
    return ocean::height_converged_static(*to_batch(batch), first, level_count, xz_dev, count, tol,
                                          to_model(model), heights_dev, iters_dev, flags_dev, &error)
               ? OCEAN_OK
               : OCEAN_ERROR_UNKNOWN;
}

int ocean_amd_cascade_height_converged(const ocean_amd_cascade* stack, const float* xz_dev,
                                       uint32_t count, float tol, const ocean_amd_proposal* model,
                                       float* heights_dev, uint8_t* iters_dev, uint8_t* flags_dev)
{
    if (stack == nullptr || !(tol > 0.0f)) return OCEAN_ERROR_INVALID_ARG;
    if (count == 0) return OCEAN_OK;
    if (xz_dev == nullptr || heights_dev == nullptr) return OCEAN_ERROR_INVALID_ARG;
    if (model != nullptr &&
        (to_stack(stack)->level_count() > 3u || count > to_model(model)->max_queries()))
        return OCEAN_ERROR_INVALID_ARG;
    return ocean::height_converged(*to_stack(stack), xz_dev, count, tol, to_model(model), heights_dev,
                                   iters_dev, flags_dev)
               ? OCEAN_OK
               : OCEAN_ERROR_UNKNOWN;
}

int ocean_amd_proposal_starts(const ocean_amd_batch* batch, uint32_t first, const float* xz_dev,
                              uint32_t count, const ocean_amd_proposal* model, float* starts_dev,
                              int32_t* spacing_dev)
{
    if (!batch_range(batch, first, 3u) || model == nullptr) return OCEAN_ERROR_INVALID_ARG;
    if (count == 0) return OCEAN_OK;
    if (xz_dev == nullptr || starts_dev == nullptr || count > to_model(model)->max_queries())
        return OCEAN_ERROR_INVALID_ARG;
    return ocean::proposal_starts(*to_batch(batch), first, xz_dev, count, *to_model(model), starts_dev,
                                  spacing_dev)
               ? OCEAN_OK
               : OCEAN_ERROR_UNKNOWN;
}

int ocean_amd_proposal_pairs(const ocean_amd_batch* batch, uint32_t first, const float* xz_dev,
                             uint32_t count, float tol, float* feats_dev, float* targets_dev,
                             uint8_t* ok_dev, int32_t* spacing_dev)
{
    if (!batch_range(batch, first, 3u) || !(tol > 0.0f)) return OCEAN_ERROR_INVALID_ARG;
    if (count == 0) return OCEAN_OK;
    if (xz_dev == nullptr || feats_dev == nullptr || targets_dev == nullptr || ok_dev == nullptr ||
        spacing_dev == nullptr)
        return OCEAN_ERROR_INVALID_ARG;
    return ocean::proposal_pairs(*to_batch(batch), first, xz_dev, count, tol, feats_dev, targets_dev,
                                 ok_dev, spacing_dev)
               ? OCEAN_OK
               : OCEAN_ERROR_UNKNOWN;
}

}  // extern "C"

// ---- renderer handoff --------------------------------------------------------

extern "C" {

int ocean_amd_pack_rgba16(const ocean_sim* sim, const ocean_amd_foam* foam,
                          uint16_t* displacement, uint16_t* normal)
{
    if (sim == nullptr || (displacement == nullptr && normal == nullptr))
        return OCEAN_ERROR_INVALID_ARG;
    if (to_cpp(sim)->pack_rgba16(to_foam(foam), displacement, normal, false)) return OCEAN_OK;
    return to_cpp(sim)->last_error() != 0 ? OCEAN_ERROR_UNKNOWN : OCEAN_ERROR_INVALID_ARG;
}

int ocean_amd_pack_rgba16_device(const ocean_sim* sim, const ocean_amd_foam* foam,
                                 uint16_t* displacement_dev, uint16_t* normal_dev)
{
    if (sim == nullptr || (displacement_dev == nullptr && normal_dev == nullptr))
        return OCEAN_ERROR_INVALID_ARG;
    if (to_cpp(sim)->pack_rgba16(to_foam(foam), displacement_dev, normal_dev, true)) return OCEAN_OK;
    return to_cpp(sim)->last_error() != 0 ? OCEAN_ERROR_UNKNOWN : OCEAN_ERROR_INVALID_ARG;
}

int ocean_amd_pack_half_device(const ocean_sim* sim, const float* src_dev, uint16_t* dst_dev,
                               uint64_t count)
{
    if (sim == nullptr || src_dev == nullptr || dst_dev == nullptr) return OCEAN_ERROR_INVALID_ARG;
    if (reinterpret_cast<std::uintptr_t>(src_dev) % 16 != 0 ||
        reinterpret_cast<std::uintptr_t>(dst_dev) % 8 != 0)
        return OCEAN_ERROR_INVALID_ARG;
    if (count == 0) return OCEAN_OK;
    auto& dev = to_cpp(sim)->device();
    dev.make_current();
    const hipError_t e = ocean_hip_pack_half(src_dev, dst_dev, count, dev.stream);
    if (e == hipErrorInvalidValue) return OCEAN_ERROR_INVALID_ARG;
    dev.note(e);
    return dev.status == 0 ? OCEAN_OK : OCEAN_ERROR_UNKNOWN;
}

int ocean_amd_interaction_pack_rgba16(const ocean_interaction* field, uint16_t* out)
{
    if (field == nullptr || out == nullptr) return OCEAN_ERROR_INVALID_ARG;
    if (to_cfield(field)->pack_rgba16(out, false)) return OCEAN_OK;
    return to_cfield(field)->last_error() != 0 ? OCEAN_ERROR_UNKNOWN : OCEAN_ERROR_INVALID_ARG;
}

double ocean_amd_interaction_energy(const ocean_interaction* field)
{
    if (field == nullptr) return std::numeric_limits<double>::quiet_NaN();
    return to_cfield(field)->energy();
}

uint32_t ocean_amd_interaction_obstruction_tables(const ocean_interaction* field, uint32_t* solid,
                                                  uint32_t* source, uint32_t capacity)
{
    if (field == nullptr) return 0;
    return to_cfield(field)->obstruction_tables(solid, source, capacity);
}

int ocean_amd_interaction_pack_rgba16_device(const ocean_interaction* field, uint16_t* out_dev)
{
    if (field == nullptr || out_dev == nullptr) return OCEAN_ERROR_INVALID_ARG;
    if (to_cfield(field)->pack_rgba16(out_dev, true)) return OCEAN_OK;
    return to_cfield(field)->last_error() != 0 ? OCEAN_ERROR_UNKNOWN : OCEAN_ERROR_INVALID_ARG;
}

}  // extern "C"

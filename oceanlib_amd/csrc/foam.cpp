// FoamField on the GPU: reference src/foam.cpp, with each fixed step one
// kernel (ocean_hip_foam_step) on the ocean's stream and the field resident in
// device memory.
#include "ocean/foam.hpp"

#include "ocean_device.hpp"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <utility>

namespace ocean {
namespace {

// reference src/foam.cpp:144-165
void validate(const FoamDesc& d)
{
    if (!(d.decay >= 0.0f)) throw std::invalid_argument("ocean: foam decay must be non-negative");
    if (!(d.source_gain >= 0.0f))
        throw std::invalid_argument("ocean: foam source_gain must be non-negative");
    if (!(d.fixed_dt > 0.0f)) throw std::invalid_argument("ocean: foam fixed_dt must be positive");
    if (d.max_substeps < 1) throw std::invalid_argument("ocean: foam max_substeps must be >= 1");
    if (!std::isfinite(d.advect_scale) || !std::isfinite(d.wind_drift_x) ||
        !std::isfinite(d.wind_drift_z))
        throw std::invalid_argument("ocean: foam drift parameters must be finite");
}

}  // namespace

struct FoamField::Impl {
    std::uint32_t plan(float dt, ocean_hip_foam_params& p) noexcept;

    const Ocean* ocean = nullptr;
    FoamDesc desc;
    std::uint32_t n = 0;
    std::size_t cells = 0;
    float cell = 0.0f;

    float* dev = nullptr;       // 2*cells floats: the ping-pong pair
    float* cur = nullptr;
    float* next = nullptr;
    double* scratch = nullptr;  // coverage result
    float* q_xz = nullptr;      // one-point query staging
    float* q_out = nullptr;
    float* host = nullptr;      // pinned mirror of `cur`
    double* host_cov = nullptr;
    mutable bool stale = true;

    double accum = 0.0;
    std::uint32_t substeps_last = 0;

    ~Impl()
    {
        if (ocean) hipStreamSynchronize(ocean->device().stream);
        if (dev) hipFree(dev);
        if (host) hipHostFree(host);
        if (host_cov) hipHostFree(host_cov);
    }
};

FoamField::FoamField(const Ocean& ocean, const FoamDesc& desc)
{
    validate(desc);
    const DeviceBuffers b = ocean.device_buffers();
    if (b.velocity == nullptr)
        throw std::invalid_argument(
            "ocean: FoamField requires OceanDesc::compute_velocity to be set");
    impl_ = std::make_unique<Impl>();
    Impl& m = *impl_;
    m.ocean = &ocean;
    m.desc = desc;
    m.n = b.size;
    m.cells = static_cast<std::size_t>(b.size) * b.size;
    m.cell = b.patch_length / static_cast<float>(b.size);

    auto& od = ocean.device();
    od.make_current();
    const std::size_t bytes = m.cells * 2 * sizeof(float) + 2 * sizeof(double) + 64;
    detail::hip_check(hipMalloc(reinterpret_cast<void**>(&m.dev), bytes), "hipMalloc foam");
    m.cur = m.dev;
    m.next = m.dev + m.cells;
    m.scratch = reinterpret_cast<double*>(m.dev + 2 * m.cells);
    m.q_xz = reinterpret_cast<float*>(m.scratch + 2);
    m.q_out = m.q_xz + 2;
    detail::hip_check(hipHostMalloc(reinterpret_cast<void**>(&m.host), m.cells * sizeof(float),
                                    hipHostMallocDefault),
                      "hipHostMalloc foam");
    detail::hip_check(hipHostMalloc(reinterpret_cast<void**>(&m.host_cov), 4 * sizeof(double),
                                    hipHostMallocDefault),
                      "hipHostMalloc foam");
    detail::hip_check(hipMemsetAsync(m.dev, 0, m.cells * 2 * sizeof(float), od.stream), "clear");
    detail::hip_check(hipStreamSynchronize(od.stream), "clear");
}

FoamField::~FoamField() = default;
FoamField::FoamField(FoamField&&) noexcept = default;
FoamField& FoamField::operator=(FoamField&&) noexcept = default;

void FoamField::update(float dt) noexcept
{
    Impl& m = *impl_;
    ocean_hip_foam_params p;
    const std::uint32_t steps = m.plan(dt, p);
    if (steps == 0) return;
    auto& od = m.ocean->device();
    for (std::uint32_t s = 0; s < steps; ++s) {
        od.note(ocean_hip_foam_step(&od.level, m.cur, m.next, &p, od.stream));
        std::swap(m.cur, m.next);
    }
    m.stale = true;
}

bool FoamField::update_and_sample(float dt, const float* xz_dev, std::uint32_t count,
                                  Surface* out_dev, float* heights_dev) noexcept
{
    if (count != 0 && (xz_dev == nullptr || (out_dev == nullptr && heights_dev == nullptr)))
        return false;
    Impl& m = *impl_;
    ocean_hip_foam_params p;
    const std::uint32_t steps = m.plan(dt, p);
    auto& od = m.ocean->device();
    if (steps == 0 || count == 0) {
        for (std::uint32_t s = 0; s < steps; ++s) {
            od.note(ocean_hip_foam_step(&od.level, m.cur, m.next, &p, od.stream));
            std::swap(m.cur, m.next);
        }
        if (steps != 0) m.stale = true;
        return count == 0 || m.ocean->sample_batch_device(xz_dev, count, out_dev, heights_dev);
    }
    for (std::uint32_t s = 0; s + 1 < steps; ++s) {
        od.note(ocean_hip_foam_step(&od.level, m.cur, m.next, &p, od.stream));
        std::swap(m.cur, m.next);
    }
    od.note(ocean_hip_foam_step_sample(&od.level, m.cur, m.next, &p, xz_dev, count,
                                       reinterpret_cast<ocean_hip_surface*>(out_dev), heights_dev,
                                       od.stream));
    std::swap(m.cur, m.next);
    m.stale = true;
    return od.status == 0;
}

// reference src/foam.cpp:218-252 (same accumulator arithmetic in double): the
// substeps `dt` asks for, with the accumulator advanced, and their constants.
std::uint32_t FoamField::Impl::plan(float dt, ocean_hip_foam_params& p) noexcept
{
    Impl& m = *this;
    if (!(dt > 0.0f)) dt = 0.0f;
    m.accum += static_cast<double>(dt);
    const double step = static_cast<double>(m.desc.fixed_dt);
    const double eps = step * 1e-6;
    std::uint32_t steps = 0;
    if (m.accum + eps >= step) steps = static_cast<std::uint32_t>((m.accum + eps) / step);
    if (steps > m.desc.max_substeps) {
        steps = m.desc.max_substeps;
        m.accum = 0.0;
    } else {
        m.accum -= static_cast<double>(steps) * step;
        if (m.accum < 0.0) m.accum = 0.0;
    }
    m.substeps_last = steps;
    if (steps == 0) return 0;

    auto& od = m.ocean->device();
    if (od.level.vel == nullptr) return 0;
    od.make_current();
    // Per-step constants, computed on the host exactly as foam.cpp:94-100.
    const float fdt = m.desc.fixed_dt;
    p.dt = fdt;
    p.inv_cell = 1.0f / m.cell;
    p.decay = std::exp(-m.desc.decay * fdt);
    p.gain = m.desc.source_gain * fdt;
    p.scale = m.desc.advect_scale;
    p.drift_x = m.desc.wind_drift_x;
    p.drift_z = m.desc.wind_drift_z;
    return steps;
}

void FoamField::clear() noexcept
{
    Impl& m = *impl_;
    auto& od = m.ocean->device();
    od.make_current();
    od.note(hipMemsetAsync(m.dev, 0, m.cells * 2 * sizeof(float), od.stream));
    m.accum = 0.0;
    m.stale = true;
}

const float* FoamField::data() const noexcept
{
    Impl& m = *impl_;
    auto& od = m.ocean->device();
    std::lock_guard<std::mutex> lock(od.query_mu);  // shares the ocean's query staging and stream
    if (m.stale) {
        od.make_current();
        od.note(hipMemcpyAsync(m.host, m.cur, m.cells * sizeof(float), hipMemcpyDeviceToHost,
                               od.stream));
        od.note(hipStreamSynchronize(od.stream));
        m.stale = false;
    }
    return m.host;
}

const float* FoamField::device_data() const noexcept { return impl_->cur; }
const Ocean& FoamField::ocean() const noexcept { return *impl_->ocean; }
std::uint32_t FoamField::size() const noexcept { return impl_->n; }
const FoamDesc& FoamField::desc() const noexcept { return impl_->desc; }
std::uint32_t FoamField::last_substeps() const noexcept { return impl_->substeps_last; }

// reference src/foam.cpp:262-285: timestep, substep clamp and scheduler are
// construction-time choices and survive a retune.
void FoamField::set_desc(const FoamDesc& d)
{
    validate(d);
    Impl& m = *impl_;
    const FoamDesc keep = m.desc;
    m.desc = d;
    m.desc.fixed_dt = keep.fixed_dt;
    m.desc.max_substeps = keep.max_substeps;
    m.desc.parallel_for = keep.parallel_for;
    m.desc.parallel_for_user = keep.parallel_for_user;
    m.desc.thread_count = keep.thread_count;
}

float FoamField::foam_at(float world_x, float world_z) const noexcept
{
    const Impl& m = *impl_;
    auto& od = m.ocean->device();
    std::lock_guard<std::mutex> lock(od.query_mu);
    od.make_current();
    // The point and the result go through the field's own pinned words
    // (host_cov[2], host_cov[3]): the ocean's host staging does not exist on
    // the levels of a cascade or batch, which allocate it only on demand.
    float* hx = reinterpret_cast<float*>(m.host_cov + 2);
    hx[0] = world_x;
    hx[1] = world_z;
    od.note(hipMemcpyAsync(m.q_xz, hx, 8, hipMemcpyHostToDevice, od.stream));
    od.note(ocean_hip_foam_query(&od.level, m.cur, m.cell, m.q_xz, 1, m.q_out, od.stream));
    float r = 0.0f;
    od.note(hipMemcpyAsync(m.host_cov + 3, m.q_out, 4, hipMemcpyDeviceToHost, od.stream));
    od.note(hipStreamSynchronize(od.stream));
    std::memcpy(&r, m.host_cov + 3, 4);
    return r;
}

double FoamField::coverage() const noexcept
{
    const Impl& m = *impl_;
    auto& od = m.ocean->device();
    std::lock_guard<std::mutex> lock(od.query_mu);
    od.make_current();
    double* out = m.scratch;
    od.note(ocean_hip_coverage(m.cur, static_cast<std::uint32_t>(m.cells), out,
                               od.stream));
    od.note(hipMemcpyAsync(m.host_cov, out, sizeof(double), hipMemcpyDeviceToHost, od.stream));
    od.note(hipStreamSynchronize(od.stream));
    return m.host_cov[0];
}

}  // namespace ocean

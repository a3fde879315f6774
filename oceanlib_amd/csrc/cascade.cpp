// CascadeStack on the GPU: reference src/cascade.cpp. The levels are an
// OceanBatch: update() steps all of them with one graph launch (one row and
// one column launch per group of equal N, the groups overlapping) instead of
// the reference's level-by-level loop, with the same bits per level; a query
// is one kernel over all levels (ocean_hip_cascade_sample), also for a single
// level, whose foam and normal then pass through the composite's own
// arithmetic as in the reference.
#include "ocean/cascade.hpp"

#include "ocean/batch.hpp"
#include "ocean_device.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <vector>

namespace ocean {

namespace {
OceanBatch make_levels(std::span<const CascadeLevel> levels)
{
    if (levels.empty())
        throw std::invalid_argument("CascadeStack: at least one level is required");
    return OceanBatch(levels);
}
}  // namespace

struct CascadeStack::Impl {
    OceanBatch batch;
    double time = 0.0;
    void* q_stage = nullptr;  // device query staging
    float* q_xz = nullptr;
    ocean_hip_surface* q_out = nullptr;
    float* h_xz = nullptr;
    ocean_hip_surface* h_out = nullptr;
    detail::MappedQueries mapped;  // a few points: no staged copies
    std::mutex query_mu;           // concurrent const queries share the staging

    explicit Impl(std::span<const CascadeLevel> levels) : batch(make_levels(levels)) {}
    ~Impl()
    {
        batch.synchronize();
        if (q_stage) hipFree(q_stage);
        if (h_xz) hipHostFree(h_xz);
        if (h_out) hipHostFree(h_out);
    }
    const ocean_hip_level* views() const noexcept
    {
        return static_cast<const ocean_hip_level*>(batch.device_levels());
    }
    hipStream_t stream() const noexcept { return static_cast<hipStream_t>(batch.stream()); }
    detail::OceanDevice& first() const noexcept { return batch.level(0).device(); }
};

CascadeStack::CascadeStack(std::span<const CascadeLevel> levels)
{
    impl_ = std::make_unique<Impl>(levels);
    Impl& m = *impl_;
    const std::size_t q = detail::kQueryStage;
    detail::hip_check(hipMalloc(&m.q_stage, q * 8 + q * sizeof(ocean_hip_surface)),
                      "hipMalloc cascade");
    m.q_xz = static_cast<float*>(m.q_stage);
    m.q_out = reinterpret_cast<ocean_hip_surface*>(m.q_xz + 2 * q);
    detail::hip_check(hipHostMalloc(reinterpret_cast<void**>(&m.h_xz), q * 8, hipHostMallocDefault),
                      "hipHostMalloc");
    detail::hip_check(hipHostMalloc(reinterpret_cast<void**>(&m.h_out),
                                    q * sizeof(ocean_hip_surface), hipHostMallocDefault),
                      "hipHostMalloc");
}

CascadeStack::CascadeStack(std::initializer_list<CascadeLevel> levels)
    : CascadeStack(std::span<const CascadeLevel>(levels.begin(), levels.size()))
{
}

CascadeStack::~CascadeStack() = default;
CascadeStack::CascadeStack(CascadeStack&&) noexcept = default;
CascadeStack& CascadeStack::operator=(CascadeStack&&) noexcept = default;

// reference src/cascade.cpp:42-55: every level to the same absolute time.
void CascadeStack::update(double time) noexcept
{
    impl_->batch.update(time);
    impl_->time = time;
}

std::size_t CascadeStack::level_count() const noexcept { return impl_->batch.level_count(); }
Buffers CascadeStack::buffers(std::size_t level) const noexcept
{
    return impl_->batch.level(level).buffers();
}
const OceanDesc& CascadeStack::desc(std::size_t level) const noexcept
{
    return impl_->batch.level(level).desc();
}
const Ocean& CascadeStack::level(std::size_t level) const noexcept
{
    return impl_->batch.level(level);
}
double CascadeStack::time() const noexcept { return impl_->time; }
void* CascadeStack::stream() const noexcept { return impl_->batch.stream(); }
const void* CascadeStack::device_levels() const noexcept { return impl_->batch.device_levels(); }

float CascadeStack::height_at(float world_x, float world_z) const noexcept
{
    return sample_at(world_x, world_z).height;
}

Surface CascadeStack::sample_at(float world_x, float world_z) const noexcept
{
    const float xz[2] = {world_x, world_z};
    Surface s;
    if (!sample_batch(xz, 1, &s)) return Surface{};
    return s;
}

bool CascadeStack::sample_batch(const float* xz, std::uint32_t count, Surface* out) const noexcept
{
    static_assert(sizeof(Surface) == sizeof(ocean_hip_surface));
    if (count == 0) return true;
    if (xz == nullptr || out == nullptr) return false;
    Impl& m = *impl_;
    std::lock_guard<std::mutex> lock(m.query_mu);
    auto& od = m.first();
    od.make_current();
    const auto levels = static_cast<std::uint32_t>(m.batch.level_count());
    if (count <= detail::kMappedQueries && m.mapped.ready()) {
        std::memcpy(m.mapped.xz_host(), xz, std::size_t(count) * 8);
        od.note(ocean_hip_cascade_sample(m.views(), levels, m.mapped.xz_dev(), count,
                                         m.mapped.out_dev<ocean_hip_surface>(), nullptr, m.stream()));
        od.note(hipStreamSynchronize(m.stream()));
        std::memcpy(out, m.mapped.out_host(), std::size_t(count) * sizeof(ocean_hip_surface));
        return od.status == 0;
    }
    for (std::uint32_t done = 0; done < count;) {
        const std::uint32_t k = std::min(detail::kQueryStage, count - done);
        std::memcpy(m.h_xz, xz + 2 * std::size_t(done), std::size_t(k) * 8);
        od.note(hipMemcpyAsync(m.q_xz, m.h_xz, std::size_t(k) * 8, hipMemcpyHostToDevice,
                               m.stream()));
        od.note(ocean_hip_cascade_sample(m.views(), levels, m.q_xz, k, m.q_out, nullptr,
                                         m.stream()));
        od.note(hipMemcpyAsync(m.h_out, m.q_out, std::size_t(k) * sizeof(ocean_hip_surface),
                               hipMemcpyDeviceToHost, m.stream()));
        od.note(hipStreamSynchronize(m.stream()));
        std::memcpy(out + done, m.h_out, std::size_t(k) * sizeof(ocean_hip_surface));
        done += k;
    }
    return od.status == 0;
}

bool CascadeStack::sample_batch_device(const float* xz_dev, std::uint32_t count, Surface* out_dev,
                                       float* heights_dev) const noexcept
{
    if (count == 0) return true;
    if (xz_dev == nullptr || (out_dev == nullptr && heights_dev == nullptr)) return false;
    Impl& m = *impl_;
    auto& od = m.first();
    od.make_current();
    od.note(ocean_hip_cascade_sample(m.views(), static_cast<std::uint32_t>(m.batch.level_count()),
                                     xz_dev, count, reinterpret_cast<ocean_hip_surface*>(out_dev),
                                     heights_dev, m.stream()));
    return od.status == 0;
}

}  // namespace ocean

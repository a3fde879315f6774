// CascadeStack on the GPU: reference src/cascade.cpp. Levels are ordinary
// Oceans sharing the stack's stream, so update() enqueues every level's
// frame back to back with no host synchronisation, and a query is one kernel
// over all levels (ocean_hip_cascade_sample).
#include "ocean/cascade.hpp"

#include "ocean_device.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace ocean {

struct CascadeStack::Impl {
    // Declared first so it is destroyed last: the levels synchronise on it.
    struct Stream {
        hipStream_t s = nullptr;
        ~Stream()
        {
            if (s) hipStreamDestroy(s);
        }
    } stream;
    std::vector<Ocean> levels;
    double time = 0.0;
    ocean_hip_level* dev_levels = nullptr;  // device copy of every level's view
    float* q_xz = nullptr;
    ocean_hip_surface* q_out = nullptr;
    float* h_xz = nullptr;
    ocean_hip_surface* h_out = nullptr;

    ~Impl()
    {
        if (stream.s) hipStreamSynchronize(stream.s);
        if (dev_levels) hipFree(dev_levels);
        if (h_xz) hipHostFree(h_xz);
        if (h_out) hipHostFree(h_out);
    }
};

CascadeStack::CascadeStack(std::span<const CascadeLevel> levels)
{
    if (levels.empty())
        throw std::invalid_argument("CascadeStack: at least one level is required");
    impl_ = std::make_unique<Impl>();
    Impl& m = *impl_;
    detail::hip_check(hipStreamCreateWithFlags(&m.stream.s, hipStreamNonBlocking),
                      "hipStreamCreate");
    m.levels.reserve(levels.size());
    for (const CascadeLevel& d : levels) m.levels.push_back(Ocean(d, m.stream.s, false));

    const std::size_t lv_bytes = (levels.size() * sizeof(ocean_hip_level) + 255) / 256 * 256;
    const std::size_t q = detail::kQueryStage;
    detail::hip_check(hipMalloc(reinterpret_cast<void**>(&m.dev_levels),
                                lv_bytes + q * 8 + q * sizeof(ocean_hip_surface)),
                      "hipMalloc cascade");
    m.q_xz = reinterpret_cast<float*>(reinterpret_cast<char*>(m.dev_levels) + lv_bytes);
    m.q_out = reinterpret_cast<ocean_hip_surface*>(m.q_xz + 2 * q);
    detail::hip_check(hipHostMalloc(reinterpret_cast<void**>(&m.h_xz), q * 8, hipHostMallocDefault),
                      "hipHostMalloc");
    detail::hip_check(hipHostMalloc(reinterpret_cast<void**>(&m.h_out),
                                    q * sizeof(ocean_hip_surface), hipHostMallocDefault),
                      "hipHostMalloc");
    std::vector<ocean_hip_level> views;
    for (const Ocean& o : m.levels) views.push_back(o.device().level);
    detail::hip_check(hipMemcpy(m.dev_levels, views.data(), views.size() * sizeof(ocean_hip_level),
                                hipMemcpyHostToDevice),
                      "upload cascade levels");
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
    for (Ocean& o : impl_->levels) o.update(time);
    impl_->time = time;
}

std::size_t CascadeStack::level_count() const noexcept { return impl_->levels.size(); }
Buffers CascadeStack::buffers(std::size_t level) const noexcept
{
    return impl_->levels[level].buffers();
}
const OceanDesc& CascadeStack::desc(std::size_t level) const noexcept
{
    return impl_->levels[level].desc();
}
const Ocean& CascadeStack::level(std::size_t level) const noexcept { return impl_->levels[level]; }
double CascadeStack::time() const noexcept { return impl_->time; }
void* CascadeStack::stream() const noexcept { return impl_->stream.s; }
const void* CascadeStack::device_levels() const noexcept { return impl_->dev_levels; }

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
    auto& od = m.levels.front().device();
    od.make_current();
    for (std::uint32_t done = 0; done < count;) {
        const std::uint32_t k = std::min(detail::kQueryStage, count - done);
        std::memcpy(m.h_xz, xz + 2 * std::size_t(done), std::size_t(k) * 8);
        od.note(hipMemcpyAsync(m.q_xz, m.h_xz, std::size_t(k) * 8, hipMemcpyHostToDevice,
                               m.stream.s));
        od.note(ocean_hip_cascade_sample(m.dev_levels,
                                         static_cast<std::uint32_t>(m.levels.size()), m.q_xz, k,
                                         m.q_out, nullptr, m.stream.s));
        od.note(hipMemcpyAsync(m.h_out, m.q_out, std::size_t(k) * sizeof(ocean_hip_surface),
                               hipMemcpyDeviceToHost, m.stream.s));
        od.note(hipStreamSynchronize(m.stream.s));
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
    auto& od = m.levels.front().device();
    od.make_current();
    od.note(ocean_hip_cascade_sample(m.dev_levels, static_cast<std::uint32_t>(m.levels.size()),
                                     xz_dev, count, reinterpret_cast<ocean_hip_surface*>(out_dev),
                                     heights_dev, m.stream.s));
    return od.status == 0;
}

}  // namespace ocean

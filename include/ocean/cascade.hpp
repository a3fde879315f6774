// Multi-scale cascades, MI355X implementation.
//
// Same class and member functions as the reference's
// include/ocean/cascade.hpp:30-92. All levels enqueue their frames on one
// shared stream; queries run one kernel over all levels and combine them in
// the reference's order (src/cascade.cpp:82-153).
#pragma once

#include "ocean/ocean.hpp"

#include <cstddef>
#include <initializer_list>
#include <memory>
#include <span>

namespace ocean {

using CascadeLevel = OceanDesc;

class CascadeStack {
public:
    explicit CascadeStack(std::span<const CascadeLevel> levels);
    CascadeStack(std::initializer_list<CascadeLevel> levels);
    ~CascadeStack();

    CascadeStack(CascadeStack&&) noexcept;
    CascadeStack& operator=(CascadeStack&&) noexcept;
    CascadeStack(const CascadeStack&) = delete;
    CascadeStack& operator=(const CascadeStack&) = delete;

    void update(double time) noexcept;
    [[nodiscard]] std::size_t level_count() const noexcept;
    [[nodiscard]] Buffers buffers(std::size_t level) const noexcept;
    [[nodiscard]] const OceanDesc& desc(std::size_t level) const noexcept;
    [[nodiscard]] const Ocean& level(std::size_t level) const noexcept;
    [[nodiscard]] float height_at(float world_x, float world_z) const noexcept;
    [[nodiscard]] Surface sample_at(float world_x, float world_z) const noexcept;
    [[nodiscard]] double time() const noexcept;

    // GPU additions: batched queries over host / device arrays.
    bool sample_batch(const float* xz, std::uint32_t count, Surface* out) const noexcept;
    bool sample_batch_device(const float* xz_dev, std::uint32_t count, Surface* out_dev,
                             float* heights_dev) const noexcept;
    [[nodiscard]] void* stream() const noexcept;
    // Device array of the levels' views (used by combined queries).
    [[nodiscard]] const void* device_levels() const noexcept;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace ocean

// Many independent ocean levels stepped together on one GPU (no reference
// counterpart: the reference updates levels one after another,
// src/cascade.cpp:42-55, each with its own four CPU stages).
//
// Levels with N in [64, 2048] are grouped by (N, velocity); every group is
// stepped by one persistent launch per pass, and the whole batch's frame is
// one instantiated hipGraph replayed by update(). Levels of other sizes are
// stepped by their own kernels right after the graph. Each level's buffers
// are bit-identical to those of a standalone Ocean with the same descriptor.
#pragma once

#include "ocean/ocean.hpp"

#include <cstddef>
#include <cstdint>
#include <memory>
#include <span>

namespace ocean {

struct BatchGroupInfo {
    std::uint32_t size = 0;         // N of the group's levels
    bool velocity = false;
    std::uint32_t levels = 0;
    std::uint32_t items = 0;        // work items per pass (N*N/2048 per level)
    std::uint32_t grid_row = 0;     // persistent grid of the row pass
    std::uint32_t grid_col = 0;
    float row_ms = 0.0f;            // device time measured by the last autotune
    float col_ms = 0.0f;
    std::uint32_t radix_row = 8;    // FFT points per lane of each pass (8, or 4:
    std::uint32_t radix_col = 8;    // stages in pairs over twice the lanes)
    std::uint32_t layout = 0;       // fields between the passes: 0 row-major,
                                    // 1 tiles as wide as an item's columns
};

class OceanBatch {
public:
    // Validates and builds every level as Ocean would; throws like Ocean.
    explicit OceanBatch(std::span<const OceanDesc> levels);
    ~OceanBatch();

    OceanBatch(OceanBatch&&) noexcept;
    OceanBatch& operator=(OceanBatch&&) noexcept;
    OceanBatch(const OceanBatch&) = delete;
    OceanBatch& operator=(const OceanBatch&) = delete;

    // Every level to the same absolute time: one graph launch. No allocation.
    void update(double time) noexcept;
    [[nodiscard]] double time() const noexcept;

    [[nodiscard]] std::size_t level_count() const noexcept;
    // The level as an Ocean, for buffers(), device_buffers() and queries.
    // Owned by the batch; update it only through the batch.
    [[nodiscard]] const Ocean& level(std::size_t i) const noexcept;

    // Queries over levels [first, first+count): one level answers as
    // Ocean::sample_at, several as a CascadeStack of those levels.
    bool sample(std::size_t first, std::size_t count, const float* xz, std::uint32_t n,
                Surface* out) const noexcept;
    bool sample_device(std::size_t first, std::size_t count, const float* xz_dev, std::uint32_t n,
                       Surface* out_dev, float* heights_dev) const noexcept;

    // Times candidate persistent grid sizes per group and pass (HIP events,
    // `reps` launches each), keeps the fastest and rebuilds the graph. The
    // choice changes only how work is spread, never the results.
    void autotune(std::uint32_t reps = 5) noexcept;
    // Opt-in fast mode of every level (values as Ocean::set_precision),
    // applied from the next update; the grid sizes are kept.
    bool set_precision(std::uint32_t mode) noexcept;
    [[nodiscard]] std::uint32_t precision() const noexcept;
    [[nodiscard]] std::size_t group_count() const noexcept;
    [[nodiscard]] BatchGroupInfo group(std::size_t g) const noexcept;
    // Sets group g's launch choices by hand (what autotune() picks): the
    // layout of the fields between the passes (0 row-major; 1 tiles as wide
    // as a work item's 2048/N columns, bit-exact mode, N = 256 .. 1024), the
    // FFT points per lane of each pass (8, or 4 for N <= 1024) and each
    // pass's grid, 1 .. its block units (the group's items, twice that with 4
    // points per lane). False, and nothing changed, for a choice that does
    // not exist.
    bool set_group_launch(std::size_t g, std::uint32_t layout, std::uint32_t radix_row,
                          std::uint32_t grid_row, std::uint32_t radix_col,
                          std::uint32_t grid_col) noexcept;

    // Checksums of levels [first, first+count) as they stand after the work
    // queued so far (blocks): per level, its buffers (displacement, normal,
    // velocity when on) as uint32 words w_i summed as sum(w_i * K_i) mod 2^64,
    // K_i = (i * 0x9E3779B97F4A7C15 + 0x632BE59BD9B4E019) | 1. Equal buffers
    // give equal sums on any machine; any single changed word changes it.
    bool checksums(std::size_t first, std::size_t count, std::uint64_t* out) const noexcept;

    void synchronize() const noexcept;
    [[nodiscard]] int last_error() const noexcept;
    [[nodiscard]] void* stream() const noexcept;
    // Device array of the levels' views (internal layout), for the query
    // kernels of this library.
    [[nodiscard]] const void* device_levels() const noexcept;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace ocean

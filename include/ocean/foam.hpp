// Persistent advected foam, MI355X implementation.
//
// Same class, descriptor fields and defaults as the reference's
// include/ocean/foam.hpp:38-142. The field lives in device memory as two
// ping-pong N*N float buffers; each fixed step is one HIP kernel enqueued on
// the ocean's stream, after the frame the ocean last enqueued.
#pragma once

#include "ocean/ocean.hpp"

#include <cstddef>
#include <cstdint>
#include <memory>

namespace ocean {

struct FoamDesc {
    float decay = 0.35f;
    float source_gain = 4.0f;
    float advect_scale = 1.0f;
    float wind_drift_x = 0.0f;
    float wind_drift_z = 0.0f;
    float fixed_dt = 1.0f / 60.0f;
    std::uint32_t max_substeps = 4;
    ParallelForFn parallel_for = nullptr;   // accepted, unused on the GPU path
    void* parallel_for_user = nullptr;
    std::uint32_t thread_count = 0;
};

class FoamField {
public:
    // Throws std::invalid_argument if the ocean has no velocity field or the
    // descriptor is malformed (reference src/foam.cpp:144-200).
    FoamField(const Ocean& ocean, const FoamDesc& desc);
    ~FoamField();

    FoamField(FoamField&&) noexcept;
    FoamField& operator=(FoamField&&) noexcept;
    FoamField(const FoamField&) = delete;
    FoamField& operator=(const FoamField&) = delete;

    void update(float dt) noexcept;
    // update(dt) followed by the ocean's sample_batch_device on the same
    // frame, with the last foam substep and the queries in one launch (no
    // reference counterpart; the same results as the two calls).
    bool update_and_sample(float dt, const float* xz_dev, std::uint32_t count, Surface* out_dev,
                           float* heights_dev) noexcept;
    void clear() noexcept;

    // Host mirror of the N*N coverage, refreshed from the device on request.
    [[nodiscard]] const float* data() const noexcept;
    [[nodiscard]] const float* device_data() const noexcept;
    // The ocean this field follows (no reference counterpart).
    [[nodiscard]] const Ocean& ocean() const noexcept;
    [[nodiscard]] std::uint32_t size() const noexcept;
    [[nodiscard]] float foam_at(float world_x, float world_z) const noexcept;
    [[nodiscard]] const FoamDesc& desc() const noexcept;
    void set_desc(const FoamDesc& desc);
    [[nodiscard]] std::uint32_t last_substeps() const noexcept;
    // Mean coverage in [0,1]; a device reduction in double, equal to the
    // reference's sequential sum to rounding (not bit for bit).
    [[nodiscard]] double coverage() const noexcept;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace ocean

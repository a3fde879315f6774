// Local interaction field (iWave), MI355X implementation.
//
// Same types, fields, defaults and member functions as the reference's
// include/ocean/interaction.hpp:52-398. The two time levels, damping
// coefficients, stencil taps and RGBA output live in device memory; every
// substep is a handful of HIP kernels (source stamps, obstruction, the
// (2P+1)^2 convolution on LDS tiles) enqueued on the field's stream, and the
// host keeps only the reference's bookkeeping (accumulator, source queue,
// integer grid origin, the nearest-fluid table built at set_obstruction).
#pragma once

#include "ocean/cascade.hpp"
#include "ocean/ocean.hpp"

#include <cstddef>
#include <cstdint>
#include <memory>

namespace ocean {

enum class Obstruction {
    Neumann,
    Dirichlet,
};

// reference include/ocean/interaction.hpp:64-139
struct InteractionDesc {
    std::uint32_t size = 256;
    float extent = 64.0f;
    std::uint32_t kernel_radius = 6;
    float damping = 0.20f;
    std::uint32_t absorb_cells = 16;
    float fixed_dt = 0.0f;
    std::uint32_t max_substeps = 8;
    std::uint32_t max_sources = 64;
    float gravity = 9.81f;
    float depth = 0.0f;
    Obstruction obstruction = Obstruction::Neumann;
    ParallelForFn parallel_for = nullptr;  // accepted, unused on the GPU path
    void* parallel_for_user = nullptr;
    std::uint32_t thread_count = 0;
};

enum class SourceKind {
    Impulse,
    Continuous,
};

// reference include/ocean/interaction.hpp:161-199
struct Disturbance {
    float world_x = 0.0f;
    float world_z = 0.0f;
    float radius = 0.35f;
    float strength = 0.25f;
    float velocity_y = 0.0f;
    float velocity_x = 0.0f;
    float velocity_z = 0.0f;
    SourceKind kind = SourceKind::Impulse;
};

struct InteractionSample {
    float height = 0.0f;
    float slope_x = 0.0f;
    float slope_z = 0.0f;
    float velocity_y = 0.0f;
};

// Host mirror of the RGBA32F field (eta, dEta/dx, dEta/dz, dEta/dt),
// refreshed from the device when asked for. Not periodic.
struct InteractionBuffers {
    const float* field = nullptr;
    std::uint32_t size = 0;
    float extent = 0.0f;
    float origin_x = 0.0f;
    float origin_z = 0.0f;
};

namespace detail {
struct InteractionDevice;
}

class InteractionField {
public:
    explicit InteractionField(const InteractionDesc& desc);
    ~InteractionField();

    InteractionField(InteractionField&&) noexcept;
    InteractionField& operator=(InteractionField&&) noexcept;
    InteractionField(const InteractionField&) = delete;
    InteractionField& operator=(const InteractionField&) = delete;

    void add(const Disturbance& d) noexcept;
    [[nodiscard]] std::uint64_t dropped_sources() const noexcept;
    void update(float dt) noexcept;
    [[nodiscard]] std::uint32_t last_substeps() const noexcept;
    [[nodiscard]] bool last_update_clamped() const noexcept;
    void recenter(float world_x, float world_z) noexcept;
    [[nodiscard]] bool recentered() const noexcept;
    [[nodiscard]] std::int32_t last_shift_x() const noexcept;
    [[nodiscard]] std::int32_t last_shift_z() const noexcept;
    void set_obstruction(const std::uint8_t* mask) noexcept;
    [[nodiscard]] InteractionSample sample_at(float world_x, float world_z) const noexcept;
    [[nodiscard]] float height_at(float world_x, float world_z) const noexcept;
    [[nodiscard]] InteractionBuffers buffers() const noexcept;
    [[nodiscard]] const InteractionDesc& desc() const noexcept;
    [[nodiscard]] float fixed_dt() const noexcept;
    [[nodiscard]] float stable_dt_limit() const noexcept;
    [[nodiscard]] float kernel_dispersion_error() const noexcept;
    // Discrete energy of the scheme (reference src/interaction.cpp:803-846):
    // a device reduction in double, equal to the reference's to rounding.
    [[nodiscard]] double energy() const noexcept;

    // GPU additions.
    [[nodiscard]] const float* device_field() const noexcept;  // N*N float4
    bool sample_batch(const float* xz, std::uint32_t count, InteractionSample* out) const noexcept;
    // FFT surface (a device array of level views; one ocean, or a cascade's
    // levels combined as a cascade) plus this field, composed as WaterSurface.
    bool sample_combined_batch(const void* device_levels, std::uint32_t level_count, bool as_cascade,
                               const float* xz, std::uint32_t count, Surface* out) const noexcept;
    // The field's RGBA32F buffer as RGBA16F (the viewer's upload, reference
    // examples/viewer/ocean_view.cpp:864-870): 4 halves per cell into `out`
    // (8-byte aligned), a device array enqueued on the field's stream, or a
    // host array filled before returning.
    bool pack_rgba16(std::uint16_t* out, bool on_device) const noexcept;
    // The obstruction tables the field applies (reference
    // src/interaction.cpp:255-292): the solid cells in index order and each
    // one's nearest fluid cell (0xFFFFFFFF: none). Copies at most `capacity`
    // of each into host arrays and returns the number of solid cells (0
    // without a mask).
    std::uint32_t obstruction_tables(std::uint32_t* solid, std::uint32_t* source,
                                     std::uint32_t capacity) const noexcept;
    void synchronize() const noexcept;
    [[nodiscard]] int last_error() const noexcept;
    [[nodiscard]] detail::InteractionDevice& device() const noexcept;

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

// reference include/ocean/interaction.hpp:374-398: FFT surface (one ocean or
// a cascade) plus the interaction field, normals composed through slopes.
class WaterSurface {
public:
    WaterSurface(const Ocean& ocean, const InteractionField& field) noexcept;
    WaterSurface(const CascadeStack& stack, const InteractionField& field) noexcept;

    [[nodiscard]] Surface sample_at(float world_x, float world_z) const noexcept;
    [[nodiscard]] float height_at(float world_x, float world_z) const noexcept;
    [[nodiscard]] InteractionSample interaction_at(float world_x, float world_z) const noexcept;
    // Batched form over host arrays (GPU addition).
    bool sample_batch(const float* xz, std::uint32_t count, Surface* out) const noexcept;

private:
    const Ocean* ocean_ = nullptr;
    const CascadeStack* stack_ = nullptr;
    const InteractionField* field_ = nullptr;
};

}  // namespace ocean

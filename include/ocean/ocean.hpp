// oceanlib C++ API, MI355X implementation.
//
// Same types, fields, defaults and member functions as the reference's
// include/ocean/ocean.hpp:18-222, so C++ callers recompile against this
// header unchanged. The simulation runs on one GPU: update() enqueues HIP
// kernels on the ocean's own stream, and buffers() returns a host mirror that
// is refreshed from device memory when it is asked for.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>

namespace ocean {

// Host scheduler hook (reference ocean.hpp:17-31). Kept for source
// compatibility; the GPU path never calls it.
using TaskFn = void (*)(void* ctx, std::uint32_t index);
using ParallelForFn = void (*)(void* user, TaskFn task, void* ctx, std::uint32_t count);

// reference ocean.hpp:39-83
struct SpectrumDesc {
    float wind_speed = 10.0f;
    float fetch = 100000.0f;
    float wind_direction = 0.0f;
    float peak_enhancement = 3.3f;
    float swell = 0.0f;
    float small_wave_cutoff = 0.5f;
    float gravity = 9.81f;
    float depth = 0.0f;  // <= 0: deep water
};

// reference ocean.hpp:85-130
struct OceanDesc {
    std::uint32_t size = 256;
    float patch_length = 200.0f;
    SpectrumDesc spectrum{};
    float choppiness = 1.0f;
    float foam_threshold = 0.5f;
    std::uint64_t seed = 1337;
    ParallelForFn parallel_for = nullptr;
    void* parallel_for_user = nullptr;
    std::uint32_t thread_count = 0;
    bool compute_velocity = false;
};

inline constexpr std::size_t kBufferAlignment = 64;

// reference ocean.hpp:141-162: RGBA32F texels, row-major i = z * size + x.
struct Buffers {
    const float* displacement = nullptr;
    const float* normal = nullptr;
    const float* velocity = nullptr;
    std::uint32_t size = 0;
    float patch_length = 0.0f;
};

// reference ocean.hpp:164-180
struct Surface {
    float height = 0.0f;
    float offset_x = 0.0f;
    float offset_z = 0.0f;
    float normal_x = 0.0f;
    float normal_y = 1.0f;
    float normal_z = 0.0f;
    float foam = 0.0f;
    float velocity_x = 0.0f;
    float velocity_y = 0.0f;
    float velocity_z = 0.0f;
};

// Device-resident view of one ocean's outputs (no reference counterpart):
// N*N float4 each, valid until the Ocean is destroyed, contents current once
// work queued on `stream` before the call has completed.
struct DeviceBuffers {
    const float* displacement = nullptr;
    const float* normal = nullptr;
    const float* velocity = nullptr;
    std::uint32_t size = 0;
    float patch_length = 0.0f;
    void* stream = nullptr;  // hipStream_t the ocean enqueues on
};

namespace detail {
struct OceanDevice;
}
class CascadeStack;
class OceanBatch;

class FoamField;

class Ocean {
public:
    // Throws std::invalid_argument for a malformed desc (reference
    // src/ocean.cpp:28-59), std::bad_alloc when device or pinned memory runs
    // out, std::runtime_error when no usable GPU is present.
    explicit Ocean(const OceanDesc& desc);
    ~Ocean();

    Ocean(Ocean&&) noexcept;
    Ocean& operator=(Ocean&&) noexcept;
    Ocean(const Ocean&) = delete;
    Ocean& operator=(const Ocean&) = delete;

    // Absolute time; enqueues the frame and returns. No allocation.
    void update(double time) noexcept;

    // Host mirror of the last update; copies from the device when stale.
    [[nodiscard]] Buffers buffers() const noexcept;

    [[nodiscard]] float height_at(float world_x, float world_z) const noexcept;
    [[nodiscard]] Surface sample_at(float world_x, float world_z) const noexcept;

    [[nodiscard]] const OceanDesc& desc() const noexcept;
    [[nodiscard]] double time() const noexcept;

    // --- additions of the GPU build -------------------------------------
    [[nodiscard]] DeviceBuffers device_buffers() const noexcept;
    // Batched queries over host arrays: xz = count (x, z) pairs.
    bool sample_batch(const float* xz, std::uint32_t count, Surface* out) const noexcept;
    // Batched queries over device arrays, enqueued on the ocean's stream.
    bool sample_batch_device(const float* xz_dev, std::uint32_t count, Surface* out_dev,
                             float* heights_dev) const noexcept;
    // Renderer handoff (reference examples/viewer/half.hpp:58-130,
    // ocean_view.cpp:818-862): RGBA16F copies of the displacement (alpha =
    // foam's persistent coverage when foam is given, which must follow this
    // ocean; else its own w) and of the normal, 4 halves per cell, either
    // pointer may be null (8-byte aligned). on_device: device arrays, enqueued
    // on the ocean's stream; else host arrays, filled before returning.
    bool pack_rgba16(const FoamField* foam, std::uint16_t* displacement, std::uint16_t* normal,
                     bool on_device) const noexcept;
    // Zero-copy outputs: the following frames write the RGBA32F outputs into
    // caller-owned device memory (N*N float4 each, 16-byte aligned, e.g. a
    // renderer's buffers imported into HIP), and every reader of this ocean
    // reads them there; null keeps or restores the ocean's own buffer for
    // that output. Re-evaluates the current frame into them. False (nothing
    // changed) for a pointer HIP does not know as device memory of that size,
    // velocity without compute_velocity, or a batch's or cascade's level.
    bool bind_outputs(float* displacement, float* normal, float* velocity) noexcept;
    // Opt-in fast mode (no reference counterpart; the reference's frame is
    // only the bit-exact one): 0 = the reference's arithmetic bit for bit
    // (the default), 1 = hardware sine/cosine of the phase, fused butterfly
    // rotations and a hardware reciprocal square root for the normal,
    // 2 = mode 1 with the fields between the two passes stored as fp16.
    // Applies from the next update; the launch shape returns to the mode's
    // default. False for another value or a batch's or cascade's level.
    bool set_precision(std::uint32_t mode) noexcept;
    [[nodiscard]] std::uint32_t precision() const noexcept;
    void synchronize() const noexcept;
    // Sticky: the first HIP error seen since construction (0 = none).
    [[nodiscard]] int last_error() const noexcept;

    [[nodiscard]] detail::OceanDevice& device() const noexcept;

private:
    friend class CascadeStack;
    friend class OceanBatch;
    // A level of a CascadeStack or OceanBatch: enqueues on the owner's
    // stream; lazy_host defers the host mirrors until first requested.
    Ocean(const OceanDesc& desc, void* shared_stream, bool lazy_host);
    // The owner enqueued this level's frame for `time` itself.
    void note_frame(double time) noexcept;

    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace ocean

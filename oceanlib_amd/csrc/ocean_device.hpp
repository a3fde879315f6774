// Device state of one ocean level, shared by Ocean, FoamField, CascadeStack
// and the batched paths. Internal header.
#pragma once

#include "hip_launch.h"
#include "ocean/ocean.hpp"

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>

namespace ocean::detail {

// Raised inside constructors; mapped to OCEAN_ERROR_* at the C boundary.
inline void hip_check(hipError_t e, const char* what)
{
    if (e == hipSuccess) return;
    if (e == hipErrorOutOfMemory) throw std::bad_alloc();
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// The reference's descriptor checks (src/ocean.cpp:28-59); throws
// std::invalid_argument. Owners of several levels run it on every level
// before touching the GPU.
void validate_desc(const OceanDesc& desc);

// Largest batch one host-array query call stages at a time.
inline constexpr std::uint32_t kQueryStage = 65536;
inline constexpr std::uint32_t kMappedQueries = 64;

// Up to kMappedQueries host queries go through mapped pinned memory: the
// query kernel reads the points from it and writes the results into it, so a
// single query is one launch and one wait. Allocated on the first use;
// ready() stays false if the runtime cannot map it (then the callers use
// their staged copies). Points at the start, results from kOut on.
struct MappedQueries {
    static constexpr std::size_t kOut = std::size_t(kMappedQueries) * 8;
    static constexpr std::size_t kResultBytes = 128;  // largest result struct per point
    void* host = nullptr;
    char* dev = nullptr;
    bool tried = false;

    bool ready() noexcept
    {
        if (!tried) {
            tried = true;
            if (hipHostMalloc(&host, kOut + std::size_t(kMappedQueries) * kResultBytes, hipHostMallocMapped) ==
                hipSuccess) {
                void* d = nullptr;
                if (hipHostGetDevicePointer(&d, host, 0) == hipSuccess) dev = static_cast<char*>(d);
            } else {
                host = nullptr;
            }
            (void)hipGetLastError();  // a refused mapping leaves the staged path; clear it
        }
        return dev != nullptr;
    }
    float* xz_host() const noexcept { return static_cast<float*>(host); }
    const float* xz_dev() const noexcept { return reinterpret_cast<const float*>(dev); }
    template <typename T>
    T* out_dev() const noexcept
    {
        static_assert(sizeof(T) <= kResultBytes);
        return reinterpret_cast<T*>(dev + kOut);
    }
    const void* out_host() const noexcept { return static_cast<const char*>(host) + kOut; }
    ~MappedQueries()
    {
        if (host) hipHostFree(host);
    }
    MappedQueries() = default;
    MappedQueries(const MappedQueries&) = delete;
    MappedQueries& operator=(const MappedQueries&) = delete;
};

struct OceanDevice {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    ocean_hip_level level{};
    ocean_hip_level* d_level = nullptr;  // device copy of `level` (for multi-level kernels)
    int fields = 4;            // complex fields per frame: 4, or 6 with velocity

    void* arena = nullptr;     // one hipMalloc holds every device buffer
    // The arena's own output buffers; level.disp / normal / vel point here
    // unless the caller bound its own (Ocean::bind_outputs).
    float* own_disp = nullptr;
    float* own_normal = nullptr;
    float* own_vel = nullptr;
    void* q_stage = nullptr;   // lazily allocated query staging (batch levels)
    float* q_xz = nullptr;     // device query staging
    ocean_hip_surface* q_out = nullptr;

    void* pack_stage = nullptr;  // lazily allocated RGBA16F staging (host-array packs)

    float* h_disp = nullptr;   // pinned host mirrors (hipHostMalloc)
    float* h_normal = nullptr;
    float* h_vel = nullptr;
    float* h_q_xz = nullptr;
    ocean_hip_surface* h_q_out = nullptr;
    MappedQueries mapped;
    // Queries are const and may run concurrently on one handle (reference
    // contract); they share the staging above, so each holds this lock.
    std::mutex query_mu;

    // Optional timing of the frame kernels inside an update loop: the row and
    // column dispatches of every `timing_every`-th frame record start and end
    // events, in a ring of kTimingSlots samples (the oldest are overwritten).
    static constexpr std::uint32_t kTimingSlots = 64;
    std::uint32_t timing_every = 0;
    std::uint32_t timing_counter = 0;
    std::uint32_t timing_next = 0;
    std::uint32_t timing_used = 0;
    hipEvent_t timing_ev[kTimingSlots][4] = {};

    bool mirror_stale = true;
    std::atomic<int> status{0};  // first HIP error seen (sticky)
    std::uint32_t init_repaired = 0;  // table entries the host recomputed at creation

    OceanDevice() = default;
    OceanDevice(const OceanDevice&) = delete;
    OceanDevice& operator=(const OceanDevice&) = delete;
    ~OceanDevice();

    // Allocates and uploads; throws on failure (state is released by the dtor).
    // lazy_host: no pinned mirrors or query staging until first asked for
    // (the levels of a large batch).
    void create(const OceanDesc& desc, hipStream_t shared_stream, bool lazy_host = false);
    void alloc_host_mirrors() noexcept;

    void note(hipError_t e) noexcept
    {
        // first error wins; queries may note concurrently on one handle
        int none = 0;
        if (e != hipSuccess) status.compare_exchange_strong(none, static_cast<int>(e));
    }
    void make_current() const noexcept;
    void enqueue_frame(double time) noexcept;
    // every = 0 stops timing; otherwise creates the events (once) and starts.
    int set_frame_timing(std::uint32_t every) noexcept;
    // Waits for the stream, averages the recorded samples and clears them.
    int read_frame_timing(float* row_ms, float* col_ms, std::uint32_t* samples) noexcept;
    // Copies the device outputs into the pinned mirrors if they are stale.
    void refresh_mirror() noexcept;
    bool sample_host(const float* xz, std::uint32_t count, Surface* out) noexcept;
};

}  // namespace ocean::detail

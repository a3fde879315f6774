// InteractionField and WaterSurface on the GPU.
//
// The host keeps exactly the reference's bookkeeping (src/interaction.cpp):
// descriptor validation, the fixed-step accumulator and substep clamp, the
// source queue and the integer grid origin. Everything per cell runs on the
// device: stamps, obstruction, the convolution step, the RGBA finalize, the
// recentre shift and the queries (iwave_kernels.hip), and the set-up work
// (iwave_init_kernels.hip): the stencil fit's normal equations and symbol
// scan, and the obstruction mask with its nearest-fluid table, rebuilt on
// set_obstruction / recenter.
#include "ocean/interaction.hpp"

#include "iwave_kernel_host.hpp"
#include "ocean_device.hpp"

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <algorithm>
#include <mutex>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace ocean {
namespace detail {

struct InteractionDevice {
    int device = 0;
    hipStream_t stream = nullptr;
    std::atomic<int> status{0};
    void note(hipError_t e) noexcept
    {
        // first error wins; queries may note concurrently on one handle
        int none = 0;
        if (e != hipSuccess) status.compare_exchange_strong(none, static_cast<int>(e));
    }
    void make_current() const noexcept
    {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != device) hipSetDevice(device);
    }
};

}  // namespace detail

namespace {

constexpr bool is_power_of_two(std::uint32_t v) noexcept { return v != 0 && (v & (v - 1)) == 0; }

// reference src/interaction.cpp:438-475
void validate(const InteractionDesc& d)
{
    if (!is_power_of_two(d.size) || d.size < 32 || d.size > 1024)
        throw std::invalid_argument(
            "ocean: interaction size must be a power of two in [32, 1024], got " +
            std::to_string(d.size));
    if (!(d.extent > 0.0f)) throw std::invalid_argument("ocean: interaction extent must be positive");
    if (d.kernel_radius < 1 || d.kernel_radius > 12)
        throw std::invalid_argument("ocean: interaction kernel_radius must be in [1, 12]");
    if (2 * d.kernel_radius >= d.size)
        throw std::invalid_argument("ocean: interaction kernel_radius is too large for this size");
    if (!(d.damping >= 0.0f))
        throw std::invalid_argument("ocean: interaction damping must be non-negative");
    if (d.absorb_cells * 2 >= d.size)
        throw std::invalid_argument("ocean: interaction absorb_cells must be less than half the size");
    if (d.max_substeps < 1) throw std::invalid_argument("ocean: interaction max_substeps must be >= 1");
    if (d.max_sources < 1) throw std::invalid_argument("ocean: interaction max_sources must be >= 1");
    if (!(d.gravity > 0.0f)) throw std::invalid_argument("ocean: gravity must be positive");
    if (d.fixed_dt < 0.0f) throw std::invalid_argument("ocean: interaction fixed_dt must be non-negative");
}

std::size_t align16(std::size_t v) { return (v + 15u) & ~std::size_t(15u); }

}  // namespace

struct InteractionField::Impl {
    InteractionDesc desc;
    oceanamd::IWaveStencil kernel;
    std::uint32_t n = 0;
    std::size_t stride = 0, cells = 0, plane_floats = 0;
    float cell = 0.0f, dt = 0.0f, dt_limit = 0.0f;
    detail::InteractionDevice dev;

    // device
    void* arena = nullptr;
    float* plane[3] = {nullptr, nullptr, nullptr};  // whole allocations incl. halo
    int cur_i = 0, old_i = 1, spare_i = 2;
    float* taps = nullptr;
    float* c1 = nullptr;
    float* c2 = nullptr;
    float* out = nullptr;
    std::uint32_t* d_solid = nullptr;   // the solid cells in index order
    std::uint32_t* d_source = nullptr;  // per solid cell: its nearest fluid cell
    std::uint32_t* d_nearest = nullptr; // per cell: its nearest fluid cell
    std::uint32_t* d_mask_work = nullptr;
    std::uint32_t* d_solid_count = nullptr;
    std::uint8_t* d_mask[2] = {nullptr, nullptr};  // the mask, and recenter's shift target
    int mask_i = 0;
    ocean_hip_stamp* d_stamps = nullptr;
    detail::MappedQueries mapped;  // host queries of a few points: no staged copies
    std::mutex query_mu;           // concurrent const queries share the staging
    float* d_means = nullptr;
    float* d_qxz = nullptr;
    float* d_qout = nullptr;
    double* d_energy = nullptr;  // 2*cells scratch + 1 result
    void* d_pack = nullptr;      // RGBA16F staging for host-array packs (on first use)
    std::size_t stamp_cap = 0;

    // pinned host
    float* h_out = nullptr;
    // Two host staging buffers for the stamps, used in turn, so that an update
    // does not wait for the previous update's upload to finish.
    ocean_hip_stamp* h_stamps_buf[2] = {nullptr, nullptr};
    ocean_hip_stamp* h_stamps = nullptr;  // the one this update fills
    std::uint32_t stamp_turn = 0;
    std::vector<std::uint32_t> first;     // per substep, its first stamp (max_substeps + 1)
    float* h_qxz = nullptr;
    float* h_qout = nullptr;
    hipEvent_t stamps_free[2] = {nullptr, nullptr};  // the last upload from each buffer is done
    bool stamps_pending[2] = {false, false};          // an upload from the buffer was enqueued
    bool mirror_stale = true;

    // host bookkeeping (reference Impl members)
    // set_obstruction's uploads: two pinned buffers used in turn, mapped into
    // the device's address space when the runtime allows (the row kernel then
    // reads the mask straight from them), each reused only once the work that
    // reads it has completed.
    std::uint8_t* h_mask_stage[2] = {nullptr, nullptr};
    const std::uint8_t* d_mask_stage[2] = {nullptr, nullptr};  // null: not mapped
    hipEvent_t mask_read[2] = {nullptr, nullptr};
    bool mask_read_pending[2] = {false, false};
    int mask_turn = 0;
    bool has_mask = false;
    std::vector<Disturbance> queue;
    std::uint32_t queue_count = 0;
    std::uint64_t dropped = 0;
    double accum = 0.0;
    std::uint32_t substeps_last = 0;
    bool clamped_last = false;
    std::int32_t origin_x = 0, origin_z = 0, shift_x = 0, shift_z = 0;
    bool moved = false;

    ~Impl()
    {
        if (dev.stream) hipStreamSynchronize(dev.stream);
        if (arena) hipFree(arena);
        if (d_pack) hipFree(d_pack);
        if (h_out) hipHostFree(h_out);
        for (ocean_hip_stamp* b : h_stamps_buf)
            if (b) hipHostFree(b);
        if (h_qxz) hipHostFree(h_qxz);
        if (h_qout) hipHostFree(h_qout);
        for (hipEvent_t e : stamps_free)
            if (e) hipEventDestroy(e);
        for (std::uint8_t* b : h_mask_stage)
            if (b) hipHostFree(b);
        for (hipEvent_t e : mask_read)
            if (e) hipEventDestroy(e);
        if (dev.stream) hipStreamDestroy(dev.stream);
    }

    float* interior(int i) const noexcept
    {
        return plane[i] + static_cast<std::size_t>(desc.kernel_radius) * stride + desc.kernel_radius;
    }

    ocean_hip_iwave view() const noexcept
    {
        ocean_hip_iwave f{};
        f.n = n;
        f.p = desc.kernel_radius;
        f.stride = static_cast<std::uint32_t>(stride);
        f.cur = interior(cur_i);
        f.old = interior(old_i);
        f.taps = taps;
        f.c1 = c1;
        f.c2 = c2;
        f.out = out;
        f.gdt2 = desc.gravity * dt * dt;
        f.dt = dt;
        f.inv_two_cell = 0.5f / cell;
        f.inv_dt = 1.0f / dt;
        f.cell = cell;
        f.origin_x = origin_x;
        f.origin_z = origin_z;
        return f;
    }

    // reference src/interaction.cpp:212-241
    void build_damping(std::vector<float>& h1, std::vector<float>& h2) const
    {
        const float a0 = desc.damping;
        const int w = static_cast<int>(desc.absorb_cells);
        const float a_edge = 4.0f / std::max(dt * static_cast<float>(std::max(w, 1)), 1e-6f);
        for (std::uint32_t z = 0; z < n; ++z) {
            for (std::uint32_t x = 0; x < n; ++x) {
                const int dx = std::min<int>(static_cast<int>(x), static_cast<int>(n - 1 - x));
                const int dz = std::min<int>(static_cast<int>(z), static_cast<int>(n - 1 - z));
                const int d = std::min(dx, dz);
                float alpha = a0;
                if (w > 0 && d < w) {
                    const float t = static_cast<float>(w - d) / static_cast<float>(w);
                    alpha += a_edge * t * t;
                }
                const float ad = alpha * dt;
                const std::size_t i = static_cast<std::size_t>(z) * n + x;
                h1[i] = 1.0f / (1.0f + ad);
                h2[i] = 1.0f - ad;
            }
        }
    }

    // reference src/interaction.cpp:255-292: the solid cells and their
    // nearest fluid cells, from the device mask (iwave_init_kernels.hip gives
    // the closed form of the reference's BFS table that the kernels compute).
    void rebuild_mask_tables() noexcept
    {
        if (!has_mask) return;
        dev.note(ocean_hip_iwave_mask_tables(d_mask[mask_i], nullptr, n, d_mask_work, d_nearest,
                                             d_solid, d_source, d_solid_count, dev.stream));
    }

    // The scalar part of inject() (reference src/interaction.cpp:322-348);
    // false when the clipped box is empty (the reference returns early).
    bool make_stamp(const Disturbance& d, float sub_time, float scale, ocean_hip_stamp& st) const
    {
        const float sigma = std::max(d.radius, 0.25f * cell);
        st.inv2s2 = 1.0f / (2.0f * sigma * sigma);
        const float wx = d.world_x + d.velocity_x * sub_time;
        const float wz = d.world_z + d.velocity_z * sub_time;
        st.gx = (wx - static_cast<float>(origin_x) * cell) / cell;
        st.gz = (wz - static_cast<float>(origin_z) * cell) / cell;
        const int rad = static_cast<int>(std::ceil(4.0f * sigma / cell));
        st.x0 = std::max(0, static_cast<int>(std::floor(st.gx)) - rad);
        st.x1 = std::min(static_cast<int>(n) - 1, static_cast<int>(std::floor(st.gx)) + rad);
        st.z0 = std::max(0, static_cast<int>(std::floor(st.gz)) - rad);
        st.z1 = std::min(static_cast<int>(n) - 1, static_cast<int>(std::floor(st.gz)) + rad);
        st.disp = -d.strength * scale;
        st.vel = d.velocity_y * scale;
        st.pad_ = 0;
        return !(st.x1 < st.x0 || st.z1 < st.z0);
    }

    void finalize() noexcept
    {
        const ocean_hip_iwave f = view();
        dev.note(ocean_hip_iwave_finalize(&f, dev.stream));
        mirror_stale = true;
    }

    void shift_planes(std::int32_t dx, std::int32_t dz) noexcept
    {
        // cur -> spare, then old -> the freed buffer; halos stay zero.
        dev.note(ocean_hip_iwave_shift(interior(cur_i), interior(spare_i), n,
                                       static_cast<std::uint32_t>(stride), dx, dz, dev.stream));
        std::swap(cur_i, spare_i);
        dev.note(ocean_hip_iwave_shift(interior(old_i), interior(spare_i), n,
                                       static_cast<std::uint32_t>(stride), dx, dz, dev.stream));
        std::swap(old_i, spare_i);
    }
};

InteractionField::InteractionField(const InteractionDesc& d)
{
    validate(d);
    impl_ = std::make_unique<Impl>();
    Impl& m = *impl_;
    m.desc = d;
    m.n = d.size;
    m.cell = d.extent / static_cast<float>(d.size);
    m.cells = static_cast<std::size_t>(m.n) * m.n;
    const std::size_t p = d.kernel_radius;
    m.stride = align16(static_cast<std::size_t>(m.n) + 2 * p);
    m.plane_floats = (static_cast<std::size_t>(m.n) + 2 * p) * m.stride;

    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        throw std::runtime_error("ocean: no HIP device available");
    detail::hip_check(hipGetDevice(&m.dev.device), "hipGetDevice");
    detail::hip_check(hipStreamCreateWithFlags(&m.dev.stream, hipStreamNonBlocking), "hipStreamCreate");

    // reference src/interaction.cpp:158-193
    switch (oceanamd::build_iwave_stencil(d.kernel_radius, m.cell, d.depth, m.dev.stream, m.kernel)) {
        case oceanamd::StencilStatus::Ok: break;
        case oceanamd::StencilStatus::Singular:
            throw std::invalid_argument("ocean: iWave kernel fit is singular (degenerate basis)");
        default: throw std::runtime_error("ocean: iWave kernel fit failed on the device");
    }
    if (!oceanamd::stencil_is_stable(m.kernel))
        throw std::invalid_argument("ocean: interaction kernel_radius " +
                                    std::to_string(d.kernel_radius) +
                                    " yields a negative Fourier symbol (unconditionally unstable)");
    m.dt_limit = static_cast<float>(2.0 / std::sqrt(static_cast<double>(d.gravity) * m.kernel.max_symbol));
    m.dt = (d.fixed_dt > 0.0f) ? d.fixed_dt : std::min(1.0f / 60.0f, 0.35f * m.dt_limit);
    if (d.fixed_dt > 0.0f && d.fixed_dt > m.dt_limit)
        throw std::invalid_argument("ocean: interaction fixed_dt exceeds the stability limit of " +
                                    std::to_string(m.dt_limit) + " s for this configuration");
    for (hipEvent_t& e : m.stamps_free)
        detail::hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    m.first.assign(std::size_t(d.max_substeps) + 1, 0u);

    m.stamp_cap = static_cast<std::size_t>(d.max_substeps) * d.max_sources;
    const std::size_t w = 2 * p + 1;
    auto up = [](std::size_t v) { return (v + 255) / 256 * 256; };
    std::size_t off = 0;
    std::size_t o_plane[3];
    for (auto& o : o_plane) { o = off; off = up(off + m.plane_floats * 4); }
    const std::size_t o_taps = off;    off = up(off + w * w * 4);
    const std::size_t o_c1 = off;      off = up(off + m.cells * 4);
    const std::size_t o_c2 = off;      off = up(off + m.cells * 4);
    const std::size_t o_out = off;     off = up(off + m.cells * 16);
    const std::size_t o_solid = off;   off = up(off + m.cells * 4);
    const std::size_t o_source = off;  off = up(off + m.cells * 4);
    const std::size_t o_near = off;    off = up(off + m.cells * 4);
    const std::size_t o_mwork = off;   off = up(off + (m.cells + 2 * std::size_t(m.n)) * 4);
    const std::size_t o_scount = off;  off = up(off + 4);
    std::size_t o_mask[2];
    for (auto& o : o_mask) { o = off; off = up(off + m.cells); }
    const std::size_t o_stamps = off;  off = up(off + m.stamp_cap * sizeof(ocean_hip_stamp));
    const std::size_t o_means = off;   off = up(off + m.stamp_cap * 4);
    const std::size_t o_qxz = off;     off = up(off + std::size_t(detail::kQueryStage) * 8);
    const std::size_t o_qout = off;    off = up(off + std::size_t(detail::kQueryStage) * sizeof(ocean_hip_surface));
    const std::size_t o_energy = off;  off = up(off + (2 * m.cells + 1) * 8);
    detail::hip_check(hipMalloc(&m.arena, off), "hipMalloc interaction");
    char* base = static_cast<char*>(m.arena);
    for (int i = 0; i < 3; ++i) m.plane[i] = reinterpret_cast<float*>(base + o_plane[i]);
    m.taps = reinterpret_cast<float*>(base + o_taps);
    m.c1 = reinterpret_cast<float*>(base + o_c1);
    m.c2 = reinterpret_cast<float*>(base + o_c2);
    m.out = reinterpret_cast<float*>(base + o_out);
    m.d_solid = reinterpret_cast<std::uint32_t*>(base + o_solid);
    m.d_source = reinterpret_cast<std::uint32_t*>(base + o_source);
    m.d_nearest = reinterpret_cast<std::uint32_t*>(base + o_near);
    m.d_mask_work = reinterpret_cast<std::uint32_t*>(base + o_mwork);
    m.d_solid_count = reinterpret_cast<std::uint32_t*>(base + o_scount);
    for (int i = 0; i < 2; ++i) m.d_mask[i] = reinterpret_cast<std::uint8_t*>(base + o_mask[i]);
    m.d_stamps = reinterpret_cast<ocean_hip_stamp*>(base + o_stamps);
    m.d_means = reinterpret_cast<float*>(base + o_means);
    m.d_qxz = reinterpret_cast<float*>(base + o_qxz);
    m.d_qout = reinterpret_cast<float*>(base + o_qout);
    m.d_energy = reinterpret_cast<double*>(base + o_energy);

    detail::hip_check(hipHostMalloc(reinterpret_cast<void**>(&m.h_out), m.cells * 16, hipHostMallocDefault), "hipHostMalloc");
    for (ocean_hip_stamp*& b : m.h_stamps_buf)
        detail::hip_check(hipHostMalloc(reinterpret_cast<void**>(&b), m.stamp_cap * sizeof(ocean_hip_stamp), hipHostMallocDefault), "hipHostMalloc");
    m.h_stamps = m.h_stamps_buf[0];
    detail::hip_check(hipHostMalloc(reinterpret_cast<void**>(&m.h_qxz), std::size_t(detail::kQueryStage) * 8, hipHostMallocDefault), "hipHostMalloc");
    detail::hip_check(hipHostMalloc(reinterpret_cast<void**>(&m.h_qout), std::size_t(detail::kQueryStage) * sizeof(ocean_hip_surface), hipHostMallocDefault), "hipHostMalloc");

    for (int i = 0; i < 2; ++i) {
        if (hipHostMalloc(reinterpret_cast<void**>(&m.h_mask_stage[i]), m.cells, hipHostMallocMapped) ==
            hipSuccess) {
            void* dp = nullptr;
            if (hipHostGetDevicePointer(&dp, m.h_mask_stage[i], 0) == hipSuccess)
                m.d_mask_stage[i] = static_cast<const std::uint8_t*>(dp);
        } else {
            m.h_mask_stage[i] = nullptr;
        }
        (void)hipGetLastError();  // a refused mapping leaves the copied upload; clear it
        if (m.h_mask_stage[i] == nullptr)
            detail::hip_check(hipHostMalloc(reinterpret_cast<void**>(&m.h_mask_stage[i]), m.cells,
                                            hipHostMallocDefault), "hipHostMalloc");
        detail::hip_check(hipEventCreateWithFlags(&m.mask_read[i], hipEventDisableTiming), "hipEventCreate");
    }
    m.queue.resize(d.max_sources);

    std::vector<float> h1(m.cells), h2(m.cells);
    m.build_damping(h1, h2);
    hipStream_t s = m.dev.stream;
    detail::hip_check(hipMemsetAsync(m.plane[0], 0, 3 * (o_plane[1] - o_plane[0]), s), "clear planes");
    detail::hip_check(hipMemsetAsync(m.d_mask[0], 0, m.cells, s), "clear mask");
    detail::hip_check(hipMemsetAsync(m.d_solid_count, 0, 4, s), "clear solid count");
    detail::hip_check(hipMemcpyAsync(m.taps, m.kernel.taps.data(), w * w * 4, hipMemcpyHostToDevice, s), "upload taps");
    detail::hip_check(hipMemcpyAsync(m.c1, h1.data(), m.cells * 4, hipMemcpyHostToDevice, s), "upload c1");
    detail::hip_check(hipMemcpyAsync(m.c2, h2.data(), m.cells * 4, hipMemcpyHostToDevice, s), "upload c2");
    m.finalize();  // a valid all-zero output before the first update (reference 507)
    detail::hip_check(hipStreamSynchronize(s), "interaction init");
    if (m.dev.status != 0) throw std::runtime_error("ocean: interaction init failed");
}

InteractionField::~InteractionField() = default;
InteractionField::InteractionField(InteractionField&&) noexcept = default;
InteractionField& InteractionField::operator=(InteractionField&&) noexcept = default;

// reference src/interaction.cpp:511-524
void InteractionField::add(const Disturbance& d) noexcept
{
    Impl& m = *impl_;
    if (m.queue_count >= m.desc.max_sources) {
        ++m.dropped;
        return;
    }
    m.queue[m.queue_count++] = d;
}

std::uint64_t InteractionField::dropped_sources() const noexcept { return impl_->dropped; }

// reference src/interaction.cpp:531-587
void InteractionField::update(float dt) noexcept
{
    Impl& m = *impl_;
    m.moved = false;
    if (!(dt > 0.0f)) dt = 0.0f;
    m.accum += static_cast<double>(dt);
    std::uint32_t steps = 0;
    if (m.accum >= m.dt) steps = static_cast<std::uint32_t>(m.accum / static_cast<double>(m.dt));
    m.clamped_last = steps > m.desc.max_substeps;
    if (m.clamped_last) {
        steps = m.desc.max_substeps;
        m.accum = 0.0;
    } else {
        m.accum -= static_cast<double>(steps) * static_cast<double>(m.dt);
    }
    m.substeps_last = steps;
    if (steps == 0) {
        std::uint32_t keep = 0;
        for (std::uint32_t i = 0; i < m.queue_count; ++i)
            if (m.queue[i].kind == SourceKind::Impulse) m.queue[keep++] = m.queue[i];
        m.queue_count = keep;
        return;
    }

    m.dev.make_current();
    hipStream_t s = m.dev.stream;
    // All substeps' stamps are staged at once, into the staging buffer whose
    // previous upload (two updates ago), if any, is done. When every substep
    // has few enough stamps they travel inside the kernel arguments; else all
    // of them are uploaded in one copy.
    m.stamp_turn ^= 1u;
    m.h_stamps = m.h_stamps_buf[m.stamp_turn];
    if (m.stamps_pending[m.stamp_turn]) {
        m.dev.note(hipEventSynchronize(m.stamps_free[m.stamp_turn]));
        m.stamps_pending[m.stamp_turn] = false;
    }
    std::uint32_t* first = m.first.data();
    std::size_t k = 0;
    for (std::uint32_t sidx = 0; sidx < steps; ++sidx) {
        first[sidx] = static_cast<std::uint32_t>(k);
        const float sub_time = static_cast<float>(sidx) * m.dt;
        for (std::uint32_t i = 0; i < m.queue_count; ++i) {
            const Disturbance& d = m.queue[i];
            ocean_hip_stamp st;
            bool live;
            if (d.kind == SourceKind::Impulse) {
                if (sidx != 0) continue;
                live = m.make_stamp(d, 0.0f, 1.0f, st);
            } else {
                live = m.make_stamp(d, sub_time, m.dt, st);
            }
            if (live) m.h_stamps[k++] = st;
        }
    }
    first[steps] = static_cast<std::uint32_t>(k);
    std::uint32_t most = 0;
    for (std::uint32_t sidx = 0; sidx < steps; ++sidx) most = std::max(most, first[sidx + 1] - first[sidx]);
    const bool by_args = most <= ocean_hip_iwave_arg_stamps();
    if (k > 0 && !by_args) {
        m.dev.note(hipMemcpyAsync(m.d_stamps, m.h_stamps, k * sizeof(ocean_hip_stamp),
                                  hipMemcpyHostToDevice, s));
        m.dev.note(hipEventRecord(m.stamps_free[m.stamp_turn], s));
        m.stamps_pending[m.stamp_turn] = true;
    }
    const int dirichlet = (m.desc.obstruction == Obstruction::Dirichlet) ? 1 : 0;
    for (std::uint32_t sidx = 0; sidx < steps; ++sidx) {
        const ocean_hip_iwave f = m.view();
        const std::uint32_t c = first[sidx + 1] - first[sidx];
        if (c > 0 && by_args)
            m.dev.note(ocean_hip_iwave_inject_args(&f, m.h_stamps + first[sidx], c,
                                                   m.d_means + first[sidx], s));
        else if (c > 0)
            m.dev.note(ocean_hip_iwave_inject(&f, m.d_stamps + first[sidx], c,
                                              m.d_means + first[sidx], s));
        if (m.has_mask)
            m.dev.note(ocean_hip_iwave_obstruct(&f, m.d_solid, m.d_source, m.d_solid_count,
                                                static_cast<std::uint32_t>(m.cells), dirichlet, s));
        m.dev.note(ocean_hip_iwave_step(&f, s));
        std::swap(m.cur_i, m.old_i);
    }
    m.queue_count = 0;
    m.finalize();
}

std::uint32_t InteractionField::last_substeps() const noexcept { return impl_->substeps_last; }
bool InteractionField::last_update_clamped() const noexcept { return impl_->clamped_last; }

// reference src/interaction.cpp:601-705
void InteractionField::recenter(float world_x, float world_z) noexcept
{
    Impl& m = *impl_;
    const float half = 0.5f * m.desc.extent;
    const std::int32_t tx = static_cast<std::int32_t>(std::lround((world_x - half) / m.cell));
    const std::int32_t tz = static_cast<std::int32_t>(std::lround((world_z - half) / m.cell));
    const std::int32_t dx = tx - m.origin_x;
    const std::int32_t dz = tz - m.origin_z;
    m.shift_x = dx;
    m.shift_z = dz;
    if (dx == 0 && dz == 0) {
        m.moved = false;
        return;
    }
    m.moved = true;
    const std::int32_t n = static_cast<std::int32_t>(m.n);
    m.origin_x = tx;
    m.origin_z = tz;
    m.dev.make_current();
    if (std::abs(dx) >= n || std::abs(dz) >= n) {
        for (int i = 0; i < 3; ++i)
            m.dev.note(hipMemsetAsync(m.plane[i], 0, m.plane_floats * 4, m.dev.stream));
        m.dev.note(hipMemsetAsync(m.d_mask[m.mask_i], 0, m.cells, m.dev.stream));
        m.rebuild_mask_tables();
        m.finalize();
        return;
    }
    m.shift_planes(dx, dz);
    if (m.has_mask) {
        m.dev.note(ocean_hip_iwave_mask_shift(m.d_mask[m.mask_i], m.d_mask[1 - m.mask_i], m.n, dx, dz,
                                              m.dev.stream));
        m.mask_i = 1 - m.mask_i;
        m.rebuild_mask_tables();
    }
    m.finalize();
}

bool InteractionField::recentered() const noexcept { return impl_->moved; }
std::int32_t InteractionField::last_shift_x() const noexcept { return impl_->shift_x; }
std::int32_t InteractionField::last_shift_z() const noexcept { return impl_->shift_z; }

// reference src/interaction.cpp:711-726
void InteractionField::set_obstruction(const std::uint8_t* m8) noexcept
{
    Impl& m = *impl_;
    m.dev.make_current();
    if (m8 == nullptr) {
        m.dev.note(hipMemsetAsync(m.d_mask[m.mask_i], 0, m.cells, m.dev.stream));
        m.has_mask = false;
        return;
    }
    const int t = m.mask_turn;
    m.mask_turn = 1 - t;
    if (m.mask_read_pending[t]) {
        m.dev.note(hipEventSynchronize(m.mask_read[t]));
        m.mask_read_pending[t] = false;
    }
    std::memcpy(m.h_mask_stage[t], m8, m.cells);
    m.has_mask = true;
    if (m.d_mask_stage[t] != nullptr) {
        // the row kernel reads the staging and keeps a device copy of the mask
        m.dev.note(ocean_hip_iwave_mask_tables(m.d_mask_stage[t], m.d_mask[m.mask_i], m.n,
                                               m.d_mask_work, m.d_nearest, m.d_solid, m.d_source,
                                               m.d_solid_count, m.dev.stream));
    } else {
        m.dev.note(hipMemcpyAsync(m.d_mask[m.mask_i], m.h_mask_stage[t], m.cells,
                                  hipMemcpyHostToDevice, m.dev.stream));
        m.rebuild_mask_tables();
    }
    m.dev.note(hipEventRecord(m.mask_read[t], m.dev.stream));
    m.mask_read_pending[t] = true;
}

bool InteractionField::pack_rgba16(std::uint16_t* out, bool on_device) const noexcept
{
    if (out == nullptr || reinterpret_cast<std::uintptr_t>(out) % 8 != 0) return false;
    Impl& m = *impl_;
    std::lock_guard<std::mutex> lock(m.query_mu);
    m.dev.make_current();
    const std::uint32_t cells = m.n * m.n;
    const std::size_t bytes = std::size_t(cells) * 8;
    std::uint16_t* dst = out;
    if (!on_device) {
        if (m.d_pack == nullptr && hipMalloc(&m.d_pack, bytes) != hipSuccess) {
            m.d_pack = nullptr;
            return false;
        }
        dst = static_cast<std::uint16_t*>(m.d_pack);
    }
    m.dev.note(ocean_hip_pack_rgba16(m.out, nullptr, dst, cells, m.dev.stream));
    if (!on_device) {
        m.dev.note(hipMemcpyAsync(out, dst, bytes, hipMemcpyDeviceToHost, m.dev.stream));
        m.dev.note(hipStreamSynchronize(m.dev.stream));
    }
    return m.dev.status == 0;
}

bool InteractionField::sample_batch(const float* xz, std::uint32_t count,
                                    InteractionSample* out) const noexcept
{
    static_assert(sizeof(InteractionSample) == 16);
    if (count == 0) return true;
    if (xz == nullptr || out == nullptr) return false;
    Impl& m = *impl_;
    std::lock_guard<std::mutex> lock(m.query_mu);
    m.dev.make_current();
    const ocean_hip_iwave f = m.view();
    if (count <= detail::kMappedQueries && m.mapped.ready()) {
        std::memcpy(m.mapped.xz_host(), xz, std::size_t(count) * 8);
        m.dev.note(ocean_hip_iwave_sample(&f, m.mapped.xz_dev(), count, m.mapped.out_dev<float>(),
                                          m.dev.stream));
        m.dev.note(hipStreamSynchronize(m.dev.stream));
        std::memcpy(out, m.mapped.out_host(), std::size_t(count) * 16);
        return m.dev.status == 0;
    }
    for (std::uint32_t done = 0; done < count;) {
        const std::uint32_t k = std::min(detail::kQueryStage, count - done);
        std::memcpy(m.h_qxz, xz + 2 * std::size_t(done), std::size_t(k) * 8);
        m.dev.note(hipMemcpyAsync(m.d_qxz, m.h_qxz, std::size_t(k) * 8, hipMemcpyHostToDevice, m.dev.stream));
        m.dev.note(ocean_hip_iwave_sample(&f, m.d_qxz, k, m.d_qout, m.dev.stream));
        m.dev.note(hipMemcpyAsync(m.h_qout, m.d_qout, std::size_t(k) * 16, hipMemcpyDeviceToHost, m.dev.stream));
        m.dev.note(hipStreamSynchronize(m.dev.stream));
        std::memcpy(out + done, m.h_qout, std::size_t(k) * 16);
        done += k;
    }
    return m.dev.status == 0;
}

InteractionSample InteractionField::sample_at(float world_x, float world_z) const noexcept
{
    const float xz[2] = {world_x, world_z};
    InteractionSample s;
    if (!sample_batch(xz, 1, &s)) return InteractionSample{};
    return s;
}

float InteractionField::height_at(float world_x, float world_z) const noexcept
{
    return sample_at(world_x, world_z).height;
}

InteractionBuffers InteractionField::buffers() const noexcept
{
    Impl& m = *impl_;
    std::lock_guard<std::mutex> lock(m.query_mu);
    if (m.mirror_stale) {
        m.dev.make_current();
        m.dev.note(hipMemcpyAsync(m.h_out, m.out, m.cells * 16, hipMemcpyDeviceToHost, m.dev.stream));
        m.dev.note(hipStreamSynchronize(m.dev.stream));
        m.mirror_stale = false;
    }
    InteractionBuffers b;
    b.field = m.h_out;
    b.size = m.n;
    b.extent = m.desc.extent;
    b.origin_x = static_cast<float>(m.origin_x) * m.cell;
    b.origin_z = static_cast<float>(m.origin_z) * m.cell;
    return b;
}

const InteractionDesc& InteractionField::desc() const noexcept { return impl_->desc; }
float InteractionField::fixed_dt() const noexcept { return impl_->dt; }
float InteractionField::stable_dt_limit() const noexcept { return impl_->dt_limit; }
float InteractionField::kernel_dispersion_error() const noexcept
{
    return static_cast<float>(impl_->kernel.peak_relative_error);
}

double InteractionField::energy() const noexcept
{
    Impl& m = *impl_;
    std::lock_guard<std::mutex> lock(m.query_mu);
    m.dev.make_current();
    const ocean_hip_iwave f = m.view();
    double* res = m.d_energy + 2 * m.cells;
    m.dev.note(ocean_hip_iwave_energy(&f, static_cast<double>(m.desc.gravity), m.d_energy, res, m.dev.stream));
    double e = 0.0;
    m.dev.note(hipMemcpyAsync(&e, res, 8, hipMemcpyDeviceToHost, m.dev.stream));
    m.dev.note(hipStreamSynchronize(m.dev.stream));
    return e;
}

std::uint32_t InteractionField::obstruction_tables(std::uint32_t* solid, std::uint32_t* source,
                                                   std::uint32_t capacity) const noexcept
{
    Impl& m = *impl_;
    if (!m.has_mask) return 0;
    std::lock_guard<std::mutex> lock(m.query_mu);
    m.dev.make_current();
    std::uint32_t count = 0;
    m.dev.note(hipMemcpyAsync(&count, m.d_solid_count, 4, hipMemcpyDeviceToHost, m.dev.stream));
    m.dev.note(hipStreamSynchronize(m.dev.stream));
    const std::uint32_t k = std::min(count, capacity);
    if (k > 0 && solid != nullptr)
        m.dev.note(hipMemcpyAsync(solid, m.d_solid, std::size_t(k) * 4, hipMemcpyDeviceToHost, m.dev.stream));
    if (k > 0 && source != nullptr)
        m.dev.note(hipMemcpyAsync(source, m.d_source, std::size_t(k) * 4, hipMemcpyDeviceToHost, m.dev.stream));
    m.dev.note(hipStreamSynchronize(m.dev.stream));
    return count;
}

bool InteractionField::sample_combined_batch(const void* device_levels, std::uint32_t level_count,
                                             bool as_cascade, const float* xz, std::uint32_t count,
                                             Surface* out) const noexcept
{
    static_assert(sizeof(Surface) == sizeof(ocean_hip_surface));
    if (count == 0) return true;
    if (xz == nullptr || out == nullptr || device_levels == nullptr) return false;
    Impl& m = *impl_;
    std::lock_guard<std::mutex> lock(m.query_mu);
    m.dev.make_current();
    const ocean_hip_iwave f = m.view();
    auto* d_out = reinterpret_cast<ocean_hip_surface*>(m.d_qout);
    if (count <= detail::kMappedQueries && m.mapped.ready()) {
        std::memcpy(m.mapped.xz_host(), xz, std::size_t(count) * 8);
        m.dev.note(ocean_hip_combined_sample(static_cast<const ocean_hip_level*>(device_levels),
                                             level_count, as_cascade ? 1 : 0, &f, m.mapped.xz_dev(), count,
                                             m.mapped.out_dev<ocean_hip_surface>(), m.dev.stream));
        m.dev.note(hipStreamSynchronize(m.dev.stream));
        std::memcpy(out, m.mapped.out_host(), std::size_t(count) * sizeof(ocean_hip_surface));
        return m.dev.status == 0;
    }
    for (std::uint32_t done = 0; done < count;) {
        const std::uint32_t k = std::min(detail::kQueryStage, count - done);
        std::memcpy(m.h_qxz, xz + 2 * std::size_t(done), std::size_t(k) * 8);
        m.dev.note(hipMemcpyAsync(m.d_qxz, m.h_qxz, std::size_t(k) * 8, hipMemcpyHostToDevice, m.dev.stream));
        m.dev.note(ocean_hip_combined_sample(static_cast<const ocean_hip_level*>(device_levels),
                                             level_count, as_cascade ? 1 : 0, &f, m.d_qxz, k, d_out,
                                             m.dev.stream));
        m.dev.note(hipMemcpyAsync(m.h_qout, d_out, std::size_t(k) * sizeof(ocean_hip_surface),
                                  hipMemcpyDeviceToHost, m.dev.stream));
        m.dev.note(hipStreamSynchronize(m.dev.stream));
        std::memcpy(out + done, m.h_qout, std::size_t(k) * sizeof(ocean_hip_surface));
        done += k;
    }
    return m.dev.status == 0;
}

const float* InteractionField::device_field() const noexcept { return impl_->out; }
void InteractionField::synchronize() const noexcept
{
    impl_->dev.make_current();
    impl_->dev.note(hipStreamSynchronize(impl_->dev.stream));
}
int InteractionField::last_error() const noexcept { return impl_->dev.status; }
detail::InteractionDevice& InteractionField::device() const noexcept { return impl_->dev; }

// ---------------------------------------------------------------------------
// WaterSurface (reference src/interaction.cpp:851-907)
// ---------------------------------------------------------------------------

WaterSurface::WaterSurface(const Ocean& o, const InteractionField& f) noexcept : ocean_(&o), field_(&f) {}
WaterSurface::WaterSurface(const CascadeStack& s, const InteractionField& f) noexcept
    : stack_(&s), field_(&f)
{
}

InteractionSample WaterSurface::interaction_at(float x, float z) const noexcept
{
    return field_->sample_at(x, z);
}

bool WaterSurface::sample_batch(const float* xz, std::uint32_t count, Surface* out) const noexcept
{
    // The FFT surface's frames are on another stream: drain it first, then
    // query on the field's stream.
    if (stack_) {
        hipStreamSynchronize(static_cast<hipStream_t>(stack_->stream()));
        return field_->sample_combined_batch(stack_->device_levels(),
                                             static_cast<std::uint32_t>(stack_->level_count()),
                                             true, xz, count, out);
    }
    ocean_->synchronize();
    return field_->sample_combined_batch(ocean_->device().d_level, 1, false, xz, count, out);
}

Surface WaterSurface::sample_at(float world_x, float world_z) const noexcept
{
    const float xz[2] = {world_x, world_z};
    Surface s;
    if (!sample_batch(xz, 1, &s)) return Surface{};
    return s;
}

float WaterSurface::height_at(float world_x, float world_z) const noexcept
{
    return sample_at(world_x, world_z).height;
}

}  // namespace ocean

// ocean::Ocean on one MI355X.
//
// Mirrors the reference's Ocean (src/ocean.cpp:65-289): validate, allocate
// everything once, evaluate t = 0 before returning; update() is absolute-time
// and allocation-free. The difference is that the frame is two HIP kernels
// enqueued on the ocean's stream (ocean_hip_frame) and the outputs live in
// device memory, copied to a pinned host mirror only when buffers() is asked.
#include "ocean/foam.hpp"
#include "ocean/ocean.hpp"

#include "ocean_device.hpp"
#include "spectrum_host.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace ocean {
namespace {

constexpr double kInvTwoPiD = 0.159154943091895335769;

constexpr bool is_power_of_two(std::uint32_t v) noexcept { return v != 0 && (v & (v - 1)) == 0; }

// reference src/ocean.cpp:28-59 (same checks, same order, same exception type)
void validate(const OceanDesc& d)
{
    if (!is_power_of_two(d.size) || d.size < 16 || d.size > 4096) {
        throw std::invalid_argument("ocean: size must be a power of two in [16, 4096], got " +
                                    std::to_string(d.size));
    }
    if (!(d.patch_length > 0.0f)) throw std::invalid_argument("ocean: patch_length must be positive");
    if (!(d.spectrum.gravity > 0.0f)) throw std::invalid_argument("ocean: gravity must be positive");
    if (!(d.spectrum.fetch > 0.0f)) throw std::invalid_argument("ocean: fetch must be positive");
    if (!(d.spectrum.wind_speed >= 0.0f))
        throw std::invalid_argument("ocean: wind_speed must be non-negative");
    if (!(d.spectrum.peak_enhancement >= 1.0f))
        throw std::invalid_argument("ocean: peak_enhancement must be >= 1 (1 = Pierson-Moskowitz)");
    if (!(d.choppiness >= 0.0f)) throw std::invalid_argument("ocean: choppiness must be non-negative");
}

std::size_t align_up(std::size_t v, std::size_t a) { return (v + a - 1) / a * a; }

// One cell the device could not certify, recomputed by the host chain.
struct TableFix {
    std::uint32_t index;
    float h0r, h0i, omega;
};

// h0, omega and the twiddles by the device kernels (spectrum_kernels.hip);
// the cells and twiddles they flag are recomputed by the host's libm chain
// (spectrum_host.cpp) and written over the device's values, so the tables
// equal the host build bit for bit. Returns the number of repaired entries.
std::uint32_t build_tables_on_device(const OceanDesc& desc, float* h0, float* omega, float* tw,
                                     unsigned char* scratch, hipStream_t stream)
{
    using detail::hip_check;
    const std::uint32_t n = desc.size;
    const std::size_t cells = std::size_t(n) * n;
    const oceanamd::SpectrumConstants c = oceanamd::spectrum_constants(desc);
    ocean_hip_spectrum_params p{};
    p.n = n;
    p.patch_length = desc.patch_length;
    p.seed = desc.seed;
    p.dk = c.dk;
    p.gravity = c.gravity;
    p.depth = c.depth;
    p.omega_p = c.omega_p;
    p.alpha = c.alpha;
    p.gamma = c.gamma;
    p.wind = c.wind;
    p.swell = c.swell;
    p.wind_direction = c.wind_direction;
    p.cutoff_l = c.cutoff_l;
    p.mu = c.mu;
    p.q_den = c.q_den;
    unsigned char* flags = scratch;
    unsigned char* tflags = scratch + cells;
    hip_check(ocean_hip_spectrum(&p, h0, omega, flags, stream), "device spectrum");
    hip_check(ocean_hip_twiddles(n, tw, tflags, stream), "device twiddles");
    std::vector<unsigned char> hf(cells + n / 2);
    hip_check(hipMemcpyAsync(hf.data(), flags, hf.size(), hipMemcpyDeviceToHost, stream), "flags");
    hip_check(hipStreamSynchronize(stream), "device tables");
    std::vector<TableFix> fixes;
    for (std::size_t i = 0; i < cells; ++i) {
        if (!hf[i]) continue;
        TableFix f{std::uint32_t(i), 0.0f, 0.0f, 0.0f};
        float hv[2];
        oceanamd::spectrum_cell(desc, std::uint32_t(i % n), std::uint32_t(i / n), hv, &f.omega);
        f.h0r = hv[0];
        f.h0i = hv[1];
        fixes.push_back(f);
    }
    std::vector<float> twfix;
    std::vector<std::uint32_t> twidx;
    for (std::uint32_t j = 0; j < n / 2; ++j) {
        if (!hf[cells + j]) continue;
        float v[2];
        oceanamd::twiddle_at(n, j, v);
        twidx.push_back(j);
        twfix.push_back(v[0]);
        twfix.push_back(v[1]);
    }
    if (!fixes.empty()) {
        hip_check(hipMemcpyAsync(scratch, fixes.data(), fixes.size() * sizeof(TableFix),
                                 hipMemcpyHostToDevice, stream), "upload fixes");
        hip_check(ocean_hip_spectrum_fix(scratch, std::uint32_t(fixes.size()), h0, omega, stream),
                  "apply fixes");
    }
    for (std::size_t k = 0; k < twidx.size(); ++k)
        hip_check(hipMemcpyAsync(tw + 2 * std::size_t(twidx[k]), twfix.data() + 2 * k, 8,
                                 hipMemcpyHostToDevice, stream), "upload twiddle fix");
    hip_check(hipStreamSynchronize(stream), "table fixes");
    return std::uint32_t(fixes.size() + twidx.size());
}



}  // namespace

namespace detail {

OceanDevice::~OceanDevice()
{
    if (stream) hipStreamSynchronize(stream);
    if (arena) hipFree(arena);
    if (q_stage) hipFree(q_stage);
    if (pack_stage) hipFree(pack_stage);
    if (h_disp) hipHostFree(h_disp);
    if (h_normal) hipHostFree(h_normal);
    if (h_vel) hipHostFree(h_vel);
    if (h_q_xz) hipHostFree(h_q_xz);
    if (h_q_out) hipHostFree(h_q_out);
    for (auto& slot : timing_ev)
        for (auto& e : slot)
            if (e) hipEventDestroy(e);
    if (own_stream && stream) hipStreamDestroy(stream);
}

void OceanDevice::make_current() const noexcept
{
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != device) hipSetDevice(device);
}

void OceanDevice::create(const OceanDesc& desc, hipStream_t shared_stream, bool lazy_host)
{
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        throw std::runtime_error("ocean: no HIP device available");
    hip_check(hipGetDevice(&device), "hipGetDevice");

    const std::uint32_t n = desc.size;
    const std::size_t cells = static_cast<std::size_t>(n) * n;
    fields = desc.compute_velocity ? 6 : 4;

    if (shared_stream) {
        stream = shared_stream;
    } else {
        hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        own_stream = true;
    }

    // Device arena layout (256-byte aligned pieces).
    const std::size_t a = 256;
    std::size_t off = 0;
    const std::size_t o_h0 = off;     off = align_up(off + cells * 8, a);
    const std::size_t o_omega = off;  off = align_up(off + cells * 4, a);
    const std::size_t o_kinv = off;   off = align_up(off + cells * 4, a);
    const std::size_t o_kx = off;     off = align_up(off + std::size_t(n) * 4, a);
    const std::size_t o_tw = off;     off = align_up(off + std::max<std::size_t>(n / 2, 1) * 8, a);
    const std::size_t o_spec = off;   off = align_up(off + cells * 8 * fields, a);
    const std::size_t o_disp = off;   off = align_up(off + cells * 16, a);
    const std::size_t o_norm = off;   off = align_up(off + cells * 16, a);
    std::size_t o_vel = 0;
    if (desc.compute_velocity) { o_vel = off; off = align_up(off + cells * 16, a); }
    // Query staging only for oceans that answer host queries themselves; the
    // levels of a batch get it on first use.
    const std::size_t q = lazy_host ? 0 : std::size_t(kQueryStage);
    const std::size_t o_qxz = off;    off = align_up(off + q * 8, a);
    const std::size_t o_qout = off;   off = align_up(off + q * sizeof(ocean_hip_surface), a);
    const std::size_t o_lvl = off;    off = align_up(off + sizeof(ocean_hip_level), a);

    hip_check(hipMalloc(&arena, off), "hipMalloc");
    char* base = static_cast<char*>(arena);

    level.n = n;
    level.patch_length = desc.patch_length;
    level.choppiness = desc.choppiness;
    level.foam_threshold = desc.foam_threshold;
    // reference src/core/evolve.cpp:247-248
    level.inv_foam_threshold = (desc.foam_threshold > 1e-6f) ? (1.0f / desc.foam_threshold) : 0.0f;
    level.h0 = reinterpret_cast<const float*>(base + o_h0);
    level.omega = reinterpret_cast<const float*>(base + o_omega);
    level.twiddles = reinterpret_cast<const float*>(base + o_tw);
    level.kinv = reinterpret_cast<const float*>(base + o_kinv);
    level.kx = reinterpret_cast<const float*>(base + o_kx);
    level.spec = reinterpret_cast<float*>(base + o_spec);
    level.disp = reinterpret_cast<float*>(base + o_disp);
    level.normal = reinterpret_cast<float*>(base + o_norm);
    level.vel = desc.compute_velocity ? reinterpret_cast<float*>(base + o_vel) : nullptr;
    own_disp = level.disp;
    own_normal = level.normal;
    own_vel = level.vel;
    if (!lazy_host) {
        q_xz = reinterpret_cast<float*>(base + o_qxz);
        q_out = reinterpret_cast<ocean_hip_surface*>(base + o_qout);
    }
    d_level = reinterpret_cast<ocean_hip_level*>(base + o_lvl);

    if (!lazy_host) {
        alloc_host_mirrors();
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&h_q_xz), std::size_t(kQueryStage) * 8, hipHostMallocDefault), "hipHostMalloc");
        hip_check(hipHostMalloc(reinterpret_cast<void**>(&h_q_out), std::size_t(kQueryStage) * sizeof(ocean_hip_surface), hipHostMallocDefault), "hipHostMalloc");
        if (status != 0) throw std::bad_alloc();
    }

    // Constant tables, built on the device; the spec scratch (unused until the
    // first frame) holds the certification flags and the repaired cells.
    init_repaired = build_tables_on_device(desc, reinterpret_cast<float*>(base + o_h0),
                                           reinterpret_cast<float*>(base + o_omega),
                                           reinterpret_cast<float*>(base + o_tw),
                                           reinterpret_cast<unsigned char*>(base + o_spec), stream);
    hip_check(ocean_hip_wavenumbers(n, desc.patch_length, reinterpret_cast<float*>(base + o_kinv),
                                    reinterpret_cast<float*>(base + o_kx), stream),
              "wavenumber planes");
    hip_check(hipMemcpyAsync(d_level, &level, sizeof(level), hipMemcpyHostToDevice, stream), "upload level");
    hip_check(hipStreamSynchronize(stream), "upload");
}

void OceanDevice::enqueue_frame(double time) noexcept
{
    make_current();
    // Hoisted exactly as the reference hoists it (src/core/evolve.cpp:33).
    const double t2 = time * kInvTwoPiD;
    if (timing_every != 0 && timing_counter++ % timing_every == 0) {
        hipEvent_t* ev = timing_ev[timing_next];
        timing_next = (timing_next + 1) % kTimingSlots;
        if (timing_used < kTimingSlots) ++timing_used;
        note(ocean_hip_frame_timed(&level, t2, stream, ev));
    } else {
        note(ocean_hip_frame(&level, t2, stream));
    }
    mirror_stale = true;
}

int OceanDevice::set_frame_timing(std::uint32_t every) noexcept
{
    make_current();
    if (every != 0 && timing_ev[0][0] == nullptr) {
        for (auto& slot : timing_ev)
            for (auto& e : slot)
                if (hipEventCreate(&e) != hipSuccess) return -1;
    }
    timing_every = every;
    timing_counter = 0;
    timing_next = 0;
    timing_used = 0;
    return 0;
}

int OceanDevice::read_frame_timing(float* row_ms, float* col_ms, std::uint32_t* samples) noexcept
{
    make_current();
    if (hipStreamSynchronize(stream) != hipSuccess) return -1;
    double row = 0.0, col = 0.0;
    for (std::uint32_t i = 0; i < timing_used; ++i) {
        float a = 0.0f, b = 0.0f;
        if (hipEventElapsedTime(&a, timing_ev[i][0], timing_ev[i][1]) != hipSuccess ||
            hipEventElapsedTime(&b, timing_ev[i][2], timing_ev[i][3]) != hipSuccess)
            return -1;
        row += a;
        col += b;
    }
    const std::uint32_t n = timing_used;
    if (row_ms) *row_ms = n ? static_cast<float>(row / n) : 0.0f;
    if (col_ms) *col_ms = n ? static_cast<float>(col / n) : 0.0f;
    if (samples) *samples = n;
    timing_used = 0;
    timing_next = 0;
    return 0;
}

void OceanDevice::alloc_host_mirrors() noexcept
{
    const std::size_t bytes = static_cast<std::size_t>(level.n) * level.n * 16;
    if (!h_disp) note(hipHostMalloc(reinterpret_cast<void**>(&h_disp), bytes, hipHostMallocDefault));
    if (!h_normal) note(hipHostMalloc(reinterpret_cast<void**>(&h_normal), bytes, hipHostMallocDefault));
    if (level.vel && !h_vel) note(hipHostMalloc(reinterpret_cast<void**>(&h_vel), bytes, hipHostMallocDefault));
}

void OceanDevice::refresh_mirror() noexcept
{
    std::lock_guard<std::mutex> lock(query_mu);
    if (!mirror_stale) return;
    make_current();
    alloc_host_mirrors();
    if (status != 0) return;
    const std::size_t bytes = static_cast<std::size_t>(level.n) * level.n * 16;
    note(hipMemcpyAsync(h_disp, level.disp, bytes, hipMemcpyDeviceToHost, stream));
    note(hipMemcpyAsync(h_normal, level.normal, bytes, hipMemcpyDeviceToHost, stream));
    if (level.vel) note(hipMemcpyAsync(h_vel, level.vel, bytes, hipMemcpyDeviceToHost, stream));
    note(hipStreamSynchronize(stream));
    mirror_stale = false;
}

bool OceanDevice::sample_host(const float* xz, std::uint32_t count, Surface* out) noexcept
{
    static_assert(sizeof(Surface) == sizeof(ocean_hip_surface));
    std::lock_guard<std::mutex> lock(query_mu);
    make_current();
    if (count <= kMappedQueries && mapped.ready()) {
        std::memcpy(mapped.xz_host(), xz, std::size_t(count) * 8);
        note(ocean_hip_sample(&level, mapped.xz_dev(), count, mapped.out_dev<ocean_hip_surface>(), nullptr,
                              stream));
        note(hipStreamSynchronize(stream));
        std::memcpy(out, mapped.out_host(), std::size_t(count) * sizeof(Surface));
        return status == 0;
    }
    if (q_xz == nullptr) {  // a batch level's first host query: its staging, once
        // All three or none: on a partial failure the ones that succeeded are
        // released, so a later call starts from nothing again.
        void* d = nullptr;
        void* hx = nullptr;
        void* ho = nullptr;
        const hipError_t e0 = hipMalloc(&d, std::size_t(kQueryStage) * (8 + sizeof(ocean_hip_surface)));
        const hipError_t e1 = hipHostMalloc(&hx, std::size_t(kQueryStage) * 8, hipHostMallocDefault);
        const hipError_t e2 = hipHostMalloc(&ho, std::size_t(kQueryStage) * sizeof(ocean_hip_surface),
                                            hipHostMallocDefault);
        if (e0 != hipSuccess || e1 != hipSuccess || e2 != hipSuccess) {
            if (e0 == hipSuccess) hipFree(d);
            if (e1 == hipSuccess) hipHostFree(hx);
            if (e2 == hipSuccess) hipHostFree(ho);
            note(e0 != hipSuccess ? e0 : e1 != hipSuccess ? e1 : e2);
            return false;
        }
        q_stage = d;
        h_q_xz = static_cast<float*>(hx);
        h_q_out = static_cast<ocean_hip_surface*>(ho);
        q_xz = static_cast<float*>(d);
        q_out = reinterpret_cast<ocean_hip_surface*>(q_xz + 2 * std::size_t(kQueryStage));
    }
    for (std::uint32_t done = 0; done < count;) {
        const std::uint32_t m = std::min(kQueryStage, count - done);
        std::memcpy(h_q_xz, xz + 2 * std::size_t(done), std::size_t(m) * 8);
        note(hipMemcpyAsync(q_xz, h_q_xz, std::size_t(m) * 8, hipMemcpyHostToDevice, stream));
        note(ocean_hip_sample(&level, q_xz, m, q_out, nullptr, stream));
        note(hipMemcpyAsync(h_q_out, q_out, std::size_t(m) * sizeof(ocean_hip_surface),
                            hipMemcpyDeviceToHost, stream));
        note(hipStreamSynchronize(stream));
        std::memcpy(out + done, h_q_out, std::size_t(m) * sizeof(ocean_hip_surface));
        done += m;
    }
    return status == 0;
}

}  // namespace detail

struct Ocean::Impl {
    OceanDesc desc;
    double time = 0.0;
    detail::OceanDevice dev;
};

Ocean::Ocean(const OceanDesc& desc) : Ocean(desc, nullptr, false) {}

Ocean::Ocean(const OceanDesc& desc, void* shared_stream, bool lazy_host)
{
    validate(desc);
    impl_ = std::make_unique<Impl>();
    impl_->desc = desc;
    impl_->dev.create(desc, static_cast<hipStream_t>(shared_stream), lazy_host);
    // Valid, fully evaluated state at t = 0 before returning (reference
    // src/ocean.cpp:271-278).
    impl_->dev.enqueue_frame(0.0);
    impl_->time = 0.0;
    detail::hip_check(hipStreamSynchronize(impl_->dev.stream), "initial frame");
    if (impl_->dev.status != 0)
        throw std::runtime_error("ocean: initial frame failed: " +
                                 std::string(hipGetErrorString(hipError_t(impl_->dev.status.load()))));
}

Ocean::~Ocean() = default;
Ocean::Ocean(Ocean&&) noexcept = default;
Ocean& Ocean::operator=(Ocean&&) noexcept = default;

void Ocean::update(double time) noexcept
{
    impl_->dev.enqueue_frame(time);
    impl_->time = time;
}

Buffers Ocean::buffers() const noexcept
{
    impl_->dev.refresh_mirror();
    Buffers b;
    b.displacement = impl_->dev.h_disp;
    b.normal = impl_->dev.h_normal;
    b.velocity = impl_->dev.h_vel;
    b.size = impl_->desc.size;
    b.patch_length = impl_->desc.patch_length;
    return b;
}

DeviceBuffers Ocean::device_buffers() const noexcept
{
    DeviceBuffers b;
    b.displacement = impl_->dev.level.disp;
    b.normal = impl_->dev.level.normal;
    b.velocity = impl_->dev.level.vel;
    b.size = impl_->desc.size;
    b.patch_length = impl_->desc.patch_length;
    b.stream = impl_->dev.stream;
    return b;
}

float Ocean::height_at(float world_x, float world_z) const noexcept
{
    return sample_at(world_x, world_z).height;
}

Surface Ocean::sample_at(float world_x, float world_z) const noexcept
{
    const float xz[2] = {world_x, world_z};
    Surface s;
    if (!impl_->dev.sample_host(xz, 1, &s)) return Surface{};
    return s;
}

bool Ocean::sample_batch(const float* xz, std::uint32_t count, Surface* out) const noexcept
{
    if (count == 0) return true;
    if (xz == nullptr || out == nullptr) return false;
    return impl_->dev.sample_host(xz, count, out);
}

bool Ocean::sample_batch_device(const float* xz_dev, std::uint32_t count, Surface* out_dev,
                                float* heights_dev) const noexcept
{
    if (count == 0) return true;
    if (xz_dev == nullptr || (out_dev == nullptr && heights_dev == nullptr)) return false;
    impl_->dev.make_current();
    impl_->dev.note(ocean_hip_sample(&impl_->dev.level, xz_dev, count,
                                     reinterpret_cast<ocean_hip_surface*>(out_dev), heights_dev,
                                     impl_->dev.stream));
    return impl_->dev.status == 0;
}

bool Ocean::pack_rgba16(const FoamField* foam, std::uint16_t* displacement, std::uint16_t* normal,
                        bool on_device) const noexcept
{
    if (displacement == nullptr && normal == nullptr) return false;
    if (foam != nullptr && &foam->ocean() != this) return false;
    auto aligned = [](const void* p) { return reinterpret_cast<std::uintptr_t>(p) % 8 == 0; };
    if (!aligned(displacement) || !aligned(normal)) return false;
    detail::OceanDevice& d = impl_->dev;
    std::lock_guard<std::mutex> lock(d.query_mu);
    d.make_current();
    const std::uint32_t cells = d.level.n * d.level.n;
    const std::size_t bytes = std::size_t(cells) * 8;
    const float* alpha = foam != nullptr ? foam->device_data() : nullptr;
    std::uint16_t* dd = displacement;
    std::uint16_t* dn = normal;
    if (!on_device) {
        if (d.pack_stage == nullptr && hipMalloc(&d.pack_stage, 2 * bytes) != hipSuccess) {
            d.pack_stage = nullptr;
            return false;
        }
        dd = static_cast<std::uint16_t*>(d.pack_stage);
        dn = dd + std::size_t(cells) * 4;
    }
    if (displacement != nullptr)
        d.note(ocean_hip_pack_rgba16(d.level.disp, alpha, dd, cells, d.stream));
    if (normal != nullptr) d.note(ocean_hip_pack_rgba16(d.level.normal, nullptr, dn, cells, d.stream));
    if (!on_device) {
        if (displacement != nullptr)
            d.note(hipMemcpyAsync(displacement, dd, bytes, hipMemcpyDeviceToHost, d.stream));
        if (normal != nullptr) d.note(hipMemcpyAsync(normal, dn, bytes, hipMemcpyDeviceToHost, d.stream));
        d.note(hipStreamSynchronize(d.stream));
    }
    return d.status == 0;
}

namespace {

// p is 16-byte aligned device memory with at least `bytes` from p to the end
// of the allocation HIP knows it by.
bool device_span_ok(const void* p, std::size_t bytes) noexcept
{
    if (reinterpret_cast<std::uintptr_t>(p) % 16 != 0) return false;
    hipDeviceptr_t base = nullptr;
    std::size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, const_cast<void*>(p)) != hipSuccess) {
        (void)hipGetLastError();  // a refused pointer is the caller's error, not the ocean's
        return false;
    }
    const auto b = reinterpret_cast<std::uintptr_t>(base), q = reinterpret_cast<std::uintptr_t>(p);
    return q >= b && q - b <= size && size - (q - b) >= bytes;
}

}  // namespace

bool Ocean::set_precision(std::uint32_t mode) noexcept
{
    detail::OceanDevice& d = impl_->dev;
    if (!d.own_stream || mode > 2u) return false;  // a batch's or cascade's level: its owner's mode
    d.make_current();
    d.level.shape = mode << 16;  // the mode's default launch shapes
    d.note(hipMemcpyAsync(d.d_level, &d.level, sizeof(d.level), hipMemcpyHostToDevice, d.stream));
    d.note(hipStreamSynchronize(d.stream));
    return d.status == 0;
}

std::uint32_t Ocean::precision() const noexcept { return (impl_->dev.level.shape >> 16) & 0xffu; }

bool Ocean::bind_outputs(float* displacement, float* normal, float* velocity) noexcept
{
    detail::OceanDevice& d = impl_->dev;
    if (!d.own_stream) return false;  // a batch's or a cascade's level: its group holds the views
    d.make_current();
    const std::size_t bytes = std::size_t(d.level.n) * d.level.n * 16;
    if (velocity != nullptr && d.own_vel == nullptr) return false;
    for (const float* p : {displacement, normal, velocity})
        if (p != nullptr && !device_span_ok(p, bytes)) return false;
    d.level.disp = displacement != nullptr ? displacement : d.own_disp;
    d.level.normal = normal != nullptr ? normal : d.own_normal;
    d.level.vel = velocity != nullptr ? velocity : d.own_vel;
    d.note(hipMemcpyAsync(d.d_level, &d.level, sizeof(d.level), hipMemcpyHostToDevice, d.stream));
    // The current frame again, into the buffers now bound (not a timed frame).
    d.note(ocean_hip_frame(&d.level, impl_->time * kInvTwoPiD, d.stream));
    d.mirror_stale = true;
    return d.status == 0;
}

void Ocean::synchronize() const noexcept
{
    impl_->dev.make_current();
    impl_->dev.note(hipStreamSynchronize(impl_->dev.stream));
}

int Ocean::last_error() const noexcept { return impl_->dev.status; }
const OceanDesc& Ocean::desc() const noexcept { return impl_->desc; }
double Ocean::time() const noexcept { return impl_->time; }
detail::OceanDevice& Ocean::device() const noexcept { return impl_->dev; }

void Ocean::note_frame(double time) noexcept
{
    impl_->time = time;
    impl_->dev.mirror_stale = true;
}

void detail::validate_desc(const OceanDesc& desc) { validate(desc); }

}  // namespace ocean

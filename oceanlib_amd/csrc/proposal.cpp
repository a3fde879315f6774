// HeightProposal and the converged-query entry points (proposal_kernels.hip).
#include "ocean/proposal.hpp"

#include "ocean/batch.hpp"
#include "ocean/cascade.hpp"
#include "ocean_device.hpp"

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <cmath>
#include <cstring>
#include <new>
#include <stdexcept>

namespace ocean {

static_assert(ProposalShape::params == OCEAN_HIP_MLP_PARAMS);
static_assert(ProposalShape::inputs == OCEAN_HIP_MLP_IN && ProposalShape::outputs == OCEAN_HIP_MLP_OUT);

struct HeightProposal::Impl {
    struct Stream {
        hipStream_t s = nullptr;
        ~Stream()
        {
            if (s) hipStreamDestroy(s);
        }
    } stream;
    void* arena = nullptr;
    ocean_hip_mlp mlp{};
    std::int32_t* spacing = nullptr;   // 3 per call (queries)
    double* spacing_partial = nullptr; // 3 * 64
    float* coarse = nullptr;           // OCEAN_HIP_MLP_COARSE_BYTES of coarse-grid records
    float* prop = nullptr;             // max_queries * OCEAN_HIP_MLP_PROP
    float* partial = nullptr;          // (max_batch / 128) * params
    float* loss_partial = nullptr;
    float* loss = nullptr;
    float* h_loss = nullptr;           // pinned
    float* h_grad = nullptr;           // pinned, params
    hipEvent_t done = nullptr;         // last work enqueued on `stream`
    std::uint32_t max_queries = 0;
    std::uint32_t step = 0;            // Adam step count
    std::atomic<int> status{0};

    void note(hipError_t e) noexcept
    {
        // first error wins; queries may note concurrently on one handle
        int none = 0;
        if (e != hipSuccess) status.compare_exchange_strong(none, static_cast<int>(e));
    }
    void mark() noexcept { note(hipEventRecord(done, stream.s)); }

    ~Impl()
    {
        if (stream.s) hipStreamSynchronize(stream.s);
        if (done) hipEventDestroy(done);
        if (arena) hipFree(arena);
        if (h_loss) hipHostFree(h_loss);
        if (h_grad) hipHostFree(h_grad);
    }
};

namespace {

std::size_t up256(std::size_t v) { return (v + 255) / 256 * 256; }

void upload_params(HeightProposal::Impl& m, const float* params)
{
    const std::size_t bytes = std::size_t(ProposalShape::params) * 4;
    if (params != nullptr)
        detail::hip_check(hipMemcpyAsync(m.mlp.master, params, bytes, hipMemcpyHostToDevice, m.stream.s),
                          "upload proposal");
    else
        detail::hip_check(hipMemsetAsync(m.mlp.master, 0, bytes, m.stream.s), "clear proposal");
    detail::hip_check(hipMemsetAsync(m.mlp.m1, 0, bytes, m.stream.s), "clear moments");
    detail::hip_check(hipMemsetAsync(m.mlp.m2, 0, bytes, m.stream.s), "clear moments");
    detail::hip_check(ocean_hip_mlp_sync_bf16(&m.mlp, m.stream.s), "proposal bf16");
    detail::hip_check(hipStreamSynchronize(m.stream.s), "proposal upload");
    m.step = 0;
}

bool wait_for(const HeightProposal* model, hipStream_t s) noexcept
{
    if (model == nullptr) return true;
    return hipStreamWaitEvent(s, model->impl().done, 0) == hipSuccess;
}

// The proposal's scratch for one call: the model's own (sized at its
// creation), or the library's static device array (the allocation-free form).
struct ProposalScratch {
    std::int32_t* spacing;
    double* spacing_partial;
    float* coarse;
    float* prop;
    std::uint32_t max_queries;
};

bool converged(const ocean_hip_level* levels, std::size_t count_levels, hipStream_t s,
               const float* xz_dev, std::uint32_t count, float tol, const HeightProposal* model,
               float* heights_dev, std::uint8_t* iters_dev, std::uint8_t* flags_dev,
               const ProposalScratch* use = nullptr) noexcept
{
    if (count == 0) return true;
    if (xz_dev == nullptr || heights_dev == nullptr || count_levels == 0 || !(tol > 0.0f)) return false;
    if (model != nullptr) {
        HeightProposal::Impl& m = model->impl();
        const ProposalScratch own{m.spacing, m.spacing_partial, m.coarse, m.prop, m.max_queries};
        const ProposalScratch& sc = use != nullptr ? *use : own;
        if (count_levels > ProposalShape::levels || count > sc.max_queries) return false;
        if (!wait_for(model, s)) return false;
        if (ocean_hip_mlp_spacing_partial(levels, static_cast<std::uint32_t>(count_levels),
                                          sc.spacing_partial, s) != hipSuccess)
            return false;
        const bool ok = ocean_hip_height_converged(levels, static_cast<std::uint32_t>(count_levels),
                                                   xz_dev, count, tol, &m.mlp, sc.spacing_partial,
                                                   sc.spacing, sc.coarse, sc.prop, heights_dev,
                                                   iters_dev, flags_dev, s) == hipSuccess;
        // The proposal's parameters (and its own scratch) are in use until
        // this query ends: later work on the model waits for it.
        return ok && hipEventRecord(m.done, s) == hipSuccess &&
               hipStreamWaitEvent(m.stream.s, m.done, 0) == hipSuccess;
    }
    return ocean_hip_height_converged(levels, static_cast<std::uint32_t>(count_levels), xz_dev, count,
                                      tol, nullptr, nullptr, nullptr, nullptr, nullptr, heights_dev,
                                      iters_dev, flags_dev, s) == hipSuccess;
}

}  // namespace

HeightProposal::HeightProposal(const float* params, std::uint32_t max_queries)
{
    if (max_queries == 0) throw std::invalid_argument("HeightProposal: max_queries must be positive");
    if (ocean_hip_mlp_configure() != 0) throw std::runtime_error("HeightProposal: LDS configuration failed");
    impl_ = std::make_unique<Impl>();
    Impl& m = *impl_;
    m.max_queries = max_queries;
    detail::hip_check(hipStreamCreateWithFlags(&m.stream.s, hipStreamNonBlocking), "hipStreamCreate");
    detail::hip_check(hipEventCreateWithFlags(&m.done, hipEventDisableTiming), "hipEventCreate");
    const std::size_t P = ProposalShape::params;
    const std::size_t blocks = ProposalShape::max_batch / 128;
    const std::size_t sizes[] = {
        up256(std::size_t(OCEAN_HIP_MLP_BF16) * 2),     // bf16
        up256(P * 4), up256(P * 4), up256(P * 4), up256(P * 4),  // master, m1, m2, grad
        256,                                            // spacing
        up256(ProposalShape::levels * OCEAN_HIP_MLP_SPACING_PARTIALS * 8),
        up256(OCEAN_HIP_MLP_COARSE_BYTES),              // coarse grids
        up256(std::size_t(max_queries) * OCEAN_HIP_MLP_PROP * 4),  // prop
        up256(blocks * P * 4),                          // partial
        up256(blocks * 4),                              // loss partial
        256,                                            // loss
    };
    std::size_t total = 0;
    for (std::size_t v : sizes) total += v;
    detail::hip_check(hipMalloc(&m.arena, total), "hipMalloc proposal");
    char* p = static_cast<char*>(m.arena);
    std::size_t k = 0;
    auto take = [&](auto*& dst) {
        dst = reinterpret_cast<std::remove_reference_t<decltype(dst)>>(p);
        p += sizes[k++];
    };
    take(m.mlp.bf16);
    take(m.mlp.master);
    take(m.mlp.m1);
    take(m.mlp.m2);
    take(m.mlp.grad);
    take(m.spacing);
    take(m.spacing_partial);
    take(m.coarse);
    take(m.prop);
    take(m.partial);
    take(m.loss_partial);
    take(m.loss);
    detail::hip_check(hipHostMalloc(reinterpret_cast<void**>(&m.h_loss), 256, hipHostMallocDefault),
                      "hipHostMalloc");
    detail::hip_check(hipHostMalloc(reinterpret_cast<void**>(&m.h_grad), P * 4, hipHostMallocDefault),
                      "hipHostMalloc");
    upload_params(m, params);
    m.mark();
}

HeightProposal::~HeightProposal() = default;
HeightProposal::HeightProposal(HeightProposal&&) noexcept = default;
HeightProposal& HeightProposal::operator=(HeightProposal&&) noexcept = default;

HeightProposal::Impl& HeightProposal::impl() const noexcept { return *impl_; }
std::uint32_t HeightProposal::max_queries() const noexcept { return impl_->max_queries; }
int HeightProposal::last_error() const noexcept { return impl_->status; }
void HeightProposal::synchronize() const noexcept { impl_->note(hipStreamSynchronize(impl_->stream.s)); }

bool HeightProposal::set_params(const float* params) noexcept
{
    try {
        upload_params(*impl_, params);
        impl_->mark();
        return true;
    } catch (...) {
        impl_->note(hipErrorUnknown);
        return false;
    }
}

bool HeightProposal::get_params(float* out) const noexcept
{
    Impl& m = *impl_;
    if (out == nullptr) return false;
    m.note(hipMemcpyAsync(m.h_grad, m.mlp.master, std::size_t(ProposalShape::params) * 4,
                          hipMemcpyDeviceToHost, m.stream.s));
    m.note(hipStreamSynchronize(m.stream.s));
    if (m.status != 0) return false;
    std::memcpy(out, m.h_grad, std::size_t(ProposalShape::params) * 4);
    return true;
}

bool HeightProposal::forward(const float* feats_dev, std::uint32_t rows, float* out_dev) const noexcept
{
    Impl& m = *impl_;
    if (rows == 0) return true;
    if (feats_dev == nullptr || out_dev == nullptr) return false;
    m.note(ocean_hip_mlp_forward(&m.mlp, feats_dev, rows, out_dev, m.stream.s));
    m.mark();
    return m.status == 0;
}

bool HeightProposal::train_step(const float* feats_dev, const float* targets_dev,
                                const std::uint32_t* rows_dev, std::uint32_t batch, float lr,
                                float* loss) noexcept
{
    Impl& m = *impl_;
    if (feats_dev == nullptr || targets_dev == nullptr || rows_dev == nullptr) return false;
    if (batch == 0 || batch % 128 != 0 || batch > ProposalShape::max_batch || !(lr > 0.0f)) return false;
    ++m.step;
    const float c1 = static_cast<float>(1.0 - std::pow(0.9, double(m.step)));
    const float c2 = static_cast<float>(1.0 - std::pow(0.999, double(m.step)));
    m.note(ocean_hip_mlp_step(&m.mlp, feats_dev, targets_dev, rows_dev, batch, lr, c1, c2, m.partial,
                              m.loss_partial, loss ? m.loss : nullptr, m.stream.s));
    m.mark();
    if (loss != nullptr) {
        m.note(hipMemcpyAsync(m.h_loss, m.loss, 4, hipMemcpyDeviceToHost, m.stream.s));
        m.note(hipStreamSynchronize(m.stream.s));
        *loss = m.h_loss[0] / (float(batch) * float(2 * ProposalShape::levels));
    }
    return m.status == 0;
}

bool HeightProposal::gradient(const float* feats_dev, const float* targets_dev,
                              const std::uint32_t* rows_dev, std::uint32_t batch, float* grad,
                              float* loss) noexcept
{
    Impl& m = *impl_;
    if (feats_dev == nullptr || targets_dev == nullptr || rows_dev == nullptr || grad == nullptr)
        return false;
    if (batch == 0 || batch % 128 != 0 || batch > ProposalShape::max_batch) return false;
    m.note(ocean_hip_mlp_step(&m.mlp, feats_dev, targets_dev, rows_dev, batch, 0.0f, 1.0f, 1.0f,
                              m.partial, m.loss_partial, m.loss, m.stream.s));
    m.note(hipMemcpyAsync(m.h_grad, m.mlp.grad, std::size_t(ProposalShape::params) * 4,
                          hipMemcpyDeviceToHost, m.stream.s));
    m.note(hipMemcpyAsync(m.h_loss, m.loss, 4, hipMemcpyDeviceToHost, m.stream.s));
    m.note(hipStreamSynchronize(m.stream.s));
    m.mark();
    if (m.status != 0) return false;
    std::memcpy(grad, m.h_grad, std::size_t(ProposalShape::params) * 4);
    if (loss != nullptr) *loss = m.h_loss[0] / (float(batch) * float(2 * ProposalShape::levels));
    return true;
}

bool height_converged(const OceanBatch& batch, std::size_t first, std::size_t level_count,
                      const float* xz_dev, std::uint32_t count, float tol,
                      const HeightProposal* model, float* heights_dev, std::uint8_t* iters_dev,
                      std::uint8_t* flags_dev) noexcept
{
    if (level_count == 0 || first > batch.level_count() || level_count > batch.level_count() - first)
        return false;
    batch.level(0).device().make_current();
    const auto* levels = static_cast<const ocean_hip_level*>(batch.device_levels()) + first;
    return converged(levels, level_count, static_cast<hipStream_t>(batch.stream()), xz_dev, count, tol,
                     model, heights_dev, iters_dev, flags_dev);
}

bool height_converged_static(const OceanBatch& batch, std::size_t first, std::size_t level_count,
                             const float* xz_dev, std::uint32_t count, float tol,
                             const HeightProposal* model, float* heights_dev,
                             std::uint8_t* iters_dev, std::uint8_t* flags_dev,
                             int* error) noexcept
{
    *error = 0;
    if (level_count == 0 || first > batch.level_count() || level_count > batch.level_count() - first)
        return false;
    if (count > OCEAN_HIP_STATIC_MAX_QUERIES) return false;
    batch.level(0).device().make_current();
    ProposalScratch sc{};
    if (model != nullptr) {
        ocean_hip_static_scratch_view v{};
        const hipError_t e = ocean_hip_static_scratch(&v);
        if (e != hipSuccess) {
            *error = static_cast<int>(e);
            return false;
        }
        sc = ProposalScratch{v.spacing, v.spacing_partial, v.coarse, v.prop, v.max_queries};
    }
    const auto* levels = static_cast<const ocean_hip_level*>(batch.device_levels()) + first;
    return converged(levels, level_count, static_cast<hipStream_t>(batch.stream()), xz_dev, count, tol,
                     model, heights_dev, iters_dev, flags_dev, model != nullptr ? &sc : nullptr);
}

bool height_converged(const CascadeStack& stack, const float* xz_dev, std::uint32_t count, float tol,
                      const HeightProposal* model, float* heights_dev, std::uint8_t* iters_dev,
                      std::uint8_t* flags_dev) noexcept
{
    stack.level(0).device().make_current();
    const auto* levels = static_cast<const ocean_hip_level*>(stack.device_levels());
    return converged(levels, stack.level_count(), static_cast<hipStream_t>(stack.stream()), xz_dev,
                     count, tol, model, heights_dev, iters_dev, flags_dev);
}

bool proposal_starts(const OceanBatch& batch, std::size_t first, const float* xz_dev,
                     std::uint32_t count, const HeightProposal& model, float* starts_dev,
                     std::int32_t* spacing_dev) noexcept
{
    if (first > batch.level_count() || ProposalShape::levels > batch.level_count() - first) return false;
    if (count == 0) return true;
    HeightProposal::Impl& m = model.impl();
    if (xz_dev == nullptr || starts_dev == nullptr || count > m.max_queries) return false;
    batch.level(0).device().make_current();
    const auto* levels = static_cast<const ocean_hip_level*>(batch.device_levels()) + first;
    const auto s = static_cast<hipStream_t>(batch.stream());
    if (!wait_for(&model, s)) return false;
    if (ocean_hip_mlp_spacing_partial(levels, ProposalShape::levels, m.spacing_partial, s) != hipSuccess ||
        ocean_hip_mlp_propose(levels, ProposalShape::levels, xz_dev, count, &m.mlp, m.spacing_partial,
                              m.spacing, m.coarse, starts_dev, s) != hipSuccess)
        return false;
    if (spacing_dev != nullptr &&
        hipMemcpyAsync(spacing_dev, m.spacing, ProposalShape::levels * sizeof(std::int32_t),
                       hipMemcpyDeviceToDevice, s) != hipSuccess)
        return false;
    return hipEventRecord(m.done, s) == hipSuccess && hipStreamWaitEvent(m.stream.s, m.done, 0) == hipSuccess;
}

bool proposal_pairs(const OceanBatch& batch, std::size_t first, const float* xz_dev,
                    std::uint32_t count, float tol, float* feats_dev, float* targets_dev,
                    std::uint8_t* ok_dev, std::int32_t* spacing_dev) noexcept
{
    if (first > batch.level_count() || ProposalShape::levels > batch.level_count() - first) return false;
    if (count == 0) return true;
    if (xz_dev == nullptr || feats_dev == nullptr || targets_dev == nullptr || ok_dev == nullptr ||
        spacing_dev == nullptr || !(tol > 0.0f))
        return false;
    batch.level(0).device().make_current();
    const auto* levels = static_cast<const ocean_hip_level*>(batch.device_levels()) + first;
    const auto s = static_cast<hipStream_t>(batch.stream());
    if (ocean_hip_mlp_spacing(levels, ProposalShape::levels, spacing_dev, nullptr, s) != hipSuccess)
        return false;
    return ocean_hip_mlp_pairs(levels, xz_dev, count, tol, spacing_dev, feats_dev, targets_dev, ok_dev,
                               s) == hipSuccess;
}

}  // namespace ocean

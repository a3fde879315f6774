// Converged height queries and the learned start-point proposal (no reference
// counterpart; the reference's height_at takes four fixed-point steps,
// src/ocean.cpp:375-400, and reports its accuracy in ARCHITECTURE.md:620-632).
//
// A converged query inverts each level's displacement with the reference's
// fixed-point map until the step is at most `tol` metres (at most 64 taps per
// level) and sums the levels' heights at their solved points, as a cascade
// does (src/cascade.cpp:118-130). With a HeightProposal, each of the first
// three levels starts from the network's proposal instead of the query point;
// the plain iteration refines it to the same tolerance and restarts from the
// query point if its step does not fall.
//
// The network: 60 features -> 64 -> 64 -> 6 (ReLU), bf16 weights and
// activations on MFMA with fp32 accumulation. Features of level l: the
// (dx, dz) displacement of the 3x3 grid nodes at spacing s_l cells around the
// node nearest the query, in units of s_l cells, then the query's offset
// from that node in the same units; s_l is the power of two in 1..16 nearest
// (in log scale) to 2 rms_l / cell_l, from the level's displacement RMS.
// Outputs: each level's start offset (u - x, v - z), in units of s_l cells.
// It is trained on the device by train_step() on pairs from pairs(), which
// solves with the plain iteration.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>

namespace ocean {

class CascadeStack;
class OceanBatch;

struct ProposalShape {
    static constexpr std::uint32_t levels = 3;
    static constexpr std::uint32_t inputs = 64;    // 60 used, 4 zero
    static constexpr std::uint32_t hidden = 64;
    static constexpr std::uint32_t outputs = 16;   // 6 used, 10 zero
    // fp32 parameters in this order: W1 [inputs][hidden], b1 [hidden],
    // W2 [hidden][hidden], b2, W3 [hidden][outputs], b3 [outputs].
    static constexpr std::uint32_t params = 9360;
    static constexpr std::uint32_t max_batch = 8192;  // train_step rows per call
};

class HeightProposal {
public:
    // `params`: ProposalShape::params floats, or nullptr for all zero.
    // `max_queries`: the largest query count a converged query may ask of it.
    HeightProposal(const float* params, std::uint32_t max_queries);
    ~HeightProposal();
    HeightProposal(HeightProposal&&) noexcept;
    HeightProposal& operator=(HeightProposal&&) noexcept;
    HeightProposal(const HeightProposal&) = delete;
    HeightProposal& operator=(const HeightProposal&) = delete;

    // Replaces the parameters and clears the optimizer state.
    bool set_params(const float* params) noexcept;
    bool get_params(float* out) const noexcept;
    std::uint32_t max_queries() const noexcept;

    // rows x inputs fp32 features (device) -> rows x outputs fp32 (device).
    bool forward(const float* feats_dev, std::uint32_t rows, float* out_dev) const noexcept;
    // One Adam step (lr) on the minibatch rows_dev[batch] of feats/targets
    // (batch a multiple of 128, at most max_batch); the loss is the mean
    // squared error over the 6 outputs; *loss (may be null) gets it, which
    // waits for the step.
    bool train_step(const float* feats_dev, const float* targets_dev, const std::uint32_t* rows_dev,
                    std::uint32_t batch, float lr, float* loss) noexcept;
    // The same minibatch's gradient (ProposalShape::params floats, host) and
    // loss, without a step.
    bool gradient(const float* feats_dev, const float* targets_dev, const std::uint32_t* rows_dev,
                  std::uint32_t batch, float* grad, float* loss) noexcept;
    void synchronize() const noexcept;
    [[nodiscard]] int last_error() const noexcept;

    struct Impl;
    [[nodiscard]] Impl& impl() const noexcept;

private:
    std::unique_ptr<Impl> impl_;
};

// Converged heights at `count` device (x, z) pairs over levels [first,
// first + level_count) of a batch, or over a cascade, enqueued on the owner's
// stream (which first waits for the proposal's own work). `model` may be null
// (plain); with a model, level_count <= 3 and count <= its max_queries.
// iters_dev (count * level_count bytes: taps per level) and flags_dev (count
// bytes: 1 = a level reached 64 taps, 2 = a proposal was abandoned) may be
// null.
bool height_converged(const OceanBatch& batch, std::size_t first, std::size_t level_count,
                      const float* xz_dev, std::uint32_t count, float tol,
                      const HeightProposal* model, float* heights_dev, std::uint8_t* iters_dev,
                      std::uint8_t* flags_dev) noexcept;
bool height_converged(const CascadeStack& stack, const float* xz_dev, std::uint32_t count, float tol,
                      const HeightProposal* model, float* heights_dev, std::uint8_t* iters_dev,
                      std::uint8_t* flags_dev) noexcept;
// The batch form above with the proposal's per-call scratch taken from a
// library-wide static device array instead of the model's own, so the call
// allocates nothing: count <= 262144 (OCEAN_AMD_STATIC_MAX_QUERIES); calls
// that overlap on one GPU share that array and race. *error receives the HIP
// error of the array's one-time address lookup on this device, if any.
bool height_converged_static(const OceanBatch& batch, std::size_t first, std::size_t level_count,
                             const float* xz_dev, std::uint32_t count, float tol,
                             const HeightProposal* model, float* heights_dev,
                             std::uint8_t* iters_dev, std::uint8_t* flags_dev,
                             int* error) noexcept;

// The network's proposal alone for batch levels [first, first + 3): each
// level's proposed start offset from the query point, in cells, level-major
// (starts_dev[2 * (lv * count + q)] = du, [+1] = dv), and the stencil
// spacings used (spacing_dev, 3 int32, may be null). count <= the model's
// max_queries. Enqueued on the batch's stream.
bool proposal_starts(const OceanBatch& batch, std::size_t first, const float* xz_dev,
                     std::uint32_t count, const HeightProposal& model, float* starts_dev,
                     std::int32_t* spacing_dev) noexcept;

// Training pairs from batch levels [first, first + 3): per point the 60
// features (count x inputs floats) and the plain solution's start offsets
// to `tol` (count x outputs floats); ok_dev[q] = 1 when all three converged.
// spacing_dev (3 int32, device) receives the stencil spacings used. Enqueued
// on the batch's stream.
bool proposal_pairs(const OceanBatch& batch, std::size_t first, const float* xz_dev,
                    std::uint32_t count, float tol, float* feats_dev, float* targets_dev,
                    std::uint8_t* ok_dev, std::int32_t* spacing_dev) noexcept;

}  // namespace ocean

// Frame kernels: the row pass and the column pass of one frame
// (frame_device.hpp), launched either per ocean level or batched.
//
// Per level: row_pass_kernel / col_pass_kernel with block shapes chosen per N
// (mirror-paired row teams; C columns per block, lanes column-fastest so each
// global access is C consecutive float2).
//
// Batched: the levels of a batch are grouped by (N, F); each group is stepped
// by one persistent launch per pass. A 256-lane block processes one work item
// at a time - PER = 2048/N rows, or PER columns, of one level - and strides
// through the group's item list by gridDim.x, so the grid size is a tuning
// knob (the autotuner's) that never changes arithmetic. In the batch's
// hipGraph each group is a row -> column chain and the chains of different
// groups are independent, so they overlap on the device.
#include "frame_launch.hpp"

#include <vector>
namespace oceanamd {
namespace {

// The batched kernel of a group's mode (the fast modes' kernels live in
// frame_fast_kernels.hip).
void* batch_kernel(uint32_t n, int fields, bool row, bool per_item = false, uint32_t mode = 0,
                   uint32_t radix = 8, uint32_t layout = 0)
{
    const int r = radix == 0u ? 8 : int(radix);
    if (mode == 1u || mode == 2u) {
        if (layout != 0u) return nullptr;
        return mode == 1u ? ocean_hip_batch_kernel_fast_1(n, fields, row ? 1 : 0, per_item ? 1 : 0, r)
                          : ocean_hip_batch_kernel_fast_2(n, fields, row ? 1 : 0, per_item ? 1 : 0, r);
    }
    if (mode != 0u) return nullptr;
    return batch_kernel_mode<kModeExact>(n, fields, row, per_item, r, int(layout));
}

// Blocks a pass's work takes: an item per block, two with 4 points per lane.
uint32_t pass_units(const ocean_hip_batch_group& g, bool row)
{
    return (row ? g.radix_row : g.radix_col) == 4u ? 2u * g.n_items : g.n_items;
}

// Node parameters of one group's pass; `args` must outlive the call using them.
hipKernelNodeParams group_params(ocean_hip_batch_group& g, bool row, double* t2, void** args)
{
    args[0] = &g.levels;
    args[1] = &g.items;
    args[2] = &g.n_items;
    args[3] = t2;
    hipKernelNodeParams p{};
    const uint32_t grid = row ? g.grid_row : g.grid_col;
    p.func = batch_kernel(g.n, static_cast<int>(g.fields), row, grid >= pass_units(g, row), g.mode,
                          row ? g.radix_row : g.radix_col, g.layout);
    p.gridDim = dim3(grid);
    p.blockDim = dim3(kBlock);
    p.sharedMemBytes = 0;
    p.kernelParams = args;
    p.extra = nullptr;
    return p;
}

}  // namespace
}  // namespace oceanamd

using namespace oceanamd;

namespace {

__global__ __launch_bounds__(kBlock) void wavenumber_kernel(uint32_t n, float patch_length,
                                                            float* __restrict__ kinv,
                                                            float* __restrict__ kx)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n * n) return;
    const uint32_t x = i % n, z = i / n;
    const float kxv = wave_component(x, n, patch_length);
    kinv[i] = inv_wavenumber(kxv, wave_component(z, n, patch_length));
    if (z == 0u) kx[x] = kxv;
}

hipError_t launch_frame(const ocean_hip_level* L, double t2, hipStream_t s, unsigned passes,
                        hipEvent_t* events)
{
    switch ((L->shape >> 16) & 0xffu) {
        case 0: return launch_frame_mode<kModeExact>(L, t2, s, passes, events);
        case 1: return ocean_hip_frame_fast_1(L, t2, s, passes, events);
        case 2: return ocean_hip_frame_fast_2(L, t2, s, passes, events);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

extern "C" hipError_t ocean_hip_wavenumbers(uint32_t n, float patch_length, float* kinv,
                                            float* kx, hipStream_t s)
{
    const uint32_t cells = n * n;
    hipLaunchKernelGGL(wavenumber_kernel, dim3((cells + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       n, patch_length, kinv, kx);
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_frame_passes(const ocean_hip_level* L, double time_over_two_pi,
                                             hipStream_t s, unsigned passes)
{
    return launch_frame(L, time_over_two_pi, s, passes, nullptr);
}

extern "C" hipError_t ocean_hip_frame(const ocean_hip_level* L, double time_over_two_pi,
                                      hipStream_t s)
{
    return launch_frame(L, time_over_two_pi, s, 3u, nullptr);
}

extern "C" hipError_t ocean_hip_frame_timed(const ocean_hip_level* L, double time_over_two_pi,
                                            hipStream_t s, hipEvent_t* events)
{
    return launch_frame(L, time_over_two_pi, s, 3u, events);
}

extern "C" int ocean_hip_frame_shapes(uint32_t n, uint32_t fields, int* row_variants,
                                      int* col_variants)
{
    return ocean_hip_frame_shapes_mode(n, fields, 0u, row_variants, col_variants);
}

extern "C" int ocean_hip_frame_shapes_mode(uint32_t n, uint32_t fields, uint32_t mode,
                                           int* row_variants, int* col_variants)
{
    int r = 0, c = 0, rc = -1;
    if (mode == 0u) rc = frame_shapes_mode<kModeExact>(n, int(fields), r, c);
    if (mode == 1u) rc = ocean_hip_frame_shapes_fast_1(n, fields, &r, &c);
    if (mode == 2u) rc = ocean_hip_frame_shapes_fast_2(n, fields, &r, &c);
    if (rc != 0) return -1;
    if (row_variants) *row_variants = r;
    if (col_variants) *col_variants = c;
    return 0;
}

extern "C" int ocean_hip_frame_shapes_layout(uint32_t n, uint32_t fields, uint32_t mode,
                                             uint32_t layout, int* row_variants, int* col_variants)
{
    if (layout == 0u) return ocean_hip_frame_shapes_mode(n, fields, mode, row_variants, col_variants);
    if (mode != 0u) return -1;
    int r = 0, c = 0;
    if (frame_shapes_mode<kModeExact>(n, int(fields), r, c, int(layout)) != 0) return -1;
    if (row_variants) *row_variants = r;
    if (col_variants) *col_variants = c;
    return 0;
}

extern "C" int ocean_hip_batchable(uint32_t n) { return batch_kernel(n, 4, true) != nullptr; }

extern "C" uint32_t ocean_hip_batch_items(uint32_t n) { return n * n / 2048u; }

extern "C" uint32_t ocean_hip_batch_item(uint32_t level, uint32_t block)
{
    return (level << kItemBlockBits) | block;
}

extern "C" hipError_t ocean_hip_batch_group_frame(const ocean_hip_batch_group* group,
                                                  double time_over_two_pi, unsigned passes,
                                                  hipStream_t s)
{
    ocean_hip_batch_group g = *group;
    if (g.n_items == 0) return hipSuccess;
    for (int pass = 0; pass < 2; ++pass) {
        if (!(passes & (1u << pass))) continue;
        double t2 = time_over_two_pi;
        void* args[4];
        const hipKernelNodeParams p = group_params(g, pass == 0, &t2, args);
        if (p.func == nullptr) return hipErrorInvalidValue;
        const hipError_t e = hipLaunchKernel(p.func, p.gridDim, p.blockDim, p.kernelParams, 0, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// The batch's frame as one instantiated graph: per group a row node and a
// column node depending on it; the only per-frame input, the time, is a
// kernel argument of the row nodes, replaced before each launch.
struct ocean_hip_batch_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<ocean_hip_batch_group> groups;
    std::vector<hipGraphNode_t> row_nodes;
};

extern "C" hipError_t ocean_hip_batch_graph_create(const ocean_hip_batch_group* groups,
                                                   uint32_t count, ocean_hip_batch_graph** out)
{
    auto* g = new ocean_hip_batch_graph();
    g->groups.assign(groups, groups + count);
    hipError_t e = hipGraphCreate(&g->graph, 0);
    for (uint32_t i = 0; i < count && e == hipSuccess; ++i) {
        ocean_hip_batch_group& gr = g->groups[i];
        if (gr.n_items == 0) continue;
        double t2 = 0.0;
        void* rargs[4];
        void* cargs[4];
        const hipKernelNodeParams rp = group_params(gr, true, &t2, rargs);
        const hipKernelNodeParams cp = group_params(gr, false, &t2, cargs);
        if (rp.func == nullptr || cp.func == nullptr) {
            e = hipErrorInvalidValue;
            break;
        }
        hipGraphNode_t rn = nullptr, cn = nullptr;
        e = hipGraphAddKernelNode(&rn, g->graph, nullptr, 0, &rp);
        if (e == hipSuccess) e = hipGraphAddKernelNode(&cn, g->graph, &rn, 1, &cp);
        g->row_nodes.push_back(rn);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        ocean_hip_batch_graph_destroy(g);
        return e;
    }
    *out = g;
    return hipSuccess;
}

extern "C" hipError_t ocean_hip_batch_graph_launch(ocean_hip_batch_graph* g,
                                                   double time_over_two_pi, hipStream_t s)
{
    std::size_t k = 0;
    for (ocean_hip_batch_group& gr : g->groups) {
        if (gr.n_items == 0) continue;
        double t2 = time_over_two_pi;
        void* args[4];
        const hipKernelNodeParams rp = group_params(gr, true, &t2, args);
        const hipError_t e = hipGraphExecKernelNodeSetParams(g->exec, g->row_nodes[k++], &rp);
        if (e != hipSuccess) return e;
    }
    return hipGraphLaunch(g->exec, s);
}

extern "C" void ocean_hip_batch_graph_destroy(ocean_hip_batch_graph* g)
{
    if (g == nullptr) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
}

extern "C" int ocean_hip_batch_occupancy(uint32_t n, uint32_t fields, int row)
{
    void* fn = batch_kernel(n, static_cast<int>(fields), row != 0);
    if (fn == nullptr) return 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fn, kBlock, 0) != hipSuccess) return 0;
    return blocks;
}

extern "C" int ocean_hip_batch_has_launch(uint32_t n, uint32_t fields, uint32_t mode, uint32_t radix,
                                          uint32_t layout)
{
    return batch_kernel(n, static_cast<int>(fields), true, false, mode, radix, layout) != nullptr &&
           batch_kernel(n, static_cast<int>(fields), false, false, mode, radix, layout) != nullptr;
}

// ---- buffer checksums -----------------------------------------------------
// Per level: the output buffers (displacement, normal, then velocity when
// present) as uint32 words w_i, i counted across the buffers in that order,
// summed as sum(w_i * K_i) mod 2^64 with K_i = (i * STEP + BASE) | 1. Addition
// mod 2^64 is commutative, so the block-level split and the atomics give the
// same value as a sequential sum.

namespace oceanamd {
namespace {

constexpr uint64_t kChecksumStep = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kChecksumBase = 0x632BE59BD9B4E019ull;
constexpr uint32_t kChecksumBlocksPerLevel = 32;

__global__ __launch_bounds__(256) void checksum_kernel(const ocean_hip_level* __restrict__ levels,
                                                       unsigned long long* __restrict__ out)
{
    __shared__ unsigned long long part[256];
    const uint32_t lv = blockIdx.x / kChecksumBlocksPerLevel;
    const uint32_t blk = blockIdx.x % kChecksumBlocksPerLevel;
    const ocean_hip_level L = levels[lv];
    const uint64_t per = 4ull * L.n * L.n;  // words per buffer
    const uint32_t nbuf = L.vel != nullptr ? 3u : 2u;
    const uint64_t total = per * nbuf;
    uint64_t sum = 0;
    for (uint64_t i = uint64_t(blk) * 256u + threadIdx.x; i < total;
         i += uint64_t(kChecksumBlocksPerLevel) * 256u) {
        const uint32_t b = uint32_t(i / per);
        const uint64_t j = i - uint64_t(b) * per;
        const float* buf = b == 0 ? L.disp : (b == 1 ? L.normal : L.vel);
        const uint32_t w = reinterpret_cast<const uint32_t*>(buf)[j];
        sum += uint64_t(w) * ((i * kChecksumStep + kChecksumBase) | 1ull);
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out + lv, part[0]);
}

__global__ void zero_u64_kernel(unsigned long long* p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

}  // namespace
}  // namespace oceanamd

extern "C" hipError_t ocean_hip_level_checksums(const ocean_hip_level* levels, uint32_t count,
                                                unsigned long long* out, hipStream_t s)
{
    if (count == 0) return hipSuccess;
    if (uint64_t(count) * kChecksumBlocksPerLevel > 0x7FFFFFFFull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(zero_u64_kernel, dim3((count + 255) / 256), dim3(256), 0, s, out, count);
    hipLaunchKernelGGL(checksum_kernel, dim3(count * kChecksumBlocksPerLevel), dim3(256), 0, s,
                       levels, out);
    return hipGetLastError();
}

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
#include "frame_device.hpp"
#include "hip_launch.h"

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <vector>

namespace oceanamd {
namespace {

constexpr int kBlock = 256;

template <int N>
struct RowShape {
    static constexpr int T = FftShape<N>::T;
    static constexpr int B0 = T > kBlock ? T : kBlock;
    static constexpr int BLOCK = B0 < N * T ? B0 : N * T;  // never more teams than rows
    static constexpr int TEAMS = BLOCK / T;
};

template <int N>
struct ColShape {
    static constexpr int T = FftShape<N>::T;
    static constexpr int C0 = (kBlock / T) > 8 ? (kBlock / T) : 8;
    static constexpr int C1 = C0 < N ? C0 : N;
    static constexpr int C2 = (C1 * T) > 512 ? (512 / T) : C1;  // <= 512 lanes: no spills at F=6
    static constexpr int C = (C2 * N * 8) > 65536 ? (65536 / (N * 8)) : C2;
    static constexpr int BLOCK = C * T;
};

// Workgroups go to the 8 XCDs round-robin by block id, and each XCD has its
// own L2. Column blocks read C*8 bytes of every row, so blocks of adjacent
// column groups share 128-byte lines (and row blocks read each other's mirror
// rows); mapping block b to group (b % 8) * (G / 8) + b / 8 keeps each XCD on
// a contiguous range of groups, so a shared line is fetched from HBM once
// instead of once per XCD.
__device__ __forceinline__ uint32_t xcd_group(uint32_t b, uint32_t g)
{
    if (g % 8u != 0u) return b;
    return (b % 8u) * (g / 8u) + b / 8u;
}

template <int N, int F, int TEAMS>
__global__ __launch_bounds__(TEAMS * FftShape<N>::T) void row_pass_kernel(ocean_hip_level L,
                                                                          double time_over_two_pi)
{
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[TEAMS * team_stride<N>()];
    row_block<N, F, TEAMS>(L, xcd_group(blockIdx.x, gridDim.x), time_over_two_pi, tw, xch);
}

template <int N, int F, int C>
__global__ __launch_bounds__(C * FftShape<N>::T) void col_pass_kernel(ocean_hip_level L)
{
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[C * team_stride<N>()];
    col_block<N, F, C>(L, xcd_group(blockIdx.x, gridDim.x), tw, xch);
}

template <int N, int F, int TEAMS, int KX>
__global__ __launch_bounds__(TEAMS * (F / 2) * FftShape<N>::T) void row_split_kernel(
    ocean_hip_level L, double time_over_two_pi)
{
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[TEAMS * (F / 2) * KX * team_stride<N>()];
    row_block_split<N, F, TEAMS, KX>(L, xcd_group(blockIdx.x, gridDim.x), time_over_two_pi, tw,
                                     xch);
}

template <int N, int F, int C, int KX>
__global__ __launch_bounds__(C * (F / 2) * FftShape<N>::T) void col_split_kernel(ocean_hip_level L)
{
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[C * (F / 2) * KX * team_stride<N>()];
    __shared__ float2 xfer[C * 2 * N];
    col_block_split<N, F, C, KX>(L, xcd_group(blockIdx.x, gridDim.x), tw, xch, xfer);
}

// Launch shapes of the per-level passes. Variant 0 is the default shape;
// the others (rows per block TEAMS, columns per block C, each 1..32) are what
// ocean_hip_frame_shapes lists for the autotuner. Every shape runs the same
// row_block / col_block code on the same cells, so results do not depend on
// the shape.
template <int N, int TEAMS>
constexpr bool row_ok()
{
    constexpr int lanes = TEAMS * FftShape<N>::T;
    return TEAMS <= N && lanes <= 512 && lanes >= 64 && lanes % 64 == 0;  // 1024 lanes spill at F = 6
}
template <int N, int C>
constexpr bool col_ok()
{
    constexpr int lanes = C * FftShape<N>::T;
    return C <= N && lanes <= 512 && lanes >= 64 && lanes % 64 == 0 &&
           C * team_stride<N>() * 8 + N * 4 <= 65536;
}

// Field-split shapes (row_split_kernel / col_split_kernel): whole-wavefront
// teams, at most 512 lanes, static LDS within 64 KB. KX = fields a team
// exchanges between the same barriers (1: one field after the other, 2: both).
template <int N, int F, int TEAMS, int KX>
constexpr bool row_split_ok()
{
    constexpr int lanes = TEAMS * (F / 2) * FftShape<N>::T;
    return FftShape<N>::T >= 64 && TEAMS <= N && lanes <= 512 &&
           (N / 2 + TEAMS * (F / 2) * KX * team_stride<N>()) * 8 <= 65536;
}
template <int N, int F, int C, int KX>
constexpr bool col_split_ok()
{
    constexpr int lanes = C * (F / 2) * FftShape<N>::T;
    return FftShape<N>::T >= 64 && C <= N && lanes <= 512 &&
           (N / 2 + C * (F / 2) * KX * team_stride<N>() + C * 2 * N) * 8 <= 65536;
}

struct Pick {
    int want, seen;
};

// Where a pass is enqueued: the stream, and optionally a pair of events that
// the dispatch itself records at its start and end (hipExtLaunchKernel), so
// the kernel is timed without marker packets between kernels.
struct Launch {
    hipStream_t s = nullptr;
    hipEvent_t start = nullptr, stop = nullptr;
};

template <typename Kernel, typename... Args>
void launch_on(const Launch& k, Kernel kernel, dim3 grid, dim3 block, Args... args)
{
    if (k.start != nullptr)
        hipExtLaunchKernelGGL(kernel, grid, block, 0, k.s, k.start, k.stop, 0u, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, 0, k.s, args...);
}

template <int N, int F, int TEAMS, int KX>
bool try_row_split(Pick& p, const ocean_hip_level& L, double t2, const Launch& k)
{
    if constexpr (!row_split_ok<N, F, TEAMS, KX>()) {
        return false;
    } else {
        if (p.seen++ != p.want) return false;
        if (k.s != nullptr || L.n != 0)
            launch_on(k, row_split_kernel<N, F, TEAMS, KX>, dim3(N / TEAMS),
                      dim3(TEAMS * (F / 2) * FftShape<N>::T), L, t2);
        return true;
    }
}

template <int N, int F, int C, int KX>
bool try_col_split(Pick& p, const ocean_hip_level& L, const Launch& k)
{
    if constexpr (!col_split_ok<N, F, C, KX>()) {
        return false;
    } else {
        if (p.seen++ != p.want) return false;
        if (k.s != nullptr || L.n != 0)
            launch_on(k, col_split_kernel<N, F, C, KX>, dim3(N / C), dim3(C * (F / 2) * FftShape<N>::T),
                      L);
        return true;
    }
}

template <int N, int F, int TEAMS>
bool try_row(Pick& p, const ocean_hip_level& L, double t2, const Launch& k, bool is_default)
{
    if constexpr (!row_ok<N, TEAMS>() && TEAMS != RowShape<N>::TEAMS) {
        return false;
    } else {
        if (!is_default && TEAMS == RowShape<N>::TEAMS) return false;  // listed first already
        if (p.seen++ != p.want) return false;
        if (k.s != nullptr || L.n != 0)
            launch_on(k, row_pass_kernel<N, F, TEAMS>, dim3(N / TEAMS), dim3(TEAMS * FftShape<N>::T),
                      L, t2);
        return true;
    }
}

template <int N, int F, int C>
bool try_col(Pick& p, const ocean_hip_level& L, const Launch& k, bool is_default)
{
    if constexpr (!col_ok<N, C>() && C != ColShape<N>::C) {
        return false;
    } else {
        if (!is_default && C == ColShape<N>::C) return false;
        if (p.seen++ != p.want) return false;
        if (k.s != nullptr || L.n != 0)
            launch_on(k, col_pass_kernel<N, F, C>, dim3(N / C), dim3(C * FftShape<N>::T), L);
        return true;
    }
}

template <int N, int F>
bool row_variant(int v, const ocean_hip_level& L, double t2, const Launch& k)
{
    Pick p{v, 0};
    return try_row<N, F, RowShape<N>::TEAMS>(p, L, t2, k, true) || try_row<N, F, 1>(p, L, t2, k, false) ||
           try_row<N, F, 2>(p, L, t2, k, false) || try_row<N, F, 4>(p, L, t2, k, false) ||
           try_row<N, F, 8>(p, L, t2, k, false) || try_row<N, F, 16>(p, L, t2, k, false) ||
           try_row<N, F, 32>(p, L, t2, k, false) || try_row_split<N, F, 1, 1>(p, L, t2, k) ||
           try_row_split<N, F, 2, 1>(p, L, t2, k) || try_row_split<N, F, 4, 1>(p, L, t2, k) ||
           try_row_split<N, F, 1, 2>(p, L, t2, k) || try_row_split<N, F, 2, 2>(p, L, t2, k);
}

template <int N, int F>
bool col_variant(int v, const ocean_hip_level& L, const Launch& k)
{
    Pick p{v, 0};
    return try_col<N, F, ColShape<N>::C>(p, L, k, true) || try_col<N, F, 1>(p, L, k, false) ||
           try_col<N, F, 2>(p, L, k, false) || try_col<N, F, 4>(p, L, k, false) ||
           try_col<N, F, 8>(p, L, k, false) || try_col<N, F, 16>(p, L, k, false) ||
           try_col<N, F, 32>(p, L, k, false) || try_col_split<N, F, 1, 1>(p, L, k) ||
           try_col_split<N, F, 2, 1>(p, L, k) || try_col_split<N, F, 4, 1>(p, L, k) ||
           try_col_split<N, F, 1, 2>(p, L, k) || try_col_split<N, F, 2, 2>(p, L, k);
}

// Number of variants of each pass (counted without launching: a null level
// view with n = 0 and no stream).
template <int N, int F>
void count_variants_f(int& rows, int& cols)
{
    ocean_hip_level none{};
    rows = 0;
    while (row_variant<N, F>(rows, none, 0.0, Launch{})) ++rows;
    cols = 0;
    while (col_variant<N, F>(cols, none, Launch{})) ++cols;
}

template <int N>
void count_variants(int fields, int& rows, int& cols)
{
    if (fields == 6)
        count_variants_f<N, 6>(rows, cols);
    else
        count_variants_f<N, 4>(rows, cols);
}

template <int N>
hipError_t launch_frame_n(const ocean_hip_level* L, double time_over_two_pi, hipStream_t s,
                          unsigned passes, hipEvent_t* events)
{
    const int rv = int(L->shape & 0xffu), cv = int((L->shape >> 8) & 0xffu);
    const Launch row{s, events ? events[0] : nullptr, events ? events[1] : nullptr};
    const Launch col{s, events ? events[2] : nullptr, events ? events[3] : nullptr};
    bool ok = true;
    if (L->vel != nullptr) {
        if (passes & 1u) ok = ok && row_variant<N, 6>(rv, *L, time_over_two_pi, row);
        if (passes & 2u) ok = ok && col_variant<N, 6>(cv, *L, col);
    } else {
        if (passes & 1u) ok = ok && row_variant<N, 4>(rv, *L, time_over_two_pi, row);
        if (passes & 2u) ok = ok && col_variant<N, 4>(cv, *L, col);
    }
    if (!ok) return hipErrorInvalidValue;
    return hipGetLastError();
}

// ---- batched persistent passes (N in [64, 2048], 256-lane blocks) --------

constexpr uint32_t kItemBlockBits = 12;
constexpr uint32_t kItemBlockMask = (1u << kItemBlockBits) - 1u;

// The lane index, re-read for every item of a persistent loop: otherwise the
// compiler hoists each lane's address arithmetic for all of an item's loads
// and stores out of the loop and keeps it live in registers across it
// (231 instead of 157 VGPRs for the 512^2 column pass), halving occupancy.
__device__ __forceinline__ uint32_t item_lane()
{
    uint32_t t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

template <int N>
struct BatchShape {
    static constexpr int PER = 2048 / N;  // rows (row pass) or columns (column pass) per item
    static_assert(PER * FftShape<N>::T == kBlock, "batched blocks are 256 lanes");
};

// Field-split items (two teams per row or column, as in the single-level
// split kernels) were measured slower here: the batch is throughput-bound,
// and the split repeats the per-cell phase and sincos in both teams.
template <int N, int F>
__global__ __launch_bounds__(kBlock) void batch_row_kernel(const ocean_hip_level* __restrict__ levels,
                                                           const uint32_t* __restrict__ items,
                                                           uint32_t n_items, double t2)
{
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[BatchShape<N>::PER * team_stride<N>()];
    for (uint32_t i = blockIdx.x; i < n_items; i += gridDim.x) {
        const uint32_t it = items[i];
        row_block<N, F, BatchShape<N>::PER>(levels[it >> kItemBlockBits], it & kItemBlockMask, t2,
                                            tw, xch, item_lane());
    }
}

template <int N, int F>
__global__ __launch_bounds__(kBlock) void batch_col_kernel(const ocean_hip_level* __restrict__ levels,
                                                           const uint32_t* __restrict__ items,
                                                           uint32_t n_items)
{
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[BatchShape<N>::PER * team_stride<N>()];
    for (uint32_t i = blockIdx.x; i < n_items; i += gridDim.x) {
        const uint32_t it = items[i];
        col_block<N, F, BatchShape<N>::PER, true>(levels[it >> kItemBlockBits],
                                                  it & kItemBlockMask, tw, xch, item_lane());
    }
}

// The same passes with one item per block, for a grid of one block per item
// (what the autotuner picks for the large groups): without the item loop the
// compiler needs fewer registers (512^2 column pass 157 instead of 187 VGPRs,
// 3 waves per SIMD instead of 2).
template <int N, int F>
__global__ __launch_bounds__(kBlock) void batch_row_item_kernel(
    const ocean_hip_level* __restrict__ levels, const uint32_t* __restrict__ items,
    uint32_t n_items, double t2)
{
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[BatchShape<N>::PER * team_stride<N>()];
    if (blockIdx.x >= n_items) return;
    const uint32_t it = items[blockIdx.x];
    row_block<N, F, BatchShape<N>::PER>(levels[it >> kItemBlockBits], it & kItemBlockMask, t2, tw,
                                        xch);
}

template <int N, int F>
__global__ __launch_bounds__(kBlock) void batch_col_item_kernel(
    const ocean_hip_level* __restrict__ levels, const uint32_t* __restrict__ items,
    uint32_t n_items)
{
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[BatchShape<N>::PER * team_stride<N>()];
    if (blockIdx.x >= n_items) return;
    const uint32_t it = items[blockIdx.x];
    col_block<N, F, BatchShape<N>::PER, true>(levels[it >> kItemBlockBits], it & kItemBlockMask,
                                              tw, xch);
}

template <int N, int F>
void* batch_fn_f(bool row, bool per_item)
{
    if (per_item)
        return row ? reinterpret_cast<void*>(batch_row_item_kernel<N, F>)
                   : reinterpret_cast<void*>(batch_col_item_kernel<N, F>);
    return row ? reinterpret_cast<void*>(batch_row_kernel<N, F>)
               : reinterpret_cast<void*>(batch_col_kernel<N, F>);
}

template <int N>
void* batch_fn(int fields, bool row, bool per_item)
{
    return fields == 6 ? batch_fn_f<N, 6>(row, per_item) : batch_fn_f<N, 4>(row, per_item);
}

// per_item: the grid has a block for every item (the loop-free kernels).
void* batch_kernel(uint32_t n, int fields, bool row, bool per_item = false)
{
    switch (n) {
        case 64: return batch_fn<64>(fields, row, per_item);
        case 128: return batch_fn<128>(fields, row, per_item);
        case 256: return batch_fn<256>(fields, row, per_item);
        case 512: return batch_fn<512>(fields, row, per_item);
        case 1024: return batch_fn<1024>(fields, row, per_item);
        case 2048: return batch_fn<2048>(fields, row, per_item);
        default: return nullptr;
    }
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
    p.func = batch_kernel(g.n, static_cast<int>(g.fields), row, grid >= g.n_items);
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
    switch (L->n) {
        case 16: return launch_frame_n<16>(L, t2, s, passes, events);
        case 32: return launch_frame_n<32>(L, t2, s, passes, events);
        case 64: return launch_frame_n<64>(L, t2, s, passes, events);
        case 128: return launch_frame_n<128>(L, t2, s, passes, events);
        case 256: return launch_frame_n<256>(L, t2, s, passes, events);
        case 512: return launch_frame_n<512>(L, t2, s, passes, events);
        case 1024: return launch_frame_n<1024>(L, t2, s, passes, events);
        case 2048: return launch_frame_n<2048>(L, t2, s, passes, events);
        case 4096: return launch_frame_n<4096>(L, t2, s, passes, events);
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
    int r = 0, c = 0;
    const int f = int(fields);
    switch (n) {
        case 16: count_variants<16>(f, r, c); break;
        case 32: count_variants<32>(f, r, c); break;
        case 64: count_variants<64>(f, r, c); break;
        case 128: count_variants<128>(f, r, c); break;
        case 256: count_variants<256>(f, r, c); break;
        case 512: count_variants<512>(f, r, c); break;
        case 1024: count_variants<1024>(f, r, c); break;
        case 2048: count_variants<2048>(f, r, c); break;
        case 4096: count_variants<4096>(f, r, c); break;
        default: return -1;
    }
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

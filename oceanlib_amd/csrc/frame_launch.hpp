// Launch machinery of the frame passes, shared by the bit-exact kernels
// (frame_kernels.hip) and the opt-in fast modes (frame_fast_kernels.hip):
// kernel templates over the block code of frame_device.hpp with a precision
// mode M (kModeExact / kModeFast / kModeFastHalf), the per-level launch
// shapes and their variant lists, and the batched persistent kernels. Each
// translation unit instantiates the modes it serves.
#pragma once

#include "frame_device.hpp"
#include "hip_launch.h"

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

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

// R: points per lane of the FFT (its radix grouping, fft_team.hpp).
// R: points per lane of the FFT (its radix grouping, fft_team.hpp); W: the
// layout of the fields between the passes (spec_at, frame_device.hpp).
template <int N, int F, int TEAMS, int MODE, int R, int W>
__global__ __launch_bounds__((TEAMS * FftShape<N, R>::T)) void row_pass_kernel(ocean_hip_level L,
                                                                             double time_over_two_pi)
{
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[TEAMS * team_stride<N>()];
    row_block<N, F, TEAMS, MODE, R, W>(L, xcd_group(blockIdx.x, gridDim.x), time_over_two_pi, tw,
                                       xch);
}

template <int N, int F, int C, int MODE, int R, int W>
__global__ __launch_bounds__((C * FftShape<N, R>::T)) void col_pass_kernel(ocean_hip_level L)
{
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[C * team_stride<N>()];
    col_block<N, F, C, false, MODE, R, W>(L, xcd_group(blockIdx.x, gridDim.x), tw, xch);
}

template <int N, int F, int TEAMS, int KX, int MODE, int R, int W>
__global__ __launch_bounds__((TEAMS * (F / 2) * FftShape<N, R>::T)) void row_split_kernel(
    ocean_hip_level L, double time_over_two_pi)
{
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[TEAMS * (F / 2) * KX * team_stride<N>()];
    row_block_split<N, F, TEAMS, KX, MODE, R, W>(L, xcd_group(blockIdx.x, gridDim.x),
                                                 time_over_two_pi, tw, xch);
}

template <int N, int F, int C, int KX, int MODE, int R, int W>
__global__ __launch_bounds__((C * (F / 2) * FftShape<N, R>::T)) void col_split_kernel(ocean_hip_level L)
{
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[C * (F / 2) * KX * team_stride<N>()];
    __shared__ float2 xfer[C * 2 * N];
    col_block_split<N, F, C, KX, MODE, R, W>(L, xcd_group(blockIdx.x, gridDim.x), tw, xch, xfer);
}

// Launch shapes of the per-level passes. Variant 0 is the default shape;
// the others (rows per block TEAMS, columns per block C, each 1..32, the
// field split, and the FFT's radix grouping R = 4 / 8 / 16 points per lane)
// are what ocean_hip_frame_shapes lists for the autotuner. Every shape runs
// the same butterflies in the same order on the same cells (only where the
// values live between stages changes), so results do not depend on the shape.
template <int N, int TEAMS, int R = 8>
constexpr bool row_ok()
{
    if constexpr (N < 2 * R) {
        return false;
    } else {
        constexpr int lanes = TEAMS * FftShape<N, R>::T;
        return TEAMS <= N && lanes <= 512 && lanes >= 64 && lanes % 64 == 0;  // 1024 lanes spill at F = 6
    }
}
template <int N, int C, int R = 8>
constexpr bool col_ok()
{
    if constexpr (N < 2 * R) {
        return false;
    } else {
        constexpr int lanes = C * FftShape<N, R>::T;
        return C <= N && lanes <= 512 && lanes >= 64 && lanes % 64 == 0 &&
               C * team_stride<N>() * 8 + N * 4 <= 65536;
    }
}

// Field-split shapes (row_split_kernel / col_split_kernel): whole-wavefront
// teams, at most 512 lanes, static LDS within 64 KB. KX = fields a team
// exchanges between the same barriers (1: one field after the other, 2: both).
template <int N, int F, int TEAMS, int KX, int R = 8>
constexpr bool row_split_ok()
{
    if constexpr (N < 64 * R) {
        return false;
    } else {
        constexpr int lanes = TEAMS * (F / 2) * FftShape<N, R>::T;
        return TEAMS <= N && lanes <= 512 &&
               (N / 2 + TEAMS * (F / 2) * KX * team_stride<N>()) * 8 <= 65536;
    }
}
template <int N, int F, int C, int KX, int R = 8>
constexpr bool col_split_ok()
{
    if constexpr (N < 64 * R) {
        return false;
    } else {
        constexpr int lanes = C * (F / 2) * FftShape<N, R>::T;
        return C <= N && lanes <= 512 &&
               (N / 2 + C * (F / 2) * KX * team_stride<N>() + C * 2 * N) * 8 <= 65536;
    }
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

template <int N, int F, int TEAMS, int KX, int MODE, int R = 8, int W = 0>
bool try_row_split(Pick& p, const ocean_hip_level& L, double t2, const Launch& k)
{
    if constexpr (!row_split_ok<N, F, TEAMS, KX, R>()) {
        return false;
    } else {
        if (p.seen++ != p.want) return false;
        if (k.s != nullptr || L.n != 0)
            launch_on(k, row_split_kernel<N, F, TEAMS, KX, MODE, R, W>, dim3(N / TEAMS),
                      dim3(TEAMS * (F / 2) * FftShape<N, R>::T), L, t2);
        return true;
    }
}

template <int N, int F, int C, int KX, int MODE, int R = 8, int W = 0>
bool try_col_split(Pick& p, const ocean_hip_level& L, const Launch& k)
{
    if constexpr (!col_split_ok<N, F, C, KX, R>()) {
        return false;
    } else {
        if (p.seen++ != p.want) return false;
        if (k.s != nullptr || L.n != 0)
            launch_on(k, col_split_kernel<N, F, C, KX, MODE, R, W>, dim3(N / C),
                      dim3(C * (F / 2) * FftShape<N, R>::T), L);
        return true;
    }
}

// The fast modes list only the default shape and the field-split shapes (the
// ones the autotuner picks at N >= 512), which keeps their instantiations few.
template <int N, int F, int TEAMS, int MODE, int W = 0>
bool try_row(Pick& p, const ocean_hip_level& L, double t2, const Launch& k, bool is_default)
{
    if constexpr ((!row_ok<N, TEAMS>() || MODE != kModeExact) && TEAMS != RowShape<N>::TEAMS) {
        return false;
    } else {
        if (!is_default && TEAMS == RowShape<N>::TEAMS) return false;  // listed first already
        if (p.seen++ != p.want) return false;
        if (k.s != nullptr || L.n != 0)
            launch_on(k, row_pass_kernel<N, F, TEAMS, MODE, 8, W>, dim3(N / TEAMS),
                      dim3(TEAMS * FftShape<N>::T), L, t2);
        return true;
    }
}

template <int N, int F, int C, int MODE, int W = 0>
bool try_col(Pick& p, const ocean_hip_level& L, const Launch& k, bool is_default)
{
    if constexpr ((!col_ok<N, C>() || MODE != kModeExact) && C != ColShape<N>::C) {
        return false;
    } else {
        if (!is_default && C == ColShape<N>::C) return false;
        if (p.seen++ != p.want) return false;
        if (k.s != nullptr || L.n != 0)
            launch_on(k, col_pass_kernel<N, F, C, MODE, 8, W>, dim3(N / C), dim3(C * FftShape<N>::T), L);
        return true;
    }
}

// Other radix groupings of the unsplit passes (bit-exact mode).
template <int N, int F, int TEAMS, int MODE, int R, int W = 0>
bool try_row_r(Pick& p, const ocean_hip_level& L, double t2, const Launch& k)
{
    if constexpr (!row_ok<N, TEAMS, R>() || MODE != kModeExact) {
        return false;
    } else {
        if (p.seen++ != p.want) return false;
        if (k.s != nullptr || L.n != 0)
            launch_on(k, row_pass_kernel<N, F, TEAMS, MODE, R, W>, dim3(N / TEAMS),
                      dim3(TEAMS * FftShape<N, R>::T), L, t2);
        return true;
    }
}

template <int N, int F, int C, int MODE, int R, int W = 0>
bool try_col_r(Pick& p, const ocean_hip_level& L, const Launch& k)
{
    if constexpr (!col_ok<N, C, R>() || MODE != kModeExact) {
        return false;
    } else {
        if (p.seen++ != p.want) return false;
        if (k.s != nullptr || L.n != 0)
            launch_on(k, col_pass_kernel<N, F, C, MODE, R, W>, dim3(N / C), dim3(C * FftShape<N, R>::T),
                      L);
        return true;
    }
}

template <int N, int F, int M, int W = 0>
bool row_variant(int v, const ocean_hip_level& L, double t2, const Launch& k)
{
    Pick p{v, 0};
    return try_row<N, F, RowShape<N>::TEAMS, M, W>(p, L, t2, k, true) ||
           try_row<N, F, 1, M, W>(p, L, t2, k, false) || try_row<N, F, 2, M, W>(p, L, t2, k, false) ||
           try_row<N, F, 4, M, W>(p, L, t2, k, false) || try_row<N, F, 8, M, W>(p, L, t2, k, false) ||
           try_row<N, F, 16, M, W>(p, L, t2, k, false) || try_row<N, F, 32, M, W>(p, L, t2, k, false) ||
           try_row_split<N, F, 1, 1, M, 8, W>(p, L, t2, k) ||
           try_row_split<N, F, 2, 1, M, 8, W>(p, L, t2, k) ||
           try_row_split<N, F, 4, 1, M, 8, W>(p, L, t2, k) ||
           try_row_split<N, F, 1, 2, M, 8, W>(p, L, t2, k) ||
           try_row_split<N, F, 2, 2, M, 8, W>(p, L, t2, k) ||
           // 4 points per lane: twice the lanes, stages in pairs
           try_row_r<N, F, 2, M, 4, W>(p, L, t2, k) || try_row_r<N, F, 4, M, 4, W>(p, L, t2, k) ||
           try_row_split<N, F, 1, 1, M, 4, W>(p, L, t2, k) ||
           try_row_split<N, F, 2, 1, M, 4, W>(p, L, t2, k) ||
           try_row_split<N, F, 1, 2, M, 4, W>(p, L, t2, k) ||
           // 16 points per lane: half the lanes, stages in fours
           try_row_split<N, F, 1, 1, M, 16, W>(p, L, t2, k) ||
           try_row_split<N, F, 2, 1, M, 16, W>(p, L, t2, k);
}

template <int N, int F, int M, int W = 0>
bool col_variant(int v, const ocean_hip_level& L, const Launch& k)
{
    Pick p{v, 0};
    return try_col<N, F, ColShape<N>::C, M, W>(p, L, k, true) ||
           try_col<N, F, 1, M, W>(p, L, k, false) || try_col<N, F, 2, M, W>(p, L, k, false) ||
           try_col<N, F, 4, M, W>(p, L, k, false) || try_col<N, F, 8, M, W>(p, L, k, false) ||
           try_col<N, F, 16, M, W>(p, L, k, false) || try_col<N, F, 32, M, W>(p, L, k, false) ||
           try_col_split<N, F, 1, 1, M, 8, W>(p, L, k) || try_col_split<N, F, 2, 1, M, 8, W>(p, L, k) ||
           try_col_split<N, F, 4, 1, M, 8, W>(p, L, k) || try_col_split<N, F, 1, 2, M, 8, W>(p, L, k) ||
           try_col_split<N, F, 2, 2, M, 8, W>(p, L, k) ||
           try_col_r<N, F, 2, M, 4, W>(p, L, k) || try_col_r<N, F, 4, M, 4, W>(p, L, k) ||
           try_col_split<N, F, 1, 1, M, 4, W>(p, L, k) || try_col_split<N, F, 2, 1, M, 4, W>(p, L, k) ||
           try_col_split<N, F, 1, 2, M, 4, W>(p, L, k) ||
           try_col_split<N, F, 1, 1, M, 16, W>(p, L, k) ||
           try_col_split<N, F, 2, 1, M, 16, W>(p, L, k);
}

// Layouts of the fields between the passes a level may use (bits 24-31 of
// its shape): 0 row-major; 1, 2, 3 tiles of 2, 4, 8 columns (bit-exact mode,
// N >= 256).
constexpr int kLayouts = 4;
constexpr int tile_width(int layout) { return layout == 0 ? 0 : (1 << layout); }
template <int N, int M>
constexpr bool tiled_ok()
{
    return M == kModeExact && N >= 256;
}

// Number of variants of each pass (counted without launching: a null level
// view with n = 0 and no stream).
template <int N, int F, int M, int W>
void count_variants_f(int& rows, int& cols)
{
    ocean_hip_level none{};
    rows = 0;
    while (row_variant<N, F, M, W>(rows, none, 0.0, Launch{})) ++rows;
    cols = 0;
    while (col_variant<N, F, M, W>(cols, none, Launch{})) ++cols;
}

// false: the layout is not available for (N, M).
template <int N, int M, int W>
void count_variants_w(int fields, int& rows, int& cols)
{
    if (fields == 6)
        count_variants_f<N, 6, M, W>(rows, cols);
    else
        count_variants_f<N, 4, M, W>(rows, cols);
}

template <int N, int M>
bool count_variants(int fields, int& rows, int& cols, int layout = 0)
{
    if (layout == 0) {
        count_variants_w<N, M, 0>(fields, rows, cols);
        return true;
    }
    if constexpr (!tiled_ok<N, M>()) {
        return false;
    } else {
        switch (layout) {
            case 1: count_variants_w<N, M, tile_width(1)>(fields, rows, cols); return true;
            case 2: count_variants_w<N, M, tile_width(2)>(fields, rows, cols); return true;
            case 3: count_variants_w<N, M, tile_width(3)>(fields, rows, cols); return true;
            default: return false;
        }
    }
}

template <int N, int M, int W>
bool launch_passes(const ocean_hip_level* L, int rv, int cv, double t2, unsigned passes,
                   const Launch& row, const Launch& col)
{
    bool ok = true;
    if (L->vel != nullptr) {
        if (passes & 1u) ok = ok && row_variant<N, 6, M, W>(rv, *L, t2, row);
        if (passes & 2u) ok = ok && col_variant<N, 6, M, W>(cv, *L, col);
    } else {
        if (passes & 1u) ok = ok && row_variant<N, 4, M, W>(rv, *L, t2, row);
        if (passes & 2u) ok = ok && col_variant<N, 4, M, W>(cv, *L, col);
    }
    return ok;
}

template <int N, int M>
hipError_t launch_frame_n(const ocean_hip_level* L, double time_over_two_pi, hipStream_t s,
                          unsigned passes, hipEvent_t* events)
{
    const int rv = int(L->shape & 0xffu), cv = int((L->shape >> 8) & 0xffu);
    const Launch row{s, events ? events[0] : nullptr, events ? events[1] : nullptr};
    const Launch col{s, events ? events[2] : nullptr, events ? events[3] : nullptr};
    const int layout = int((L->shape >> 24) & 0xffu);
    bool ok = false;
    if (layout == 0) {
        ok = launch_passes<N, M, 0>(L, rv, cv, time_over_two_pi, passes, row, col);
    } else if constexpr (tiled_ok<N, M>()) {
        switch (layout) {
            case 1: ok = launch_passes<N, M, tile_width(1)>(L, rv, cv, time_over_two_pi, passes, row, col); break;
            case 2: ok = launch_passes<N, M, tile_width(2)>(L, rv, cv, time_over_two_pi, passes, row, col); break;
            case 3: ok = launch_passes<N, M, tile_width(3)>(L, rv, cv, time_over_two_pi, passes, row, col); break;
            default: break;
        }
    }
    if (!ok) return hipErrorInvalidValue;
    return hipGetLastError();
}

// The per-level passes of mode M for any supported N.
template <int M>
hipError_t launch_frame_mode(const ocean_hip_level* L, double t2, hipStream_t s, unsigned passes,
                             hipEvent_t* events)
{
    switch (L->n) {
        case 16: return launch_frame_n<16, M>(L, t2, s, passes, events);
        case 32: return launch_frame_n<32, M>(L, t2, s, passes, events);
        case 64: return launch_frame_n<64, M>(L, t2, s, passes, events);
        case 128: return launch_frame_n<128, M>(L, t2, s, passes, events);
        case 256: return launch_frame_n<256, M>(L, t2, s, passes, events);
        case 512: return launch_frame_n<512, M>(L, t2, s, passes, events);
        case 1024: return launch_frame_n<1024, M>(L, t2, s, passes, events);
        case 2048: return launch_frame_n<2048, M>(L, t2, s, passes, events);
        case 4096: return launch_frame_n<4096, M>(L, t2, s, passes, events);
        default: return hipErrorInvalidValue;
    }
}

template <int M>
int frame_shapes_mode(uint32_t n, int f, int& r, int& c, int layout = 0)
{
    bool ok = false;
    switch (n) {
        case 16: ok = count_variants<16, M>(f, r, c, layout); break;
        case 32: ok = count_variants<32, M>(f, r, c, layout); break;
        case 64: ok = count_variants<64, M>(f, r, c, layout); break;
        case 128: ok = count_variants<128, M>(f, r, c, layout); break;
        case 256: ok = count_variants<256, M>(f, r, c, layout); break;
        case 512: ok = count_variants<512, M>(f, r, c, layout); break;
        case 1024: ok = count_variants<1024, M>(f, r, c, layout); break;
        case 2048: ok = count_variants<2048, M>(f, r, c, layout); break;
        case 4096: ok = count_variants<4096, M>(f, r, c, layout); break;
        default: return -1;
    }
    return ok ? 0 : -1;
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

template <int N, int R = 8>
struct BatchShape {
    static constexpr int PER = kBlock * R / N;  // rows (row pass) or columns (column pass) per block
    static_assert(PER * FftShape<N, R>::T == kBlock, "batched blocks are 256 lanes");
};

// Field-split items (two teams per row or column, as in the single-level
// split kernels) were measured slower here: the batch is throughput-bound,
// and the split repeats the per-cell phase and sincos in both teams.
template <int N, int F, int M, int W>
__global__ __launch_bounds__(kBlock) void batch_row_kernel(const ocean_hip_level* __restrict__ levels,
                                                           const uint32_t* __restrict__ items,
                                                           uint32_t n_items, double t2)
{
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[BatchShape<N>::PER * team_stride<N>()];
    for (uint32_t i = blockIdx.x; i < n_items; i += gridDim.x) {
        const uint32_t it = items[i];
        row_block<N, F, BatchShape<N>::PER, M, 8, W>(levels[it >> kItemBlockBits], it & kItemBlockMask,
                                               t2, tw, xch, item_lane());
    }
}

template <int N, int F, int M, int W>
__global__ __launch_bounds__(kBlock) void batch_col_kernel(const ocean_hip_level* __restrict__ levels,
                                                           const uint32_t* __restrict__ items,
                                                           uint32_t n_items)
{
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[BatchShape<N>::PER * team_stride<N>()];
    for (uint32_t i = blockIdx.x; i < n_items; i += gridDim.x) {
        const uint32_t it = items[i];
        col_block<N, F, BatchShape<N>::PER, true, M, 8, W>(levels[it >> kItemBlockBits],
                                                  it & kItemBlockMask, tw, xch, item_lane());
    }
}

// The same passes with one item per block, for a grid of one block per item
// (what the autotuner picks for the large groups): without the item loop the
// compiler needs fewer registers (512^2 column pass 157 instead of 187 VGPRs,
// 3 waves per SIMD instead of 2).
template <int N, int F, int M, int W>
__global__ __launch_bounds__(kBlock) void batch_row_item_kernel(
    const ocean_hip_level* __restrict__ levels, const uint32_t* __restrict__ items,
    uint32_t n_items, double t2)
{
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[BatchShape<N>::PER * team_stride<N>()];
    if (blockIdx.x >= n_items) return;
    const uint32_t it = items[blockIdx.x];
    row_block<N, F, BatchShape<N>::PER, M, 8, W>(levels[it >> kItemBlockBits], it & kItemBlockMask, t2,
                                           tw, xch);
}

template <int N, int F, int M, int W>
__global__ __launch_bounds__(kBlock) void batch_col_item_kernel(
    const ocean_hip_level* __restrict__ levels, const uint32_t* __restrict__ items,
    uint32_t n_items)
{
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[BatchShape<N>::PER * team_stride<N>()];
    if (blockIdx.x >= n_items) return;
    const uint32_t it = items[blockIdx.x];
    col_block<N, F, BatchShape<N>::PER, true, M, 8, W>(levels[it >> kItemBlockBits], it & kItemBlockMask,
                                                 tw, xch);
}

// 4 FFT points per lane (the radix grouping the batch autotuner may pick per
// pass): a work item's PER = 2048/N rows or columns are done by two blocks of
// PER/2, unit u of 2 n_items taking half u / n_items of item u % n_items, so
// with an item count that is a multiple of 8 (every group of N >= 128) the
// two halves of an item run on the XCD the item's place in the list gives it.
// Every row and column goes through the same butterflies as with 8 points per
// lane.
template <int N, int F, int M, int W>
__global__ __launch_bounds__(kBlock) void batch_row_r4_kernel(
    const ocean_hip_level* __restrict__ levels, const uint32_t* __restrict__ items,
    uint32_t n_items, double t2)
{
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[BatchShape<N, 4>::PER * team_stride<N>()];
    const uint32_t units = 2u * n_items;
    for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
        const uint32_t it = items[u % n_items];
        const uint32_t blk = 2u * (it & kItemBlockMask) + u / n_items;
        row_block<N, F, BatchShape<N, 4>::PER, M, 4, W>(levels[it >> kItemBlockBits], blk, t2, tw, xch,
                                                     item_lane());
    }
}

template <int N, int F, int M, int W>
__global__ __launch_bounds__(kBlock) void batch_col_r4_kernel(
    const ocean_hip_level* __restrict__ levels, const uint32_t* __restrict__ items,
    uint32_t n_items)
{
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[BatchShape<N, 4>::PER * team_stride<N>()];
    const uint32_t units = 2u * n_items;
    for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
        const uint32_t it = items[u % n_items];
        const uint32_t blk = 2u * (it & kItemBlockMask) + u / n_items;
        col_block<N, F, BatchShape<N, 4>::PER, true, M, 4, W>(levels[it >> kItemBlockBits], blk, tw,
                                                           xch, item_lane());
    }
}

template <int N, int F, int M, int W>
__global__ __launch_bounds__(kBlock) void batch_row_r4_item_kernel(
    const ocean_hip_level* __restrict__ levels, const uint32_t* __restrict__ items,
    uint32_t n_items, double t2)
{
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[BatchShape<N, 4>::PER * team_stride<N>()];
    const uint32_t u = blockIdx.x;
    if (u >= 2u * n_items) return;
    const uint32_t it = items[u % n_items];
    row_block<N, F, BatchShape<N, 4>::PER, M, 4, W>(levels[it >> kItemBlockBits],
                                                 2u * (it & kItemBlockMask) + u / n_items, t2, tw,
                                                 xch);
}

template <int N, int F, int M, int W>
__global__ __launch_bounds__(kBlock) void batch_col_r4_item_kernel(
    const ocean_hip_level* __restrict__ levels, const uint32_t* __restrict__ items,
    uint32_t n_items)
{
    __shared__ float2 tw[N / 2];
    __shared__ float2 xch[BatchShape<N, 4>::PER * team_stride<N>()];
    const uint32_t u = blockIdx.x;
    if (u >= 2u * n_items) return;
    const uint32_t it = items[u % n_items];
    col_block<N, F, BatchShape<N, 4>::PER, true, M, 4, W>(levels[it >> kItemBlockBits],
                                                       2u * (it & kItemBlockMask) + u / n_items, tw,
                                                       xch);
}

template <int N, int F, int M, int W>
void* batch_fn_w(bool row, bool per_item, int radix)
{
    if (radix == 4) {
        if constexpr (N > 1024) {
            return nullptr;  // 4 points per lane: N <= 1024 (one row per 256-lane block)
        } else {
            if (per_item)
                return row ? reinterpret_cast<void*>(batch_row_r4_item_kernel<N, F, M, W>)
                           : reinterpret_cast<void*>(batch_col_r4_item_kernel<N, F, M, W>);
            return row ? reinterpret_cast<void*>(batch_row_r4_kernel<N, F, M, W>)
                       : reinterpret_cast<void*>(batch_col_r4_kernel<N, F, M, W>);
        }
    }
    if (radix != 8) return nullptr;
    if (per_item)
        return row ? reinterpret_cast<void*>(batch_row_item_kernel<N, F, M, W>)
                   : reinterpret_cast<void*>(batch_col_item_kernel<N, F, M, W>);
    return row ? reinterpret_cast<void*>(batch_row_kernel<N, F, M, W>)
               : reinterpret_cast<void*>(batch_col_kernel<N, F, M, W>);
}

// Layouts of a group's fields between the passes: 0 row-major; 1 tiles as
// wide as a work item's columns (PER = 2048/N), so an item's column loads
// are one contiguous run (bit-exact mode, N = 256 .. 1024).
template <int N, int M>
constexpr bool batch_tiled_ok()
{
    return M == kModeExact && N >= 256 && N <= 1024;
}

template <int N, int F, int M>
void* batch_fn_f(bool row, bool per_item, int radix, int layout)
{
    if (layout == 0) return batch_fn_w<N, F, M, 0>(row, per_item, radix);
    if constexpr (batch_tiled_ok<N, M>()) {
        if (layout == 1) return batch_fn_w<N, F, M, BatchShape<N>::PER>(row, per_item, radix);
    }
    return nullptr;
}

template <int N, int M>
void* batch_fn(int fields, bool row, bool per_item, int radix, int layout)
{
    return fields == 6 ? batch_fn_f<N, 6, M>(row, per_item, radix, layout)
                       : batch_fn_f<N, 4, M>(row, per_item, radix, layout);
}

// The batched kernel of mode M; per_item: the grid has a block for every
// unit (the loop-free kernels); radix: FFT points per lane, 8 or 4; layout of
// the fields between the passes (NULL if not available for this N and mode).
template <int M>
void* batch_kernel_mode(uint32_t n, int fields, bool row, bool per_item, int radix = 8,
                        int layout = 0)
{
    switch (n) {
        case 64: return batch_fn<64, M>(fields, row, per_item, radix, layout);
        case 128: return batch_fn<128, M>(fields, row, per_item, radix, layout);
        case 256: return batch_fn<256, M>(fields, row, per_item, radix, layout);
        case 512: return batch_fn<512, M>(fields, row, per_item, radix, layout);
        case 1024: return batch_fn<1024, M>(fields, row, per_item, radix, layout);
        case 2048: return batch_fn<2048, M>(fields, row, per_item, radix, layout);
        default: return nullptr;
    }
}

}  // namespace
}  // namespace oceanamd

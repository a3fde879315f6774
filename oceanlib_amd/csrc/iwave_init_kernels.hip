// Interaction-field set-up and obstruction tables on the device.
//
//   fit_normal_kernel  : the normal equations of the least-squares stencil fit
//                        (reference src/core/iwave_kernel.cpp:286-336): one lane
//                        per lower-triangle entry of A^T W A and per entry of
//                        A^T W s, each summing over the fit samples in the
//                        reference's order with its rounding (wr = w * psi_r,
//                        then acc += wr * psi_c), so the sums are the
//                        reference's bits. The basis values psi (glibc's
//                        double cosine, whose bits the sums inherit) come from
//                        the host.
//   symbol_scan_kernel : the realised symbol of the fitted stencil on the
//                        (S+1)^2 wavenumber grid (iwave_kernel.cpp:161-173 and
//                        408-432) with the device's double cosine, and the
//                        true symbol there; the host takes the extremes from
//                        these values with a certified error bound and
//                        recomputes only the candidates exactly.
//   near_row_kernel, near_col_kernel, solid_scatter_kernel :
//                        the nearest-fluid table of an obstruction mask and
//                        the list of solid cells in index order (reference
//                        src/interaction.cpp:255-292), from the closed form
//                        below.
//   mask_shift_kernel  : recenter's whole-cell shift of the mask
//                        (src/interaction.cpp:672-690).
//
// The reference's table is a multi-source BFS: every fluid cell enqueued in
// index order, each dequeued cell claiming its unclaimed 4-neighbours (+x,
// -x, +z, -z) for its own nearest fluid cell. A cell at BFS level d is
// claimed from the first-dequeued neighbour at level d - 1, so the queue
// holds each level in the lexicographic order of the cells' claim paths
// (root's index, then the directions taken), and a cell's claim path is the
// lexicographically smallest of all paths of length d from a fluid cell.
// Every fluid cell at the cell's smallest L1 distance d starts such a path
// (a monotone path from it meets no other fluid cell before its end, else d
// would be smaller), and paths compare by their root first, so:
//
//   nearest(c) = the fluid cell f minimising (|f.x - c.x| + |f.z - c.z|, index(f)).
//
// That is separable: per row, the nearest fluid column (ties to the smaller
// column, the smaller index); then per column the minimum over rows z' of
// (row distance + |z - z'|, z', column) as a prefix and a suffix minimum.
// Keys pack (distance, index) into 32 bits, distance above the 20 index bits
// (n <= 1024), so an unsigned minimum is the lexicographic one.
#include "hip_launch.h"

#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>

namespace oceanamd {
namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kIdxBits = 20;
constexpr int kFitChunk = 32;  // fit samples staged in LDS at a time

// ---- the fit's normal equations ---------------------------------------------

__global__ __launch_bounds__(256) void fit_normal_kernel(const double* __restrict__ psi,
                                                         const double* __restrict__ w,
                                                         const double* __restrict__ sym,
                                                         uint32_t samples, uint32_t mm,
                                                         double* __restrict__ ata_lower,
                                                         double* __restrict__ atb)
{
    __shared__ double tile[kFitChunk * OCEAN_HIP_IWAVE_MAX_BASIS];
    __shared__ double tw[kFitChunk], ts[kFitChunk];
    const uint32_t tri = mm * (mm + 1u) / 2u;
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = e < tri + mm;
    const bool is_atb = e >= tri;
    // entry e < tri is (r, c), c <= r, in row-major lower-triangle order
    uint32_t r = 0, c = 0;
    if (live && !is_atb) {
        r = static_cast<uint32_t>((sqrt(8.0 * double(e) + 1.0) - 1.0) * 0.5);
        while (r * (r + 1u) / 2u > e) --r;
        while ((r + 1u) * (r + 2u) / 2u <= e) ++r;
        c = e - r * (r + 1u) / 2u;
    } else if (live) {
        r = e - tri;
    }
    double acc = 0.0;
    for (uint32_t base = 0; base < samples; base += kFitChunk) {
        const uint32_t cnt = min(uint32_t(kFitChunk), samples - base);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < cnt * mm; i += blockDim.x) tile[i] = psi[size_t(base) * mm + i];
        for (uint32_t i = threadIdx.x; i < cnt; i += blockDim.x) {
            tw[i] = w[base + i];
            ts[i] = sym[base + i];
        }
        __syncthreads();
        if (live) {
            for (uint32_t s = 0; s < cnt; ++s) {
                const double wr = tw[s] * tile[s * mm + r];
                const double v = is_atb ? ts[s] : tile[s * mm + c];
                acc = acc + wr * v;
            }
        }
    }
    if (live) {
        if (is_atb)
            atb[r] = acc;
        else
            ata_lower[e] = acc;
    }
}

// ---- the realised symbol on the scan grid -----------------------------------

__device__ __forceinline__ double true_symbol(double k, double depth)
{
    if (k <= 0.0) return 0.0;
    if (depth <= 0.0) return k;
    return k * tanh(k * depth);
}

__global__ __launch_bounds__(256) void symbol_scan_kernel(const float* __restrict__ taps, int p,
                                                          double dx, double k_max, double depth,
                                                          int scan, double2* __restrict__ out)
{
    const int side = scan + 1;
    const int pt = blockIdx.x * blockDim.x + threadIdx.x;
    if (pt >= side * side) return;
    const int zi = pt / side, xi = pt % side;
    const double kx = k_max * xi / scan;
    const double kz = k_max * zi / scan;
    const double kk = sqrt(kx * kx + kz * kz);
    const int width = 2 * p + 1;
    double sum = 0.0;
    for (int j = -p; j <= p; ++j)
        for (int i = -p; i <= p; ++i)
            sum += static_cast<double>(taps[(j + p) * width + (i + p)]) * cos(kx * i * dx + kz * j * dx);
    out[pt] = make_double2(sum, true_symbol(kk, depth));
}

// ---- the nearest-fluid table ------------------------------------------------

// C bytes of a mask row at an offset that is a multiple of C, in one load.
template <int C>
__device__ __forceinline__ void load_bytes(const uint8_t* p, uint8_t (&v)[C])
{
    if constexpr (C == 16) {
        const uint4 w = *reinterpret_cast<const uint4*>(p);
        const uint32_t words[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = uint8_t(words[k / 4] >> (8 * (k % 4)));
    } else if constexpr (C == 8) {
        const uint2 w = *reinterpret_cast<const uint2*>(p);
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = uint8_t((k < 4 ? w.x : w.y) >> (8 * (k % 4)));
    } else if constexpr (C == 4) {
        const uint32_t w = *reinterpret_cast<const uint32_t*>(p);
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = uint8_t(w >> (8 * k));
    } else {
#pragma unroll
        for (int k = 0; k < C; ++k) v[k] = p[k];
    }
}

template <int C>
__device__ __forceinline__ void store_bytes(uint8_t* p, const uint8_t (&v)[C])
{
#pragma unroll
    for (int k = 0; k < C; ++k) p[k] = v[k];
}

// One wavefront per row, C = n / 64 cells per lane (lanes past n idle):
// rowbest[z n + x] = (distance << 20) | (z n + column) of the row's nearest
// fluid cell, kNone if the row has none; rowcount[z] = the row's solid cells.
// The mask is read from `src` (device memory, or host memory mapped into the
// device's address space: read once, straight from the caller's upload) and,
// when `keep` is not null, copied there.
template <int C>
__global__ __launch_bounds__(256) void near_row_kernel(const uint8_t* __restrict__ src,
                                                       uint8_t* __restrict__ keep, uint32_t n,
                                                       uint32_t* __restrict__ rowbest,
                                                       uint32_t* __restrict__ rowcount)
{
    const int lane = threadIdx.x & 63;
    const uint32_t z = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (z >= n) return;  // the whole wavefront
    const int x0 = lane * C;
    const bool active = uint32_t(x0) < n;
    uint8_t v[C];
    if (active) {
        load_bytes<C>(src + size_t(z) * n + x0, v);
        if (keep != nullptr) store_bytes<C>(keep + size_t(z) * n + x0, v);
    }
    bool fluid[C];
    int last = -1, first = INT_MAX;
    uint32_t solids = 0;
#pragma unroll
    for (int k = 0; k < C; ++k) {
        fluid[k] = active && v[k] == 0;
        if (active && v[k] != 0) ++solids;
        if (fluid[k]) {
            last = x0 + k;
            if (first == INT_MAX) first = x0 + k;
        }
    }
    // the last fluid column left of this lane's cells, the first right of them
    int incl = last;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(incl, off, 64);
        if (lane >= off) incl = max(incl, y);
    }
    int left = __shfl_up(incl, 1, 64);
    if (lane == 0) left = -1;
    int incr = first;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_down(incr, off, 64);
        if (lane + off < 64) incr = min(incr, y);
    }
    int right = __shfl_down(incr, 1, 64);
    if (lane == 63) right = INT_MAX;
    uint32_t tot = solids;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) tot += __shfl_xor(tot, off, 64);
    if (lane == 0) rowcount[z] = tot;
    if (!active) return;

    int lft[C];
#pragma unroll
    for (int k = 0; k < C; ++k) {
        if (fluid[k]) left = x0 + k;
        lft[k] = left;
    }
#pragma unroll
    for (int k = C - 1; k >= 0; --k) {
        const int x = x0 + k;
        if (fluid[k]) right = x;
        uint32_t best = kNone;
        const int l = lft[k];
        if (l >= 0 && (right == INT_MAX || x - l <= right - x))
            best = (uint32_t(x - l) << kIdxBits) | (z * n + uint32_t(l));
        else if (right != INT_MAX)
            best = (uint32_t(right - x) << kIdxBits) | (z * n + uint32_t(right));
        rowbest[size_t(z) * n + x] = best;
    }
}

// 64 columns per block, the rows in 16 segments (one wavefront each):
// near[z n + x] = the index of cell (x, z)'s nearest fluid cell, kNone when
// the mask has no fluid cell.
constexpr int kColSegs = 16;

__global__ __launch_bounds__(64 * kColSegs) void near_col_kernel(const uint32_t* __restrict__ rowbest,
                                                                 uint32_t n,
                                                                 uint32_t* __restrict__ near)
{
    __shared__ uint32_t seg_f[kColSegs][64], seg_b[kColSegs][64];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t x = blockIdx.x * 64u + lane;
    const bool active = x < n;
    const uint32_t len = n / kColSegs;
    const uint32_t z0 = w * len;
    // forward key (g + n - z') : minimum over z' <= z, then + z - n
    // backward key (g + z')    : minimum over z' >= z, then - z
    uint32_t sf = kNone, sb = kNone;
    if (active) {
        for (uint32_t z = z0; z < z0 + len; ++z) {
            const uint32_t rb = rowbest[size_t(z) * n + x];
            if (rb != kNone) {
                sf = min(sf, rb + ((n - z) << kIdxBits));
                sb = min(sb, rb + (z << kIdxBits));
            }
        }
    }
    seg_f[w][lane] = sf;
    seg_b[w][lane] = sb;
    __syncthreads();
    uint32_t run = kNone;
    for (uint32_t v = 0; v < w; ++v) run = min(run, seg_f[v][lane]);
    uint32_t runb = kNone;
    for (uint32_t v = w + 1; v < uint32_t(kColSegs); ++v) runb = min(runb, seg_b[v][lane]);
    if (!active) return;
    for (uint32_t z = z0; z < z0 + len; ++z) {
        const uint32_t rb = rowbest[size_t(z) * n + x];
        if (rb != kNone) run = min(run, rb + ((n - z) << kIdxBits));
        near[size_t(z) * n + x] = run == kNone ? kNone : run - ((n - z) << kIdxBits);
    }
    for (uint32_t zz = z0 + len; zz > z0; --zz) {
        const uint32_t z = zz - 1u;
        const uint32_t rb = rowbest[size_t(z) * n + x];
        if (rb != kNone) runb = min(runb, rb + (z << kIdxBits));
        const uint32_t cb = runb == kNone ? kNone : runb - (z << kIdxBits);
        const uint32_t best = min(near[size_t(z) * n + x], cb);
        near[size_t(z) * n + x] = best == kNone ? kNone : (best & ((1u << kIdxBits) - 1u));
    }
}

// One wavefront per row: the row's solid cells, in index order, and their
// nearest fluid cells, after the solid cells of the rows before it (summed
// from rowcount by the wavefront itself); the last row's wavefront writes the
// total.
template <int C>
__global__ __launch_bounds__(256) void solid_scatter_kernel(const uint8_t* __restrict__ mask,
                                                            const uint32_t* __restrict__ near,
                                                            uint32_t n,
                                                            const uint32_t* __restrict__ rowcount,
                                                            uint32_t* __restrict__ solid,
                                                            uint32_t* __restrict__ source,
                                                            uint32_t* __restrict__ count)
{
    const int lane = threadIdx.x & 63;
    const uint32_t z = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (z >= n) return;
    uint32_t before = 0;
    for (uint32_t r = uint32_t(lane); r < z; r += 64u) before += rowcount[r];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) before += __shfl_xor(before, off, 64);
    if (z == n - 1u && lane == 0) *count = before + rowcount[z];
    const int x0 = lane * C;
    const bool active = uint32_t(x0) < n;
    const uint8_t* row = mask + size_t(z) * n;
    bool sol[C];
    uint32_t cnt = 0;
#pragma unroll
    for (int k = 0; k < C; ++k) {
        sol[k] = active && row[x0 + k] != 0;
        cnt += sol[k] ? 1u : 0u;
    }
    uint32_t incl = cnt;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    uint32_t at = before + incl - cnt;
#pragma unroll
    for (int k = 0; k < C; ++k) {
        if (sol[k]) {
            const uint32_t i = z * n + uint32_t(x0 + k);
            solid[at] = i;
            source[at] = near[i];
            ++at;
        }
    }
}

__global__ __launch_bounds__(256) void mask_shift_kernel(const uint8_t* __restrict__ src,
                                                         uint8_t* __restrict__ dst, uint32_t n,
                                                         int32_t dx, int32_t dz)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n * n) return;
    const int32_t x = int32_t(i % n), z = int32_t(i / n);
    const int32_t sx = x + dx, sz = z + dz;
    const int32_t ni = int32_t(n);
    dst[i] = (sx >= 0 && sx < ni && sz >= 0 && sz < ni) ? src[size_t(sz) * n + sx] : uint8_t(0);
}

}  // namespace
}  // namespace oceanamd

using namespace oceanamd;

extern "C" hipError_t ocean_hip_iwave_fit_normal(const double* psi, const double* w,
                                                 const double* sym, uint32_t samples, uint32_t mm,
                                                 double* ata_lower, double* atb, hipStream_t s)
{
    if (mm == 0 || mm > OCEAN_HIP_IWAVE_MAX_BASIS || samples == 0) return hipErrorInvalidValue;
    const uint32_t lanes = mm * (mm + 1u) / 2u + mm;
    hipLaunchKernelGGL(fit_normal_kernel, dim3((lanes + 255u) / 256u), dim3(256), 0, s, psi, w, sym,
                       samples, mm, ata_lower, atb);
    return hipGetLastError();
}

extern "C" hipError_t ocean_hip_iwave_symbol_scan(const float* taps, uint32_t radius, double dx,
                                                  double k_max, double depth, uint32_t scan,
                                                  double* out, hipStream_t s)
{
    if (radius < 1 || radius > 12 || scan == 0 || scan > 4096) return hipErrorInvalidValue;
    const uint32_t pts = (scan + 1u) * (scan + 1u);
    hipLaunchKernelGGL(symbol_scan_kernel, dim3((pts + 255u) / 256u), dim3(256), 0, s, taps,
                       int(radius), dx, k_max, depth, int(scan), reinterpret_cast<double2*>(out));
    return hipGetLastError();
}

namespace {

template <int C>
hipError_t launch_rows(const uint8_t* src, uint8_t* keep, uint32_t n, uint32_t* rowbest,
                       uint32_t* rowcount, hipStream_t s)
{
    hipLaunchKernelGGL(near_row_kernel<C>, dim3((n + 3u) / 4u), dim3(256), 0, s, src, keep, n,
                       rowbest, rowcount);
    return hipGetLastError();
}

template <int C>
hipError_t launch_scatter(const uint8_t* mask, const uint32_t* near, uint32_t n,
                          const uint32_t* rowcount, uint32_t* solid, uint32_t* source,
                          uint32_t* count, hipStream_t s)
{
    hipLaunchKernelGGL(solid_scatter_kernel<C>, dim3((n + 3u) / 4u), dim3(256), 0, s, mask, near,
                       n, rowcount, solid, source, count);
    return hipGetLastError();
}

}  // namespace

extern "C" hipError_t ocean_hip_iwave_mask_tables(const uint8_t* src, uint8_t* keep, uint32_t n,
                                                  uint32_t* work, uint32_t* near,
                                                  uint32_t* solid, uint32_t* source,
                                                  uint32_t* count, hipStream_t s)
{
    if (n < 32 || n > 1024 || (n & (n - 1u)) != 0u) return hipErrorInvalidValue;
    uint32_t* rowbest = work;
    uint32_t* rowcount = work + size_t(n) * n;
    const uint8_t* mask = keep != nullptr ? keep : src;  // the device copy, for the scatter
    hipError_t e;
    switch (n / 64u) {
        case 0:
        case 1: e = launch_rows<1>(src, keep, n, rowbest, rowcount, s); break;
        case 2: e = launch_rows<2>(src, keep, n, rowbest, rowcount, s); break;
        case 4: e = launch_rows<4>(src, keep, n, rowbest, rowcount, s); break;
        case 8: e = launch_rows<8>(src, keep, n, rowbest, rowcount, s); break;
        default: e = launch_rows<16>(src, keep, n, rowbest, rowcount, s); break;
    }
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(near_col_kernel, dim3((n + 63u) / 64u), dim3(64 * kColSegs), 0, s, rowbest,
                       n, near);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    switch (n / 64u) {
        case 0:
        case 1: return launch_scatter<1>(mask, near, n, rowcount, solid, source, count, s);
        case 2: return launch_scatter<2>(mask, near, n, rowcount, solid, source, count, s);
        case 4: return launch_scatter<4>(mask, near, n, rowcount, solid, source, count, s);
        case 8: return launch_scatter<8>(mask, near, n, rowcount, solid, source, count, s);
        default: return launch_scatter<16>(mask, near, n, rowcount, solid, source, count, s);
    }
}

extern "C" hipError_t ocean_hip_iwave_mask_shift(const uint8_t* src, uint8_t* dst, uint32_t n,
                                                 int32_t dx, int32_t dz, hipStream_t s)
{
    const uint32_t cells = n * n;
    hipLaunchKernelGGL(mask_shift_kernel, dim3((cells + 255u) / 256u), dim3(256), 0, s, src, dst, n,
                       dx, dz);
    return hipGetLastError();
}
